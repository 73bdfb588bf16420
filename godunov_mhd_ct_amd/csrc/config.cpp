// config.cpp -- the INI reader of config.hpp.
#include "config.hpp"

#include <cctype>
#include <cerrno>
#include <cstdlib>

namespace mhdct
{

std::string
lower(std::string s)
{
  for (char & c : s)
    c = char(std::tolower(static_cast<unsigned char>(c)));
  return s;
}

std::string
trim(const std::string & s)
{
  size_t a = 0, b = s.size();
  while (a < b && std::isspace(static_cast<unsigned char>(s[a])))
    ++a;
  while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1])))
    --b;
  return s.substr(a, b - a);
}

int
Config::parse(const std::string & text)
{
  data_.clear();
  std::string section;
  size_t      pos = 0;
  int         lineno = 0;
  while (pos <= text.size())
  {
    size_t end = text.find('\n', pos);
    if (end == std::string::npos)
      end = text.size();
    std::string line = trim(text.substr(pos, end - pos));
    pos = end + 1;
    ++lineno;
    if (line.empty() || line[0] == '#' || line[0] == ';')
      continue;
    if (line[0] == '[')
    {
      const size_t close = line.find(']');
      if (close == std::string::npos || trim(line.substr(close + 1)).size() > 0)
      {
        err = "line " + std::to_string(lineno) + ": a section header must be [name]";
        return lineno;
      }
      section = lower(trim(line.substr(1, close - 1)));
      continue;
    }
    const size_t eq = line.find('=');
    if (eq == std::string::npos || trim(line.substr(0, eq)).empty())
    {
      err = "line " + std::to_string(lineno) + ": expected key = value";
      return lineno;
    }
    data_[section][lower(trim(line.substr(0, eq)))] = trim(line.substr(eq + 1));
  }
  return 0;
}

const std::string *
Config::find(const std::string & section, const std::string & key) const
{
  auto s = data_.find(lower(section));
  if (s == data_.end())
    return nullptr;
  auto k = s->second.find(lower(key));
  return k == s->second.end() ? nullptr : &k->second;
}

std::string
Config::get_string(const std::string & section, const std::string & key, const std::string & dflt) const
{
  const std::string * v = find(section, key);
  return v ? *v : dflt;
}

double
Config::get_real(const std::string & section, const std::string & key, double dflt) const
{
  const std::string * v = find(section, key);
  if (!v || v->empty())
    return dflt;
  char * endp = nullptr;
  errno = 0;
  const double x = std::strtod(v->c_str(), &endp);
  return (endp && *endp == '\0' && errno == 0) ? x : dflt;
}

long
Config::get_integer(const std::string & section, const std::string & key, long dflt) const
{
  const std::string * v = find(section, key);
  if (!v || v->empty())
    return dflt;
  char * endp = nullptr;
  errno = 0;
  const long x = std::strtol(v->c_str(), &endp, 0);
  return (endp && *endp == '\0' && errno == 0) ? x : dflt;
}

bool
Config::get_bool(const std::string & section, const std::string & key, bool dflt) const
{
  const std::string * v = find(section, key);
  if (!v)
    return dflt;
  const std::string s = lower(*v);
  if (s == "true" || s == "yes" || s == "on" || s == "1")
    return true;
  if (s == "false" || s == "no" || s == "off" || s == "0")
    return false;
  return dflt;
}

std::vector<std::string>
Config::get_string_vector(const std::string & section, const std::string & key,
                          const std::vector<std::string> & dflt) const
{
  const std::string * v = find(section, key);
  if (!v)
    return dflt;
  std::vector<std::string> out;
  size_t                   pos = 0;
  while (pos <= v->size())
  {
    size_t end = v->find(',', pos);
    if (end == std::string::npos)
      end = v->size();
    const std::string item = trim(v->substr(pos, end - pos));
    if (!item.empty())
      out.push_back(item);
    pos = end + 1;
  }
  return out;
}

} // namespace mhdct
