// config.hpp -- the .ini settings of a run: an INI reader in place of kalypsso::core's ConfigMap
// (absent from the reference tree), and the keys and defaults the reference reads from it.
//
// Syntax (the .ini files under app/): `[section]` lines, `key = value` lines (blanks around
// the name and the value are dropped), blank lines, and comment lines whose first non-blank
// character is '#' or ';'.  Section and key names are case-insensitive; a later value of a
// key replaces an earlier one.  Typed reads follow the ConfigMap calls the reference makes:
// getReal / getInteger parse the whole value (else the default), getBool takes
// true/yes/on/1 and false/no/off/0 in any case (else the default), getStringVector splits on
// commas and drops blanks around each item.
#pragma once

#include <map>
#include <string>
#include <vector>

namespace mhdct
{

class Config
{
public:
  // 0, or the 1-based number of the first line that is none of the forms above
  int parse(const std::string & text);

  const std::string * find(const std::string & section, const std::string & key) const;
  std::string         get_string(const std::string & section, const std::string & key, const std::string & dflt) const;
  double              get_real(const std::string & section, const std::string & key, double dflt) const;
  long                get_integer(const std::string & section, const std::string & key, long dflt) const;
  bool                get_bool(const std::string & section, const std::string & key, bool dflt) const;
  std::vector<std::string> get_string_vector(const std::string & section, const std::string & key,
                                             const std::vector<std::string> & dflt) const;

  mutable std::string err;

private:
  std::map<std::string, std::map<std::string, std::string>> data_;
};

std::string lower(std::string s);
std::string trim(const std::string & s);

} // namespace mhdct
