// capi.hip -- the extern "C" boundary (include/mhdct.h).
#include <hip/hip_runtime.h>

#include <cstring>
#include <fstream>
#include <new>
#include <sstream>
#include <string>

#include "../../include/mhdct.h"
#include "ctx.hpp"
#include "fused_host.hpp"
#include "handles.hpp"

using mhdct::Context;
using mhdct::Mesh;



namespace
{

// the message of the last failed mhdct_remap_create on this thread
std::string &
remap_create_error()
{
  thread_local std::string e;
  return e;
}

// copy a host table into its device buffer, reallocated (with a margin of 1/2) only when the
// new table does not fit: an AMR cycle then mostly reuses the previous mesh's buffers
template <typename T>
int
upload(Context & c, T *& dst, size_t & cap, const std::vector<T> & src)
{
  const size_t bytes = (src.empty() ? 1 : src.size()) * sizeof(T);
  hipError_t   e = hipSuccess;
  if (!dst || cap < bytes)
  {
    if (dst)
      (void)hipFree(dst);
    dst = nullptr;
    cap = 0;
    e = hipMalloc(&dst, bytes + bytes / 2);
    if (e != hipSuccess)
    {
      dst = nullptr;
      c.err = std::string("hipMalloc: ") + hipGetErrorString(e);
      return -1;
    }
    cap = bytes + bytes / 2;
  }
  if (!src.empty())
  {
    e = hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess)
    {
      c.err = std::string("hipMemcpy: ") + hipGetErrorString(e);
      return -1;
    }
  }
  c.mesh_bytes += cap;
  return 0;
}

// the mesh's device tables, the border-edge gather list included, and the scan buffer it is
// built with
void
free_mesh(Context & c)
{
  void * ts[] = { c.dm.level, c.dm.coord, c.dm.nb, c.dm.conformal, c.dm.out_src, c.dm.out_dirs, c.dm.fix,
                  c.dm.fix_src, c.border_scan };
  for (void * p : ts)
    if (p)
      (void)hipFree(p);
  c.dm.level = nullptr;
  c.dm.coord = nullptr;
  c.dm.nb = nullptr;
  c.dm.conformal = nullptr;
  c.dm.out_src = nullptr;
  c.dm.out_dirs = nullptr;
  c.dm.fix = nullptr;
  c.dm.fix_src = nullptr;
  c.dm.nfix = 0;
  c.border_scan = nullptr;
  c.mesh_bytes = 0;
  for (auto & cp : c.table_cap)
    cp = 0;
}

} // namespace

extern "C"
{

  int
  mhdct_mesh_create(int dim, int block_size, const int nbrick[3], const int periodic[3], int level,
                    mhdct_mesh ** out)
  {
    if (!out)
      return -1;
    *out = nullptr;
    if (dim != 2 && dim != 3)
      return -1;
    if (block_size < 4 || (block_size % 2) || level < 0 || level > 12)
      return -1;
    auto * h = new (std::nothrow) mhdct_mesh();
    if (!h)
      return -1;
    h->m.dim = dim;
    h->m.b = block_size;
    for (int d = 0; d < 3; ++d)
    {
      h->m.nbrick[d] = d < dim ? (nbrick ? nbrick[d] : 1) : 1;
      h->m.periodic[d] = d < dim ? (periodic ? (periodic[d] != 0) : 1) : 1;
      if (h->m.nbrick[d] < 1)
      {
        delete h;
        return -1;
      }
    }
    h->m.uniform(level);
    *out = h;
    return 0;
  }

  int
  mhdct_mesh_set_bc(mhdct_mesh * mesh, const int bc[6])
  {
    if (!mesh || !bc)
      return -1;
    for (int f = 0; f < 6; ++f)
    {
      if (bc[f] != 0 && bc[f] != 1 && bc[f] != 2 && bc[f] != 3)
      {
        mesh->m.err = "border condition must be one of NONE(0) PERIODIC(1) ZERO_GRADIENT(2) WALL(3)";
        return -1;
      }
      mesh->m.bc[f] = bc[f];
    }
    return 0;
  }

  int
  mhdct_mesh_adapt(mhdct_mesh * mesh, const int8_t * flags, int nleaves)
  {
    if (!mesh || (nleaves > 0 && !flags))
      return -1;
    std::vector<int8_t> f(flags, flags + nleaves);
    return mesh->m.adapt(f) ? 0 : -1;
  }

  int
  mhdct_mesh_set_leaves(mhdct_mesh * mesh, int n, const int * levels, const int * coords)
  {
    if (!mesh || n < 0 || (n > 0 && (!levels || !coords)))
      return -1;
    std::vector<mhdct::Leaf> given(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i)
      given[i] = { levels[i], coords[3 * i], coords[3 * i + 1], coords[3 * i + 2] };
    return mesh->m.set_leaves(given) ? 0 : -1;
  }

  int
  mhdct_mesh_balance(mhdct_mesh * mesh)
  {
    if (!mesh)
      return -1;
    mesh->m.balance();
    return 0;
  }

  int
  mhdct_mesh_finalize(mhdct_mesh * mesh)
  {
    if (!mesh)
      return -1;
    return mesh->m.finalize() ? 0 : -2;
  }

  int
  mhdct_mesh_num_leaves(const mhdct_mesh * mesh)
  {
    return mesh ? int(mesh->m.leaves.size()) : -1;
  }

  int
  mhdct_mesh_num_outside(const mhdct_mesh * mesh)
  {
    return mesh ? int(mesh->m.outside.size()) : -1;
  }

  int
  mhdct_mesh_leaves(const mhdct_mesh * mesh, int * levels, int * coords)
  {
    if (!mesh)
      return -1;
    int n = 0;
    for (auto & L : mesh->m.leaves)
    {
      if (levels)
        levels[n] = L.level;
      if (coords)
      {
        coords[3 * n] = L.x;
        coords[3 * n + 1] = L.y;
        coords[3 * n + 2] = L.z;
      }
      ++n;
    }
    for (auto & L : mesh->m.outside)
    {
      if (levels)
        levels[n] = L.level;
      if (coords)
      {
        coords[3 * n] = L.x;
        coords[3 * n + 1] = L.y;
        coords[3 * n + 2] = L.z;
      }
      ++n;
    }
    return 0;
  }

  int
  mhdct_mesh_outside(const mhdct_mesh * mesh, int * src, int * dirs)
  {
    if (!mesh)
      return -1;
    for (size_t i = 0; i < mesh->m.outside.size(); ++i)
    {
      if (src)
        src[i] = mesh->m.outside_src[i];
      if (dirs)
        for (int d = 0; d < 3; ++d)
          dirs[3 * i + d] = mesh->m.outside_dirs[3 * i + d];
    }
    return 0;
  }

  const char *
  mhdct_mesh_last_error(const mhdct_mesh * mesh)
  {
    return mesh ? mesh->m.err.c_str() : "null mesh handle";
  }

  void
  mhdct_mesh_destroy(mhdct_mesh * mesh)
  {
    delete mesh;
  }

  int
  mhdct_create(const mhdct_settings * s, int dim, const int block_size[3], int impl_version, mhdct_ctx ** out)
  {
    if (!out)
      return -1;
    *out = nullptr;
    if (!s || !block_size || (dim != 2 && dim != 3))
      return -1;
    if (impl_version != MHDCT_IMPL_PLAIN && impl_version != MHDCT_IMPL_FUSED && impl_version != MHDCT_IMPL_FUSED_FP32)
      return -4;
    if (s->riemann != 5 || (s->prolongation != 0 && s->prolongation != 1))
      return -5;
    if (s->time_integrator != MHDCT_TIME_HANCOCK && s->time_integrator != MHDCT_TIME_RK2_SSP)
      return -5;
    // viscous fluxes run in the FP64 forms (plain and fused), not in the FP32-bulk form
    if (s->viscosity_enabled && (impl_version == MHDCT_IMPL_FUSED_FP32 || !(s->viscosity_mu >= 0.0)))
      return -5;
    for (int d = 1; d < dim; ++d)
      if (block_size[d] != block_size[0])
        return -6; // dx == dy == dz (ComputeDtMHDFunctor.cpp:98)
    auto * h = new (std::nothrow) mhdct_ctx();
    if (!h)
      return -1;
    Context & c = h->c;
    c.dim = dim;
    c.b = block_size[0];
    c.impl = impl_version;
    c.st.gamma0 = s->gamma0;
    c.st.cfl = s->cfl;
    c.st.slope_type = s->slope_type;
    c.st.smallr = s->smallr;
    c.st.smallc = s->smallc;
    c.st.scaling = s->scaling_factor;
    c.st.abort_when_negative_eint = s->abort_when_negative_eint;
    c.st.time_integrator = s->time_integrator;
    c.st.gravity_enabled = s->gravity_enabled != 0;
    for (int d = 0; d < 3; ++d)
      c.st.gravity[d] = d < dim ? s->gravity[d] : 0.0;
    c.st.viscosity_enabled = s->viscosity_enabled != 0;
    c.st.mu = s->viscosity_mu;
    c.st.viscosity_predictor = s->viscosity_hancock_predictor != 0;
    c.st.prolongation = s->prolongation;
    if (hipGetDevice(&c.device) != hipSuccess)
    {
      delete h;
      return -7;
    }
    if (hipMalloc(&c.red, 2 * sizeof(unsigned long long)) != hipSuccess)
    {
      delete h;
      return -7;
    }
    (void)hipMemset(c.red, 0, 2 * sizeof(unsigned long long));
    *out = h;
    return 0;
  }

  int
  mhdct_set_mesh(mhdct_ctx * h, const mhdct_mesh * mesh)
  {
    if (!h || !mesh)
      return -1;
    Context & c = h->c;
    if (mesh->m.dim != c.dim || mesh->m.b != c.b)
    {
      c.err = "mesh dimension / block size differ from the context's";
      return -1;
    }
    if (mesh->m.nb.empty() && !mesh->m.leaves.empty())
    {
      c.err = "mesh not finalized (call mhdct_mesh_finalize)";
      return -1;
    }
    const mhdct::Mesh & M = mesh->m;
    // the device tables are reused when large enough: work still in flight on any stream may
    // read the previous mesh's, so let it finish first
    const hipError_t se = hipDeviceSynchronize();
    if (se != hipSuccess)
    {
      c.err = std::string("set_mesh: ") + hipGetErrorString(se);
      return -1;
    }
    // the context holds a mesh again only once every table and the scratch are in place: a
    // failure below leaves it without one, so the step entry points refuse to run
    c.has_mesh = false;
    c.mesh_bytes = 0;
    c.dm.dim = c.dim;
    c.dm.b = c.b;
    c.dm.nowned = int(M.leaves.size());
    c.dm.ntotal = M.ntotal();
    c.dm.noutside = int(M.outside.size());
    c.dm.prolong = c.st.prolongation;
    c.dm.nfix = 0;
    for (int f = 0; f < 6; ++f)
      c.dm.bc[f] = M.bc[f];
    for (int d = 0; d < 3; ++d)
      c.dm.nbrick[d] = M.nbrick[d];
    std::vector<int> lv, co;
    lv.reserve(size_t(c.dm.ntotal));
    co.reserve(size_t(c.dm.ntotal) * 3);
    for (const auto * list : { &M.leaves, &M.outside })
      for (auto & L : *list)
      {
        lv.push_back(L.level);
        co.push_back(L.x);
        co.push_back(L.y);
        co.push_back(L.z);
      }
    if (upload(c, c.dm.level, c.table_cap[0], lv) || upload(c, c.dm.coord, c.table_cap[1], co) ||
        upload(c, c.dm.nb, c.table_cap[2], M.nb) || upload(c, c.dm.conformal, c.table_cap[3], M.conformal) ||
        upload(c, c.dm.out_src, c.table_cap[4], M.outside_src) ||
        upload(c, c.dm.out_dirs, c.table_cap[5], M.outside_dirs) || mhdct::build_border_list(c))
      return -1;
    int rc = 0;
    if (c.impl == MHDCT_IMPL_PLAIN)
      rc = mhdct::plain_alloc(c);
    else
      rc = mhdct::fused_alloc(c, M);
    if (rc == 0)
      c.has_mesh = true;
    return rc;
  }

  uint64_t
  mhdct_mem_bytes(const mhdct_ctx * h)
  {
    if (!h)
      return 0;
    return uint64_t(h->c.ps.bytes + h->c.mesh_bytes);
  }

  int
  mhdct_do_time_step(mhdct_ctx * h, const double * U, double * U2, const double * B, double * B2, double dt,
                     void * stream)
  {
    if (!h)
      return -1;
    Context & c = h->c;
    if (!c.has_mesh)
    {
      c.err = "no mesh set";
      return -1;
    }
    if (c.dm.nowned > 0 && (!U || !U2 || !B || !B2))
    {
      c.err = "null state pointer";
      return -1;
    }
    if (!(dt > 0.0) || dt != dt)
    {
      c.err = "dt must be positive and finite";
      return -1;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (c.impl == MHDCT_IMPL_PLAIN)
      return mhdct::plain_step(c, U, U2, B, B2, dt, s);
    return mhdct::fused_step(c, U, U2, B, B2, dt, s);
  }

  int
  mhdct_compute_dt(mhdct_ctx * h, const double * U, const double * B, double * dt_out, void * stream)
  {
    if (!h || !dt_out)
      return -1;
    Context & c = h->c;
    if (!c.has_mesh)
    {
      c.err = "no mesh set";
      return -1;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    unsigned long long host[2] = { 0, 0 };
    if (hipMemsetAsync(c.red, 0, sizeof(unsigned long long), s) != hipSuccess)
      return -1;
    if (mhdct::launch_inv_dt(c, U, B, s))
      return -1;
    if (hipMemcpyAsync(host, c.red, sizeof(host), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
    {
      c.err = "dt readback failed";
      return -1;
    }
    double inv;
    std::memcpy(&inv, &host[0], sizeof(inv));
    if (host[1] != 0)
      (void)hipMemset(c.red + 1, 0, sizeof(unsigned long long));
    *dt_out = c.st.cfl / inv;
    if (host[1] & 2ull)
    {
      c.err = "NaN state met in the dt reduction";
      return -3;
    }
    if (host[1] & 1ull)
    {
      c.err = "negative internal energy detected (abort_when_negative_eint)";
      return -2;
    }
    return 0;
  }

  int
  mhdct_fill_outside(mhdct_ctx * h, double * U, double * B, void * stream)
  {
    if (!h)
      return -1;
    if (!h->c.has_mesh)
    {
      h->c.err = "no mesh set";
      return -1;
    }
    return mhdct::launch_fill_outside(h->c, U, B, static_cast<hipStream_t>(stream));
  }

  int
  mhdct_linear_combination(mhdct_ctx * h, const double * U, double * U_RK, const double * B, double * B_RK,
                           double c0, double c1, void * stream)
  {
    if (!h)
      return -1;
    Context & c = h->c;
    if (!c.has_mesh)
    {
      c.err = "no mesh set";
      return -1;
    }
    if (c.dm.ntotal > 0 && (!U || !U_RK || !B || !B_RK))
    {
      c.err = "null state pointer";
      return -1;
    }
    return mhdct::launch_linear_combination(c, U, U_RK, B, B_RK, c0, c1, static_cast<hipStream_t>(stream));
  }

  int
  mhdct_refine_flags(mhdct_ctx * h, const double * U, const double * B, const mhdct_amr_params * p, int8_t * flags,
                     double * indicator, void * stream)
  {
    if (!h || !p || !flags)
      return -1;
    Context & c = h->c;
    mhdct::AmrParams ap;
    if (mhdct::check_amr_params(c, U, B, p, ap))
      return -1;
    return mhdct::launch_refine_flags(c, U, B, ap, flags, indicator, static_cast<hipStream_t>(stream));
  }

  int
  mhdct_mesh_clone(const mhdct_mesh * mesh, mhdct_mesh ** out)
  {
    if (!mesh || !out)
      return -1;
    *out = new (std::nothrow) mhdct_mesh(*mesh);
    return *out ? 0 : -1;
  }

  int
  mhdct_remap_create(const mhdct_mesh * old_mesh, const mhdct_mesh * new_mesh, mhdct_remap ** out)
  {
    if (!out)
      return -1;
    *out = nullptr;
    if (!old_mesh || !new_mesh)
    {
      remap_create_error() = "null mesh";
      return -1;
    }
    auto * r = new (std::nothrow) mhdct_remap();
    if (!r)
      return -1;
    r->dim = new_mesh->m.dim;
    r->b = new_mesh->m.b;
    if (!mhdct::remap_plan(old_mesh->m, new_mesh->m, r->plan, r->lists, r->counts, remap_create_error()))
    {
      delete r;
      return -1;
    }
    *out = r;
    return 0;
  }

  int
  mhdct_remap_counts(const mhdct_remap * r, int counts[3])
  {
    if (!r || !counts)
      return -1;
    for (int i = 0; i < 3; ++i)
      counts[i] = r->counts[i];
    return 0;
  }

  int
  mhdct_remap_apply(mhdct_remap * r, const double * U_old, const double * B_old, double * U_new, double * B_new,
                    void * stream)
  {
    if (!r)
      return -1;
    const int nnew = int(r->plan.size());
    if (nnew == 0)
      return 0;
    if (!U_old || !B_old || !U_new || !B_new)
    {
      r->err = "null state pointer";
      return -1;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess)
    {
      r->err = "no device";
      return -1;
    }
    if (r->dev && r->device != dev)
    {
      r->err = "the remap was first applied on another device";
      return -1;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (!r->dev)
    {
      const size_t bytes = r->plan.size() * sizeof(mhdct::RemapEntry);
      const size_t lbytes = (r->lists.ints.empty() ? 1 : r->lists.ints.size()) * sizeof(int);
      if (hipMalloc(&r->dev, bytes) != hipSuccess || hipMalloc(&r->dev_lists, lbytes) != hipSuccess)
      {
        r->err = "hipMalloc of the remap plan failed";
        return -1;
      }
      r->device = dev;
      // synchronous: the host plan is pageable
      if (hipMemcpy(r->dev, r->plan.data(), bytes, hipMemcpyHostToDevice) != hipSuccess ||
          (!r->lists.ints.empty() &&
           hipMemcpy(r->dev_lists, r->lists.ints.data(), r->lists.ints.size() * sizeof(int), hipMemcpyHostToDevice) !=
             hipSuccess))
      {
        r->err = "copy of the remap plan failed";
        return -1;
      }
    }
    return mhdct::launch_remap(r->dim, r->b, r->dev, r->lists, r->dev_lists, U_old, B_old, U_new, B_new, s, r->err);
  }

  const char *
  mhdct_remap_last_error(const mhdct_remap * r)
  {
    return r ? r->err.c_str() : remap_create_error().c_str();
  }

  void
  mhdct_remap_destroy(mhdct_remap * r)
  {
    if (!r)
      return;
    if (r->dev)
      (void)hipFree(r->dev);
    if (r->dev_lists)
      (void)hipFree(r->dev_lists);
    delete r;
  }

  const char *
  mhdct_last_error(const mhdct_ctx * h)
  {
    return h ? h->c.err.c_str() : "null context";
  }

  void
  mhdct_destroy(mhdct_ctx * h)
  {
    if (!h)
      return;
    mhdct::plain_free(h->c);
    mhdct::fused_free(h->c);
    if (h->c.amr_buf)
      (void)hipFree(h->c.amr_buf);
    free_mesh(h->c);
    if (h->c.red)
      (void)hipFree(h->c.red);
    delete h;
  }

  // ---- .ini settings ---------------------------------------------------------------

  int
  mhdct_config_parse(const char * text, mhdct_config ** out)
  {
    if (!out)
      return -1;
    *out = nullptr;
    if (!text)
      return -1;
    auto * h = new (std::nothrow) mhdct_config();
    if (!h)
      return -1;
    const int bad = h->c.parse(text);
    if (bad != 0)
    {
      delete h;
      return bad;
    }
    *out = h;
    return 0;
  }

  int
  mhdct_config_load(const char * path, mhdct_config ** out)
  {
    if (!out)
      return -1;
    *out = nullptr;
    if (!path)
      return -1;
    std::ifstream f(path, std::ios::binary);
    if (!f)
      return -2;
    std::ostringstream ss;
    ss << f.rdbuf();
    if (f.bad())
      return -2;
    return mhdct_config_parse(ss.str().c_str(), out);
  }

  const char *
  mhdct_config_get(const mhdct_config * cfg, const char * section, const char * key)
  {
    if (!cfg || !section || !key)
      return nullptr;
    const std::string * v = cfg->c.find(section, key);
    return v ? v->c_str() : nullptr;
  }

  int
  mhdct_config_settings(mhdct_config * cfg, mhdct_settings * out)
  {
    if (!cfg || !out)
      return -1;
    const mhdct::Config & c = cfg->c;
    mhdct_settings        s{};
    s.gamma0 = c.get_real("hydro", "gamma0", 1.4);
    s.cfl = c.get_real("hydro", "cfl", 0.5);
    s.slope_type = c.get_real("hydro", "slope_type", 1.0);
    s.smallr = c.get_real("hydro", "smallr", 1e-10);
    s.smallc = c.get_real("hydro", "smallc", 1e-10);
    s.scaling_factor = c.get_real("mesh", "scaling_factor", 1.0);
    // RiemannSolverType (models/riemann_solver_types.h:26-33), names matched as written
    static const char * riemann_names[] = { "UNKNOWN", "APPROX", "LLF", "HLL", "HLLC", "HLLD" };
    const std::string   rn = c.get_string("hydro", "riemann", "UNKNOWN");
    s.riemann = 0;
    for (int i = 0; i < 6; ++i)
      if (rn == riemann_names[i])
        s.riemann = i;
    const std::string pro = c.get_string("amr", "cell_centered_prolongation_type", "SIMPLE_COPY");
    if (pro == "SIMPLE_COPY")
      s.prolongation = 0;
    else if (pro == "EXTRAPOLATE_LINEAR_MINMOD")
      s.prolongation = 1;
    else
    {
      cfg->c.err = "[amr] cell_centered_prolongation_type " + pro + " is unknown";
      return -1;
    }
    s.abort_when_negative_eint = c.get_bool("hydro", "abort_when_negative_eint", false) ? 1 : 0;
    const std::string * ti = c.find("run", "time_integrator");
    if (!ti)
      ti = c.find("amr", "time_integrator");
    const std::string tname = ti ? *ti : "HANCOCK";
    if (tname == "HANCOCK")
      s.time_integrator = MHDCT_TIME_HANCOCK;
    else if (tname == "RK2_SSP")
      s.time_integrator = MHDCT_TIME_RK2_SSP;
    else
    {
      cfg->c.err = "time_integrator " + tname + " is not supported (HANCOCK or RK2_SSP)";
      return -1;
    }
    const long dim = c.get_integer("run", "dimension", 3);
    s.gravity_enabled = c.get_bool("gravity", "enabled", false) ? 1 : 0;
    static const char * gnames[3] = { "uniform_gx", "uniform_gy", "uniform_gz" };
    for (int d = 0; d < 3; ++d)
      s.gravity[d] = d < dim ? c.get_real("gravity", gnames[d], 0.0) : 0.0;
    s.viscosity_enabled = c.get_bool("viscosity", "enabled", false) ? 1 : 0;
    s.viscosity_mu = c.get_real("viscosity", "mu", 0.0);
    s.viscosity_hancock_predictor = c.get_bool("viscosity", "hancock_predictor_enabled", false) ? 1 : 0;
    *out = s;
    return 0;
  }

  int
  mhdct_config_amr_params(mhdct_config * cfg, mhdct_amr_params * out)
  {
    if (!cfg || !out)
      return -1;
    const mhdct::Config & c = cfg->c;
    if (!c.find("amr", "level_min") || !c.find("amr", "level_max"))
    {
      cfg->c.err = "[amr] level_min and level_max are required";
      return -1;
    }
    mhdct_amr_params p{};
    p.level_min = int(c.get_integer("amr", "level_min", -1));
    p.level_max = int(c.get_integer("amr", "level_max", -1));
    if (p.level_min < 0 || p.level_max < p.level_min)
    {
      cfg->c.err = "[amr] needs 0 <= level_min <= level_max";
      return -1;
    }
    p.epsilon_refine = c.get_real("amr", "epsilon_refine", 0.01);
    p.epsilon_coarsen = c.get_real("amr", "epsilon_coarsen", 0.00125);
    p.epsilon_lohner = c.get_real("amr", "epsilon_lohner", 0.02);
    const std::string crit = c.get_string("amr", "refine_criterion", "LOHNER_UNSPLIT");
    if (crit == "LOHNER_UNSPLIT")
      p.criterion = MHDCT_REFINE_LOHNER_UNSPLIT;
    else if (crit == "SIMPLE_GRADIENT")
      p.criterion = MHDCT_REFINE_SIMPLE_GRADIENT;
    else
    {
      cfg->c.err = "[amr] refine_criterion " + crit + " is unknown (LOHNER_UNSPLIT or SIMPLE_GRADIENT)";
      return -1;
    }
    const auto names = c.get_string_vector("amr", "variable_used_to_compute_refine_criterion", { "rho" });
    if (names.empty() || names.size() > 4)
    {
      cfg->c.err = "[amr] variable_used_to_compute_refine_criterion: one to four variables";
      return -1;
    }
    // SolverGodunovMHD.cpp:1117-1167: a DERIVED_QUANTITY name (utils/ComputeDerivedQuantities.h:44-53,
    // matched in any case) first, else a cell variable of U by MHD::m_id2names_all
    // (models/MHD.cpp:20-32), else rho
    static const char * derived[] = { "thermal_pressure", "speed_of_sound", "specific_ekin", "magnetic_pressure",
                                      "local_mach_number", "vx", "vy", "vz" };
    static const char * uvars[] = { "rho", "e_tot", "rho_vx", "rho_vy", "rho_vz" };
    p.nvars = int(names.size());
    for (int n = 0; n < p.nvars; ++n)
    {
      int code = 0;
      bool found = false;
      for (int d = 0; d < 8; ++d)
        if (mhdct::lower(names[n]) == derived[d])
        {
          code = MHDCT_RVAR_THERMAL_PRESSURE + d;
          found = true;
        }
      for (int v = 0; v < 5 && !found; ++v)
        if (names[n] == uvars[v])
          code = v;
      p.vars[n] = code;
    }
    p.ivar = p.vars[0] < 5 ? p.vars[0] : 0;
    *out = p;
    return 0;
  }

  int
  mhdct_config_mesh(mhdct_config * cfg, mhdct_mesh ** out)
  {
    if (!out)
      return -1;
    *out = nullptr;
    if (!cfg)
      return -1;
    const mhdct::Config & c = cfg->c;
    const long            dim = c.get_integer("run", "dimension", 0);
    if (dim != 2 && dim != 3)
    {
      cfg->c.err = "[run] dimension must be 2 or 3";
      return -1;
    }
    static const char * bnames[3] = { "bx", "by", "bz" };
    long                b = -1;
    for (int d = 0; d < dim; ++d)
    {
      const long bd = c.get_integer("amr", bnames[d], -1);
      if (bd < 0)
      {
        cfg->c.err = std::string("[amr] ") + bnames[d] + " is required";
        return -1;
      }
      if (d > 0 && bd != b)
      {
        cfg->c.err = "[amr] bx, by (and bz) must be equal: blocks are cubes here";
        return -1;
      }
      b = bd;
    }
    if (b < 4 || (b % 2))
    {
      cfg->c.err = "[amr] block size must be even and at least 4";
      return -1;
    }
    if (!c.find("amr", "level_min"))
    {
      cfg->c.err = "[amr] level_min is required";
      return -1;
    }
    const long level = c.get_integer("amr", "level_min", -1);
    static const char * ax = "xyz";
    int                 nbrick[3] = { 1, 1, 1 }, periodic[3] = { 1, 1, 1 }, bc[6] = { 0, 0, 0, 0, 0, 0 };
    for (int d = 0; d < dim; ++d)
    {
      nbrick[d] = int(c.get_integer("p4est_connectivity", std::string("nbrick_") + ax[d], 1));
      periodic[d] = c.get_integer("p4est_connectivity", std::string("periodic_") + ax[d], 1) != 0 ? 1 : 0;
    }
    // BC_MHD (border_conditions/FillOutside.h:55-61)
    static const char * bcn[4] = { "NONE", "PERIODIC", "ZERO_GRADIENT", "WALL" };
    for (int f = 0; f < 2 * dim; ++f)
    {
      const std::string key = std::string("boundary_type_") + ax[f / 2] + (f % 2 ? "max" : "min");
      const std::string v = c.get_string("mesh", key, "ZERO_GRADIENT");
      int               k = -1;
      for (int i = 0; i < 4; ++i)
        if (v == bcn[i])
          k = i;
      if (k < 0)
      {
        cfg->c.err = "[mesh] " + key + " " + v + " is not built (NONE, PERIODIC, ZERO_GRADIENT or WALL)";
        return -1;
      }
      bc[f] = k;
    }
    mhdct_mesh * m = nullptr;
    const int    rc = mhdct_mesh_create(int(dim), int(b), nbrick, periodic, int(level), &m);
    if (rc != 0)
    {
      cfg->c.err = "the mesh cannot be created (nbrick >= 1, 0 <= level_min <= 12)";
      return rc;
    }
    mhdct_mesh_set_bc(m, bc);
    *out = m;
    return 0;
  }

  const char *
  mhdct_config_last_error(const mhdct_config * cfg)
  {
    return cfg ? cfg->c.err.c_str() : "null config";
  }

  void
  mhdct_config_destroy(mhdct_config * cfg)
  {
    delete cfg;
  }
}
