// ctx.hpp -- state behind one C-ABI handle (the GodunovImplemBase counterpart).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "mesh.hpp"
#include "phys.hpp"

namespace mhdct
{

// HydroSettings / MHDSettings subset the hot path reads (models/HydroSettings.cpp:19-30,
// models/MHDSettings.cpp:20); riemann must be HLLD, prolongation SIMPLE_COPY.
struct Settings
{
  double gamma0 = 1.4;
  double cfl = 0.5;
  double slope_type = 1.0;
  double smallr = 1e-10;
  double smallc = 1e-10;
  double scaling = 1.0;
  int    abort_when_negative_eint = 0;
  int    time_integrator = 0; // 0 HANCOCK, 1 RK2_SSP
  int    gravity_enabled = 0;
  double gravity[3] = { 0, 0, 0 };
  int    viscosity_enabled = 0;   // [viscosity] enabled (plain form only)
  double mu = 0.0;                // [viscosity] mu, the dynamic viscosity
  int    viscosity_predictor = 0; // [viscosity] hancock_predictor_enabled
  int    prolongation = 0;        // [amr] cell_centered_prolongation_type: 0 SIMPLE_COPY, 1 EXTRAPOLATE_LINEAR_MINMOD
};

// Mesh tables resident in HBM.
struct DevMesh
{
  int              dim = 3;
  int              b = 16;
  int              nowned = 0;
  int              ntotal = 0;
  int              nfix = 0;
  int              noutside = 0;
  int              prolong = 0; // ghost cells beside a coarser block: 0 SIMPLE_COPY, 1 EXTRAPOLATE_LINEAR_MINMOD
  int              bc[6] = { 0, 0, 0, 0, 0, 0 };
  int              nbrick[3] = { 1, 1, 1 };
  int *            level = nullptr;     // [ntotal]
  int *            coord = nullptr;     // [ntotal][3] block coordinates at the block's level
  NbEntry *        nb = nullptr;        // [nowned][3^dim][2^dim]
  uint8_t *        conformal = nullptr; // [nowned][2*dim]
  EmfFix *         fix = nullptr;       // [nfix] corrected border edges, built on the device
  int64_t *        fix_src = nullptr;   // [nfix][8] their sources
  int *            out_src = nullptr;   // [noutside]
  int *            out_dirs = nullptr;  // [noutside*3]
};

// Per-block element counts of the V0-shaped scratch arrays (GodunovImplemV0.h:52-103).
struct Shapes
{
  int  dim, b;
  long cells() const { return dim == 3 ? long(b) * b * b : long(b) * b; }
  long gcells(int g) const
  {
    const long n = b + 2 * g;
    return dim == 3 ? n * n * n : n * n;
  }
  long fsize(int c) const
  {
    if (dim == 2)
      return c == 2 ? long(b) * b : long(b + 1) * b;
    return long(b + 1) * b * b;
  }
  long fper() const { return fsize(0) + fsize(1) + fsize(2); }
  long gfsize(int c) const
  {
    const long n = b + 4;
    if (dim == 2)
      return c == 2 ? n * n : (n + 1) * n;
    return (n + 1) * n * n;
  }
  long gfper() const { return gfsize(0) + gfsize(1) + gfsize(2); }
  long flux_cells() const { return dim == 3 ? long(b + 1) * b * b : long(b + 1) * b; }
  long edge_cells() const
  {
    const long n = b + 1;
    return dim == 3 ? n * n * n : n * n;
  }
  int ncomp_e() const { return dim == 3 ? 3 : 1; }
};

struct PlainScratch
{
  double * Q = nullptr;    // 11 x (b+4)^d
  double * Q2 = nullptr;   // 8 x (b+2)^d
  double * S[3] = { nullptr, nullptr, nullptr };
  double * Fl = nullptr;   // 8 x flux cells
  double * E = nullptr;    // (1|3) x (b+4)^d
  double * sF = nullptr;   // (2|3) x (b+2)^d
  double * emf = nullptr;  // (1|3) x (b+1)^d
  double * emf2 = nullptr;
  double * Bg = nullptr;   // ghosted face B
  size_t   bytes = 0;
};

struct Context
{
  Settings       st;
  int            dim = 3;
  int            b = 16;
  int            impl = 10;
  int            device = 0;
  bool           has_mesh = false;
  DevMesh        dm;
  size_t         mesh_bytes = 0;
  // capacity in bytes of each mesh table of dm (level, coord, nb, conformal, out_src,
  // out_dirs, fix, fix_src) and of the border-list count scan: each is reallocated only when
  // a new mesh outgrows it
  size_t         table_cap[9] = { 0, 0, 0, 0, 0, 0, 0, 0, 0 };
  int *          border_scan = nullptr;
  PlainScratch   ps;
  // reduction scalars
  unsigned long long * red = nullptr; // [0] inv_dt bits, [1] error flags
  std::string    err;
  // fused-form scratch (border EMFs etc.), owned by the fused implementation
  void *         fused = nullptr;
  // refinement-indicator scratch (per owned block: indicator bits, indicator, flag)
  void *         amr_buf = nullptr;
  size_t         amr_bytes = 0;
};

Gas<double> gas_of(const Settings & s);

// plain (one kernel per stage) form, plain.hip
int  plain_alloc(Context & c);
void plain_free(Context & c);
int  plain_step(Context & c, const double * U, double * U2, const double * B, double * B2, double dt,
                hipStream_t s);
int  launch_inv_dt(Context & c, const double * U, const double * B, hipStream_t s);
// the viscous fluxes and their conservative update, direction by direction (GodunovImplemV0.cpp:
// 179-192), from primitive records [block][qvars][(b+4)^d] whose velocities are variables 2-4;
// Fl holds 8 x flux cells per owned block
void launch_viscous(Context & c, const double * Qp, int qvars, double * Fl, double * U2, double * B2, double dt,
                    hipStream_t s);
int  launch_fill_outside(Context & c, double * U, double * B, hipStream_t s);
// the border-edge gather list of the uploaded mesh, built on the device (border.hip)
int  build_border_list(Context & c);
int  launch_linear_combination(Context & c, const double * U, double * U_RK, const double * B, double * B_RK,
                               double c0, double c1, hipStream_t s);

} // namespace mhdct
