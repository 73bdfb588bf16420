// plain.hip -- the plain form: one gfx950 kernel per stage of GodunovImplemV0::do_time_step
// (scheme/GodunovImplemV0.cpp:101-207), with the reference's ghosted scratch arrays in HBM.
// It is the drop-in counterpart of GodunovImplemV0 and the baseline the fused form is
// measured against.  Every kernel is one thread per array element (the RangePolicy shape),
// 256-thread workgroups, FP64 throughout.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>

#include "common.hpp"
#include "ctx.hpp"

namespace mhdct
{

namespace
{

constexpr int kThreads = 256;

inline unsigned
grid_of(long n)
{
  return unsigned((n + kThreads - 1) / kThreads);
}

// ---------------------------------------------------------------------------
// K1: face B with two ghost layers (FillBlockGhostFacesFunctor, GodunovImplemV0.cpp:224-235)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads)
k_fill_bface_ghosted(DevMesh m, const double * __restrict__ B, double * __restrict__ Bg)
{
  const Idx  I{ m.dim, m.b };
  const long per = I.gfper();
  const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= long(m.nowned) * per)
    return;
  const int  o = int(t / per);
  long       r = t - long(o) * per;
  int        c = 0;
  while (c < 2 && r >= I.gfsize(c))
  {
    r -= I.gfsize(c);
    ++c;
  }
  const int b = m.b, dim = m.dim;
  const int n = b + 4;
  const int nx = c == 0 ? n + 1 : n;
  const int ny = c == 1 ? n + 1 : n;
  int       f[3] = { int(r % nx) - 2, int((r / nx) % ny) - 2, dim == 3 ? int(r / (long(nx) * ny)) - 2 : 0 };
  Bg[long(o) * per + I.gface(c, f[0], f[1], f[2])] = face_value_ghosted(m, I, B, o, c, f);
}

// ---------------------------------------------------------------------------
// K3: conservative -> primitive with two ghost layers (ConvertToPrimitivesVariablesFunctor,
// fill_ghosts .cpp:639-719; prolongation SIMPLE_COPY)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads)
k_convert_to_primitives(DevMesh m, Gas<double> g, int abort_eint, const double * __restrict__ U,
                        const double * __restrict__ B, const double * __restrict__ Bg,
                        double * __restrict__ Q, unsigned long long * flags)
{
  const Idx  I{ m.dim, m.b };
  const long nc = I.gcells(2);
  const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= long(m.nowned) * nc)
    return;
  const int  o = int(t / nc);
  const long ci = t - long(o) * nc;
  const int  b = m.b, dim = m.dim, n = b + 4;
  const int  i = int(ci % n) - 2, j = int((ci / n) % n) - 2, k = dim == 3 ? int(ci / (long(n) * n)) - 2 : 0;
  const long ncell = I.cells();
  const long gper = I.gfper();
  const double * G = Bg + long(o) * gper;
  auto bgv = [&](int c, int ii, int jj, int kk) { return G[I.gface(c, ii, jj, kk)]; };
  double u[8];
  const bool inside = i >= 0 && i < b && j >= 0 && j < b && (dim == 2 || (k >= 0 && k < b));
  if (inside)
  {
    const long cidx = I.cell(i, j, k);
    for (int v = 0; v < 5; ++v)
      u[v] = U[(long(o) * 5 + v) * ncell + cidx];
    if (dim == 2)
    {
      u[kA] = 0.5 * (bgv(0, i, j, 0) + bgv(0, i + 1, j, 0));
      u[kB] = 0.5 * (bgv(1, i, j, 0) + bgv(1, i, j + 1, 0));
      u[kC] = 0.5 * (bgv(2, i, j, 0) + bgv(2, i, j, 0));
    }
    else
    {
      u[kA] = 0.5 * (bgv(0, i, j, k) + bgv(0, i + 1, j, k));
      u[kB] = 0.5 * (bgv(1, i, j, k) + bgv(1, i, j + 1, k));
      u[kC] = 0.5 * (bgv(2, i, j, k) + bgv(2, i, j, k + 1));
    }
  }
  else
  {
    const int c[3] = { i, j, k };
    const Hit h = neighbour(m, o, c);
    const bool lin = h.rel < 0 && m.prolong == 1;
    if (h.rel > 0 || lin)
    {
      if (lin)
        extrapolate_linear_minmod(m, I, U, h, i, j, k, u);
      else
        for (int v = 0; v < 5; ++v)
        {
          const double * Un = U + (long(h.oct) * 5 + v) * ncell;
          double         s = 0;
          for (int dk = 0; dk < (dim == 3 ? 2 : 1); ++dk)
            for (int dj = 0; dj < 2; ++dj)
              for (int di = 0; di < 2; ++di)
                s += Un[I.cell(h.l[0] + di, h.l[1] + dj, h.l[2] + dk)];
          u[v] = s / double(1 << dim);
        }
      // the cell-centred field from this block's ghosted faces (finer neighbour, and the linear
      // prolongation: ConvertToPrimitivesVariablesFunctor.cpp:536-541, 623-628)
      if (dim == 2)
      {
        u[kA] = 0.5 * (bgv(0, i, j, 0) + bgv(0, i + 1, j, 0));
        u[kB] = 0.5 * (bgv(1, i, j, 0) + bgv(1, i, j + 1, 0));
        u[kC] = bgv(2, i, j, 0);
      }
      else
      {
        u[kA] = 0.5 * (bgv(0, i, j, k) + bgv(0, i + 1, j, k));
        u[kB] = 0.5 * (bgv(1, i, j, k) + bgv(1, i, j + 1, k));
        u[kC] = 0.5 * (bgv(2, i, j, k) + bgv(2, i, j, k + 1));
      }
    }
    else
    {
      const long cidx = I.cell(h.l[0], h.l[1], h.l[2]);
      for (int v = 0; v < 5; ++v)
        u[v] = U[(long(h.oct) * 5 + v) * ncell + cidx];
      const double * Bn = B + long(h.oct) * I.fper();
      const int      a = h.l[0], bb = h.l[1], cc = h.l[2];
      if (dim == 2)
      {
        u[kA] = 0.5 * (Bn[I.face(0, a, bb, 0)] + Bn[I.face(0, a + 1, bb, 0)]);
        u[kB] = 0.5 * (Bn[I.face(1, a, bb, 0)] + Bn[I.face(1, a, bb + 1, 0)]);
        u[kC] = 0.5 * (Bn[I.face(2, a, bb, 0)] + Bn[I.face(2, a, bb, 0)]);
      }
      else
      {
        u[kA] = 0.5 * (Bn[I.face(0, a, bb, cc)] + Bn[I.face(0, a + 1, bb, cc)]);
        u[kB] = 0.5 * (Bn[I.face(1, a, bb, cc)] + Bn[I.face(1, a, bb + 1, cc)]);
        u[kC] = 0.5 * (Bn[I.face(2, a, bb, cc)] + Bn[I.face(2, a, bb, cc + 1)]);
      }
    }
  }
  double    q[8];
  const double eint = to_primitive(u, q, g);
  if (abort_eint && eint < 0)
    atomicOr(flags, 1ull);
  double * Qo = Q + long(o) * 11 * nc;
  Qo[0 * nc + ci] = q[kD];
  Qo[1 * nc + ci] = q[kP];
  Qo[2 * nc + ci] = q[kU];
  Qo[3 * nc + ci] = q[kV];
  Qo[4 * nc + ci] = q[kW];
  if (dim == 2)
  {
    Qo[5 * nc + ci] = bgv(0, i, j, 0);
    Qo[8 * nc + ci] = bgv(0, i + 1, j, 0);
    Qo[6 * nc + ci] = bgv(1, i, j, 0);
    Qo[9 * nc + ci] = bgv(1, i, j + 1, 0);
    Qo[7 * nc + ci] = bgv(2, i, j, 0);
    Qo[10 * nc + ci] = bgv(2, i, j, 0);
  }
  else
  {
    Qo[5 * nc + ci] = bgv(0, i, j, k);
    Qo[8 * nc + ci] = bgv(0, i + 1, j, k);
    Qo[6 * nc + ci] = bgv(1, i, j, k);
    Qo[9 * nc + ci] = bgv(1, i, j + 1, k);
    Qo[7 * nc + ci] = bgv(2, i, j, k);
    Qo[10 * nc + ci] = bgv(2, i, j, k + 1);
  }
}

// Q (ghost 2) accessor
struct QView
{
  const double * p;
  Idx            I;
  long           nc;
  int            nv = 11; // variables per block: 11 here, 5 in the fused form's primitive records
  __device__ double operator()(int i, int j, int k, int v, int o) const
  {
    return p[(long(o) * nv + v) * nc + I.gcell(2, i, j, k)];
  }
};
// ghost-1 arrays with nv variables
struct G1View
{
  double * p;
  Idx      I;
  long     nc;
  int      nv;
  __device__ double & operator()(int i, int j, int k, int v, int o) const
  {
    return p[(long(o) * nv + v) * nc + I.gcell(1, i, j, k)];
  }
};

// ---------------------------------------------------------------------------
// K5: limited slopes (ComputeLimitedSlopesFunctor.cpp:225-314)
// ---------------------------------------------------------------------------
__device__ inline double
slope_mag(const QView & q, int i, int j, int k, int dir, int v1, int v2, int o, double st, int dim)
{
  int uv[3] = { 0, 0, 0 };
  uv[dir] = 1;
  if (v1 == 5 + dir)
    return q(i, j, k, v2, o) - q(i, j, k, v1, o);
  const double dc = (q(i, j, k, v1, o) + q(i, j, k, v2, o)) * 0.5;
  const double dp = (q(i + uv[0], j + uv[1], k + uv[2], v1, o) + q(i + uv[0], j + uv[1], k + uv[2], v2, o)) * 0.5;
  const double dm = (q(i - uv[0], j - uv[1], k - uv[2], v1, o) + q(i - uv[0], j - uv[1], k - uv[2], v2, o)) * 0.5;
  return limited_slope(dc, dp, dm, st);
}

__global__ void __launch_bounds__(kThreads)
k_slopes(DevMesh m, double st, const double * __restrict__ Qp, double * Sx, double * Sy, double * Sz)
{
  const Idx  I{ m.dim, m.b };
  const long nc = I.gcells(1);
  const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= long(m.nowned) * nc)
    return;
  const int  o = int(t / nc);
  const long ci = t - long(o) * nc;
  const int  b = m.b, dim = m.dim, n = b + 2;
  const int  i = int(ci % n) - 1, j = int((ci / n) % n) - 1, k = dim == 3 ? int(ci / (long(n) * n)) - 1 : 0;
  const QView q{ Qp, I, I.gcells(2) };
  const G1View sx{ Sx, I, nc, 8 }, sy{ Sy, I, nc, 8 }, sz{ Sz, I, nc, 8 };
  const int nh = dim == 3 ? 5 : 4;
  for (int v = 0; v < nh; ++v)
  {
    const double qc = q(i, j, k, v, o);
    sx(i, j, k, v, o) = limited_slope(qc, q(i + 1, j, k, v, o), q(i - 1, j, k, v, o), st);
    sy(i, j, k, v, o) = limited_slope(qc, q(i, j + 1, k, v, o), q(i, j - 1, k, v, o), st);
    if (dim == 3)
      sz(i, j, k, v, o) = limited_slope(qc, q(i, j, k + 1, v, o), q(i, j, k - 1, v, o), st);
  }
  for (int d = 0; d < dim; ++d)
  {
    const G1View & s = d == 0 ? sx : (d == 1 ? sy : sz);
    s(i, j, k, 5, o) = slope_mag(q, i, j, k, d, 5, 8, o, st, dim);
    s(i, j, k, 6, o) = slope_mag(q, i, j, k, d, 6, 9, o, st, dim);
    s(i, j, k, 7, o) = slope_mag(q, i, j, k, d, 7, 10, o, st, dim);
  }
}

// ---------------------------------------------------------------------------
// K6: Hancock predictor (ComputeUpdatedPrimitiveVariablesFunctor.cpp:284-546)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads)
k_predictor(DevMesh m, double scaling, Gas<double> gs, double dt, int rk2, double visc_mu,
            const double * __restrict__ Qp, const double * __restrict__ Sx, const double * __restrict__ Sy,
            const double * __restrict__ Sz, double * __restrict__ Q2)
{
  const Idx  I{ m.dim, m.b };
  const long nc = I.gcells(1);
  const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= long(m.nowned) * nc)
    return;
  const int  o = int(t / nc);
  const long ci = t - long(o) * nc;
  const int  b = m.b, dim = m.dim, n = b + 2;
  const int  i = int(ci % n) - 1, j = int((ci / n) % n) - 1, k = dim == 3 ? int(ci / (long(n) * n)) - 1 : 0;
  const QView q{ Qp, I, I.gcells(2) };
  const long  s0 = long(o) * 8 * nc + ci;
  auto        SX = [&](int v) { return Sx[s0 + v * nc]; };
  auto        SY = [&](int v) { return Sy[s0 + v * nc]; };
  auto        SZ = [&](int v) { return Sz[s0 + v * nc]; };
  const double dx = dx_of(m, scaling, o);
  const double dtdx = dt / dx, dtdy = dt / dx, dtdz = dt / dx;
  const double gamma = gs.gamma0;
  const double r = q(i, j, k, 0, o), p = q(i, j, k, 1, o);
  double       u = q(i, j, k, 2, o), v = q(i, j, k, 3, o), w = q(i, j, k, 4, o);
  const double A = (q(i, j, k, 5, o) + q(i, j, k, 8, o)) * 0.5;
  const double B = (q(i, j, k, 6, o) + q(i, j, k, 9, o)) * 0.5;
  const double C = (q(i, j, k, 7, o) + q(i, j, k, 10, o)) * 0.5;
  if (rk2)
  {
    // RK2_SSP: no half-step predictor, Q2 holds the cell-centred state
    // (ComputeUpdatedPrimitiveVariablesFunctor.cpp:396-407, 531-542)
    double * out = Q2 + s0;
    out[0 * nc] = r;
    out[2 * nc] = u;
    out[3 * nc] = v;
    out[4 * nc] = w;
    out[1 * nc] = p;
    out[5 * nc] = A;
    out[6 * nc] = B;
    out[7 * nc] = C;
    return;
  }
  const double drx = SX(0), dpx = SX(1), dux = SX(2), dvx = SX(3), dwx = SX(4), dAx = SX(5), dBx = SX(6), dCx = SX(7);
  const double dry = SY(0), dpy = SY(1), duy = SY(2), dvy = SY(3), dwy = SY(4), dAy = SY(5), dBy = SY(6), dCy = SY(7);
  double sr0, su0, sv0, sw0, sp0, sA0, sB0, sC0;
  if (dim == 2)
  {
    sr0 = (-u * drx - dux * r) * dtdx + (-v * dry - dvy * r) * dtdy;
    su0 = (-u * dux - dpx / r - B * dBx / r - C * dCx / r) * dtdx + (-v * duy + B * dAy / r) * dtdy;
    sv0 = (-u * dvx + A * dBx / r) * dtdx + (-v * dvy - dpy / r - A * dAy / r - C * dCy / r) * dtdy;
    sw0 = (-u * dwx + A * dCx / r) * dtdx + (-v * dwy + B * dCy / r) * dtdy;
    sp0 = (-u * dpx - dux * gamma * p) * dtdx + (-v * dpy - dvy * gamma * p) * dtdy;
    sA0 = (u * dBy + B * duy - v * dAy - A * dvy) * dtdy;
    sB0 = (-u * dBx - B * dux + v * dAx + A * dvx) * dtdx;
    sC0 = (w * dAx + A * dwx - u * dCx - C * dux) * dtdx + (-v * dCy - C * dvy + w * dBy + B * dwy) * dtdy;
  }
  else
  {
    const double drz = SZ(0), dpz = SZ(1), duz = SZ(2), dvz = SZ(3), dwz = SZ(4), dAz = SZ(5), dBz = SZ(6),
                 dCz = SZ(7);
    sr0 = (-u * drx - dux * r) * dtdx + (-v * dry - dvy * r) * dtdy + (-w * drz - dwz * r) * dtdz;
    su0 = (-u * dux - (dpx + B * dBx + C * dCx) / r) * dtdx + (-v * duy + B * dAy / r) * dtdy +
          (-w * duz + C * dAz / r) * dtdz;
    sv0 = (-u * dvx + A * dBx / r) * dtdx + (-v * dvy - (dpy + A * dAy + C * dCy) / r) * dtdy +
          (-w * dvz + C * dBz / r) * dtdz;
    sw0 = (-u * dwx + A * dCx / r) * dtdx + (-v * dwy + B * dCy / r) * dtdy +
          (-w * dwz - (dpz + A * dAz + B * dBz) / r) * dtdz;
    sp0 = (-u * dpx - dux * gamma * p) * dtdx + (-v * dpy - dvy * gamma * p) * dtdy + (-w * dpz - dwz * gamma * p) * dtdz;
    sA0 = (u * dBy + B * duy - v * dAy - A * dvy) * dtdy + (u * dCz + C * duz - w * dAz - A * dwz) * dtdz;
    sB0 = (v * dAx + A * dvx - u * dBx - B * dux) * dtdx + (v * dCz + C * dvz - w * dBz - B * dwz) * dtdz;
    sC0 = (w * dAx + A * dwx - u * dCx - C * dux) * dtdx + (w * dBy + B * dwy - v * dCy - C * dvy) * dtdy;
  }
  // gravity, half time step, after the source terms (ComputeUpdatedPrimitiveVariablesFunctor.cpp:372-377,
  // 505-511)
  if (gs.grav)
  {
    u += gs.g[0] * 0.5 * dt;
    v += gs.g[1] * 0.5 * dt;
    if (dim == 3)
      w += gs.g[2] * 0.5 * dt;
  }
  // viscous predictor, after the gravity kick (add_viscous_predictor_{2d,3d},
  // ComputeUpdatedPrimitiveVariablesFunctor.cpp:178-272, called at :380-383, 514-517): per
  // velocity component (2D: u and v), per axis, 1/2 nu (second difference) / dx * dt / dx with
  // nu = mu / rho; visc_mu < 0 leaves it out
  if (visc_mu >= 0.0)
  {
    const double nu = visc_mu / r;
    const double dtd[3] = { dtdx, dtdy, dtdz };
    double *     vel[3] = { &u, &v, &w };
    for (int c = 0; c < dim; ++c)
      for (int a = 0; a < dim; ++a)
      {
        const int    e0 = a == 0, e1 = a == 1, e2 = a == 2;
        const double d2 = q(i + e0, j + e1, k + e2, 2 + c, o) + q(i - e0, j - e1, k - e2, 2 + c, o) -
                          2 * q(i, j, k, 2 + c, o);
        *vel[c] += 0.5 * nu * d2 / dx * dtd[a];
      }
  }
  double * out = Q2 + s0;
  out[0 * nc] = r + 0.5 * sr0;
  out[2 * nc] = u + 0.5 * su0;
  out[3 * nc] = v + 0.5 * sv0;
  out[4 * nc] = w + 0.5 * sw0;
  out[1 * nc] = p + 0.5 * sp0;
  out[5 * nc] = A + 0.5 * sA0;
  out[6 * nc] = B + 0.5 * sB0;
  out[7 * nc] = C + 0.5 * sC0;
}

// ---------------------------------------------------------------------------
// K7: edge electric field E = v x B at t_n (ComputeElectricFieldFunctor.cpp:97-235)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads)
k_elec_field(DevMesh m, const double * __restrict__ Qp, double * __restrict__ E)
{
  const Idx  I{ m.dim, m.b };
  const long nc = I.gcells(2);
  const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= long(m.nowned) * nc)
    return;
  const int  o = int(t / nc);
  const long ci = t - long(o) * nc;
  const int  b = m.b, dim = m.dim, n = b + 4;
  const int  i = int(ci % n) - 2, j = int((ci / n) % n) - 2, k = dim == 3 ? int(ci / (long(n) * n)) - 2 : 0;
  const QView q{ Qp, I, nc };
  if (i <= -2 || j <= -2 || (dim == 3 && k <= -2))
    return;
  if (dim == 2)
  {
    const double u = 0.25 * (q(i - 1, j - 1, 0, 2, o) + q(i - 1, j, 0, 2, o) + q(i, j - 1, 0, 2, o) + q(i, j, 0, 2, o));
    const double v = 0.25 * (q(i - 1, j - 1, 0, 3, o) + q(i - 1, j, 0, 3, o) + q(i, j - 1, 0, 3, o) + q(i, j, 0, 3, o));
    const double A = 0.5 * (q(i, j - 1, 0, 5, o) + q(i, j, 0, 5, o));
    const double B = 0.5 * (q(i - 1, j, 0, 6, o) + q(i, j, 0, 6, o));
    E[long(o) * nc + ci] = u * B - v * A;
    return;
  }
  {
    const double v = 0.25 * (q(i, j - 1, k - 1, 3, o) + q(i, j - 1, k, 3, o) + q(i, j, k - 1, 3, o) + q(i, j, k, 3, o));
    const double w = 0.25 * (q(i, j - 1, k - 1, 4, o) + q(i, j - 1, k, 4, o) + q(i, j, k - 1, 4, o) + q(i, j, k, 4, o));
    const double B = 0.5 * (q(i, j, k - 1, 6, o) + q(i, j, k, 6, o));
    const double C = 0.5 * (q(i, j - 1, k, 7, o) + q(i, j, k, 7, o));
    E[(long(o) * 3 + 0) * nc + ci] = v * C - w * B;
  }
  {
    const double u = 0.25 * (q(i - 1, j, k - 1, 2, o) + q(i - 1, j, k, 2, o) + q(i, j, k - 1, 2, o) + q(i, j, k, 2, o));
    const double w = 0.25 * (q(i - 1, j, k - 1, 4, o) + q(i - 1, j, k, 4, o) + q(i, j, k - 1, 4, o) + q(i, j, k, 4, o));
    const double A = 0.5 * (q(i, j, k - 1, 5, o) + q(i, j, k, 5, o));
    const double C = 0.5 * (q(i - 1, j, k, 7, o) + q(i, j, k, 7, o));
    E[(long(o) * 3 + 1) * nc + ci] = w * A - u * C;
  }
  {
    const double u = 0.25 * (q(i - 1, j - 1, k, 2, o) + q(i - 1, j, k, 2, o) + q(i, j - 1, k, 2, o) + q(i, j, k, 2, o));
    const double v = 0.25 * (q(i - 1, j - 1, k, 3, o) + q(i - 1, j, k, 3, o) + q(i, j - 1, k, 3, o) + q(i, j, k, 3, o));
    const double A = 0.5 * (q(i, j - 1, k, 5, o) + q(i, j, k, 5, o));
    const double B = 0.5 * (q(i - 1, j, k, 6, o) + q(i, j, k, 6, o));
    E[(long(o) * 3 + 2) * nc + ci] = u * B - v * A;
  }
}

// ---------------------------------------------------------------------------
// K8: face-B half-step source (ComputeSourceFaceMagFunctor.cpp:114-203)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads)
k_sfacemag(DevMesh m, double scaling, double dt, const double * __restrict__ E, double * __restrict__ sF)
{
  const Idx  I{ m.dim, m.b };
  const long nc = I.gcells(1);
  const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= long(m.nowned) * nc)
    return;
  const int  o = int(t / nc);
  const long ci = t - long(o) * nc;
  const int  b = m.b, dim = m.dim, n = b + 2;
  const int  i = int(ci % n) - 1, j = int((ci / n) % n) - 1, k = dim == 3 ? int(ci / (long(n) * n)) - 1 : 0;
  const long ne = I.gcells(2);
  const int  ncE = dim == 3 ? 3 : 1;
  auto e = [&](int ii, int jj, int kk, int c) { return E[(long(o) * ncE + c) * ne + I.gcell(2, ii, jj, kk)]; };
  const double dx = dx_of(m, scaling, o);
  const double dtdx = dt / dx, dtdy = dt / dx, dtdz = dt / dx;
  if (dim == 2)
  {
    sF[(long(o) * 2 + 0) * nc + ci] = (e(i, j + 1, 0, 0) - e(i, j, 0, 0)) * 0.5 * dtdy;
    sF[(long(o) * 2 + 1) * nc + ci] = -(e(i + 1, j, 0, 0) - e(i, j, 0, 0)) * 0.5 * dtdx;
    return;
  }
  sF[(long(o) * 3 + 0) * nc + ci] =
    (e(i, j + 1, k, 2) - e(i, j, k, 2)) * 0.5 * dtdy - (e(i, j, k + 1, 1) - e(i, j, k, 1)) * 0.5 * dtdz;
  sF[(long(o) * 3 + 1) * nc + ci] =
    -(e(i + 1, j, k, 2) - e(i, j, k, 2)) * 0.5 * dtdx + (e(i, j, k + 1, 0) - e(i, j, k, 0)) * 0.5 * dtdz;
  sF[(long(o) * 3 + 2) * nc + ci] =
    (e(i + 1, j, k, 1) - e(i, j, k, 1)) * 0.5 * dtdx - (e(i, j + 1, k, 0) - e(i, j, k, 0)) * 0.5 * dtdy;
}

// ---------------------------------------------------------------------------
// K9: HLLD fluxes along one direction (ComputeHydroFluxesAndStoreFunctor.cpp:218-503)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads)
k_fluxes(DevMesh m, Gas<double> g, double scaling, double dt, int dir, const double * __restrict__ Qp,
         const double * __restrict__ Q2, const double * __restrict__ Sd, const double * __restrict__ sF,
         double * __restrict__ Fl)
{
  const Idx  I{ m.dim, m.b };
  const int  b = m.b, dim = m.dim;
  int        sh[3] = { b, b, dim == 3 ? b : 1 };
  sh[dir] = b + 1;
  const long nf = long(sh[0]) * sh[1] * sh[2];
  const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= long(m.nowned) * nf)
    return;
  const int  o = int(t / nf);
  const long ci = t - long(o) * nf;
  const int  i = int(ci % sh[0]), j = int((ci / sh[0]) % sh[1]), k = dim == 3 ? int(ci / (long(sh[0]) * sh[1])) : 0;
  const long n1 = I.gcells(1);
  const QView q{ Qp, I, I.gcells(2) };
  int        dl[3] = { 0, 0, 0 };
  dl[dir] = 1;
  const long cR = long(o) * 8 * n1 + I.gcell(1, i, j, k);
  const long cL = long(o) * 8 * n1 + I.gcell(1, i - dl[0], j - dl[1], k - dl[2]);
  double     qL[8], qR[8], f[8];
  for (int v = 0; v < 8; ++v)
  {
    qL[v] = Q2[cL + v * n1] + 0.5 * Sd[cL + v * n1];
    qR[v] = Q2[cR + v * n1] + -0.5 * Sd[cR + v * n1];
  }
  qL[kD] = fmax(g.smallr, qL[kD]);
  qL[kP] = fmax(g.smallp * qL[kD], qL[kP]);
  qR[kD] = fmax(g.smallr, qR[kD]);
  qR[kP] = fmax(g.smallp * qR[kD], qR[kP]);
  const int    ncS = dim == 3 ? 3 : 2;
  const double bn = q(i, j, k, 5 + dir, o) + sF[(long(o) * ncS + dir) * n1 + I.gcell(1, i, j, k)];
  qR[kA + dir] = bn;
  qL[kA + dir] = bn;
  if (dir != 0)
  {
    const int vt = kU + dir, bt = kA + dir;
    double    x;
    x = qL[kU]; qL[kU] = qL[vt]; qL[vt] = x;
    x = qR[kU]; qR[kU] = qR[vt]; qR[vt] = x;
    x = qL[kA]; qL[kA] = qL[bt]; qL[bt] = x;
    x = qR[kA]; qR[kA] = qR[bt]; qR[bt] = x;
  }
  hlld(qL, qR, f, g);
  if (dir != 0)
  {
    const int vt = kU + dir, bt = kA + dir;
    double    x;
    x = f[kU]; f[kU] = f[vt]; f[vt] = x;
    x = f[kA]; f[kA] = f[bt]; f[bt] = x;
  }
  const double dtdS = dt / dx_of(m, scaling, o);
  double *     out = Fl + long(o) * 8 * nf + ci;
  for (int v = 0; v < 8; ++v)
  {
    if (dim == 2 && (v == kA || v == kB))
      continue;
    out[v * nf] = f[v] * dtdS;
  }
}

// ---------------------------------------------------------------------------
// K9v: viscous fluxes (ComputeViscousFluxesAndStoreFunctor.cpp:96-352) at the low face of every
// flux cell (i,j,k) along dir, from Q's velocities: every gradient is the backward difference
// at cell (i,j,k) along each axis; the stress is mu (grad v + grad v^T) minus its trace part
// (1/2 in 2D, 1/3 in 3D); the energy flux uses cell (i,j,k)'s velocity.  Stored times dt/dx
// where k_update reads fluxes; the density (and 2D Bz) flux is 0 (set_flux, .h:208-259).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads)
k_viscous_fluxes(DevMesh m, double scaling, double dt, double mu, int dir, const double * __restrict__ Qp,
                 int qvars, double * __restrict__ Fl)
{
  const Idx  I{ m.dim, m.b };
  const int  b = m.b, dim = m.dim;
  int        sh[3] = { b, b, dim == 3 ? b : 1 };
  sh[dir] = b + 1;
  const long nf = long(sh[0]) * sh[1] * sh[2];
  const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= long(m.nowned) * nf)
    return;
  const int  o = int(t / nf);
  const long ci = t - long(o) * nf;
  const int  i = int(ci % sh[0]), j = int((ci / sh[0]) % sh[1]), k = dim == 3 ? int(ci / (long(sh[0]) * sh[1])) : 0;
  const QView  q{ Qp, I, I.gcells(2), qvars };
  const double dx = dx_of(m, scaling, o);
  // g[c][a] = d v_c / d x_a
  double g[3][3] = { { 0, 0, 0 }, { 0, 0, 0 }, { 0, 0, 0 } };
  const double vel[3] = { q(i, j, k, kU, o), q(i, j, k, kV, o), dim == 3 ? q(i, j, k, kW, o) : 0.0 };
  for (int a = 0; a < dim; ++a)
  {
    const int e0 = a == 0, e1 = a == 1, e2 = a == 2;
    for (int c = 0; c < dim; ++c)
      g[c][a] = (q(i, j, k, kU + c, o) - q(i - e0, j - e1, k - e2, kU + c, o)) / dx;
  }
  double fU, fV, fW = 0.0, fE;
  if (dim == 2)
  {
    const double dv = g[0][0] + g[1][1];
    const double txy = mu * (g[1][0] + g[0][1]);
    if (dir == 0)
    {
      const double txx = 2 * mu * (g[0][0] - 0.5 * dv);
      fU = -txx;
      fV = -txy;
      fE = -vel[0] * txx - vel[1] * txy;
    }
    else
    {
      const double tyy = 2 * mu * (g[1][1] - 0.5 * dv);
      fU = -txy;
      fV = -tyy;
      fE = -vel[0] * txy - vel[1] * tyy;
    }
  }
  else
  {
    const double third = 1.0 / 3.0;
    const double dv = g[0][0] + g[1][1] + g[2][2];
    double       t0, t1, t2;
    if (dir == 0)
    {
      t0 = 2 * mu * (g[0][0] - third * dv);
      t1 = mu * (g[1][0] + g[0][1]);
      t2 = mu * (g[2][0] + g[0][2]);
    }
    else if (dir == 1)
    {
      t0 = mu * (g[0][1] + g[1][0]);
      t1 = 2 * mu * (g[1][1] - third * dv);
      t2 = mu * (g[2][1] + g[1][2]);
    }
    else
    {
      t0 = mu * (g[0][2] + g[2][0]);
      t1 = mu * (g[1][2] + g[2][1]);
      t2 = 2 * mu * (g[2][2] - third * dv);
    }
    fU = -t0;
    fV = -t1;
    fW = -t2;
    fE = -vel[0] * t0 - vel[1] * t1 - vel[2] * t2;
  }
  const double dtdS = dt / dx;
  double *     out = Fl + long(o) * 8 * nf + ci;
  out[kD * nf] = 0.0;
  out[kP * nf] = fE * dtdS;
  out[kU * nf] = fU * dtdS;
  out[kV * nf] = fV * dtdS;
  out[kW * nf] = fW * dtdS;
  out[kC * nf] = 0.0;
  if (dim == 3)
  {
    out[kA * nf] = 0.0;
    out[kB * nf] = 0.0;
  }
}

// ---------------------------------------------------------------------------
// K10: conservative update (ReadFluxesAndConservativeUpdateFunctor.cpp:105-436)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads)
k_update(DevMesh m, int dir, const double * __restrict__ Fl, double * __restrict__ U2, double * __restrict__ B2)
{
  const Idx  I{ m.dim, m.b };
  const int  b = m.b, dim = m.dim;
  const long ncell = I.cells();
  const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= long(m.nowned) * ncell)
    return;
  const int  o = int(t / ncell);
  const long ci = t - long(o) * ncell;
  const int  c[3] = { int(ci % b), int((ci / b) % b), dim == 3 ? int(ci / (long(b) * b)) : 0 };
  int        sh[3] = { b, b, dim == 3 ? b : 1 };
  sh[dir] = b + 1;
  const long nf = long(sh[0]) * sh[1] * sh[2];
  auto fidx = [&](int i, int j, int k) { return i + long(sh[0]) * (j + long(sh[1]) * k); };
  const int  vars[6] = { kD, kP, kU, kV, kW, kC };
  const int  nv = dim == 2 ? 6 : 5;
  const bool lo_fine = m.conformal[long(o) * 2 * dim + 2 * dir] != 0;
  const bool hi_fine = m.conformal[long(o) * 2 * dim + 2 * dir + 1] != 0;
  const double vr = double(1 << dim);
  double acc[8] = { 0, 0, 0, 0, 0, 0, 0, 0 };
  for (int side = 0; side < 2; ++side)
  {
    const bool at_border = side == 0 ? (c[dir] == 0 && lo_fine) : (c[dir] == b - 1 && hi_fine);
    if (at_border)
    {
      int cc[3] = { c[0], c[1], c[2] };
      cc[dir] = side == 0 ? -1 : b;
      const Hit h = neighbour(m, o, cc);
      int       base[3] = { h.l[0], h.l[1], h.l[2] };
      base[dir] = side == 0 ? b : 0;
      int tt[2], nt = 0;
      for (int d = 0; d < dim; ++d)
        if (d != dir)
          tt[nt++] = d;
      const double * Fn = Fl + long(h.oct) * 8 * nf;
      for (int vi = 0; vi < nv; ++vi)
      {
        const int v = vars[vi];
        double    s = 0;
        if (dim == 2)
        {
          for (int a = 0; a < 2; ++a)
          {
            int q[3] = { base[0], base[1], 0 };
            q[tt[0]] += a;
            s += Fn[v * nf + fidx(q[0], q[1], 0)];
          }
        }
        else
        {
          for (int a1 = 0; a1 < 2; ++a1)
            for (int a0 = 0; a0 < 2; ++a0)
            {
              int q[3] = { base[0], base[1], base[2] };
              q[tt[0]] += a0;
              q[tt[1]] += a1;
              s += Fn[v * nf + fidx(q[0], q[1], q[2])];
            }
        }
        acc[v] = side == 0 ? acc[v] + s / vr : acc[v] - s / vr;
      }
    }
    else
    {
      int q[3] = { c[0], c[1], c[2] };
      q[dir] += side;
      const double * Fo = Fl + long(o) * 8 * nf + fidx(q[0], q[1], q[2]);
      for (int vi = 0; vi < nv; ++vi)
      {
        const int v = vars[vi];
        acc[v] = side == 0 ? acc[v] + Fo[v * nf] : acc[v] - Fo[v * nf];
      }
    }
  }
  for (int v = 0; v < 5; ++v)
    U2[(long(o) * 5 + v) * ncell + ci] += acc[v];
  if (dim == 2)
    B2[long(o) * I.fper() + I.face(2, c[0], c[1], 0)] += acc[kC];
}

// ---------------------------------------------------------------------------
// K11: CT EMF from the 2D HLLD magnetic Riemann solver (ComputeEmfAndStoreFunctor.cpp:105-590)
// ---------------------------------------------------------------------------
struct EmfIn
{
  QView          q;
  const double * Q2;
  const double * S[3];
  const double * sF; // face-B half-step source, added to the edge face values under HANCOCK only
  long           n1;
  int            ncS;
  Gas<double>    g;
  int            hancock;
};

__device__ inline double
qslope(const EmfIn & in, int i, int j, int k, int var, int o, int dir)
{
  const QView & q = in.q;
  if (dir == 0)
    return limited_slope(q(i, j, k, var, o), q(i + 1, j, k, var, o), q(i - 1, j, k, var, o), in.g.slope_type);
  if (dir == 1)
    return limited_slope(q(i, j, k, var, o), q(i, j + 1, k, var, o), q(i, j - 1, k, var, o), in.g.slope_type);
  return limited_slope(q(i, j, k, var, o), q(i, j, k + 1, var, o), q(i, j, k - 1, var, o), in.g.slope_type);
}

// corner kinds: 0 LB, 1 RT, 2 RB, 3 LT
__device__ inline void
edge_state_2d(const EmfIn & in, const Idx & I, int is, int js, int o, int kind, double * qo)
{
  const long c0 = long(o) * 8 * in.n1 + I.gcell(1, is, js, 0);
  for (int v = 0; v < 8; ++v)
    qo[v] = in.Q2[c0 + v * in.n1];
  int sx = 1, sy = 1, ia = is, ja = js, ib = is, jb = js, sa = 1, sb = 1;
  if (kind == 0)
  {
    sx = -1; sy = -1; sa = -1; sb = -1;
  }
  else if (kind == 1)
  {
    ia = is + 1; jb = js + 1;
  }
  else if (kind == 2)
  {
    ia = is + 1; sy = -1; sa = -1;
  }
  else
  {
    jb = js + 1; sx = -1; sb = -1;
  }
  double A = in.q(ia, ja, 0, 5, o);
  double B = in.q(ib, jb, 0, 6, o);
  if (in.hancock) // ComputeEmfAndStoreFunctor.cpp:189-193
  {
    A += in.sF[(long(o) * 2 + 0) * in.n1 + I.gcell(1, ia, ja, 0)];
    B += in.sF[(long(o) * 2 + 1) * in.n1 + I.gcell(1, ib, jb, 0)];
  }
  const double dAy = qslope(in, ia, ja, 0, 5, o, 1);
  const double dBx = qslope(in, ib, jb, 0, 6, o, 0);
  const double fx = double(sx), fy = double(sy);
  for (int v = 0; v < 5; ++v)
    qo[v] += 0.5 * (fx * in.S[0][c0 + v * in.n1] + fy * in.S[1][c0 + v * in.n1]);
  qo[kA] = A + 0.5 * (double(sa) * dAy);
  qo[kB] = B + 0.5 * (double(sb) * dBx);
  qo[kC] += 0.5 * (fx * in.S[0][c0 + 7 * in.n1] + fy * in.S[1][c0 + 7 * in.n1]);
  qo[kD] = fmax(in.g.smallr, qo[kD]);
  qo[kP] = fmax(in.g.smallp * qo[kD], qo[kP]);
}

__device__ inline void
edge_state_3d(const EmfIn & in, const Idx & I, int is, int js, int ks, int o, int kind, int d0, int d1, double * qo)
{
  const long c0 = long(o) * 8 * in.n1 + I.gcell(1, is, js, ks);
  for (int v = 0; v < 8; ++v)
    qo[v] = in.Q2[c0 + v * in.n1];
  int s0 = 1, s1 = 1, b0 = 1, b1 = 1;
  int p0[3] = { is, js, ks }, p1[3] = { is, js, ks };
  if (kind == 0)
  {
    s0 = -1; s1 = -1; b0 = -1; b1 = -1;
  }
  else if (kind == 1)
  {
    p0[d0] += 1; p1[d1] += 1;
  }
  else if (kind == 2)
  {
    p0[d0] += 1; s1 = -1; b0 = -1;
  }
  else
  {
    p1[d1] += 1; s0 = -1; b1 = -1;
  }
  double B0 = in.q(p0[0], p0[1], p0[2], 5 + d0, o);
  double B1 = in.q(p1[0], p1[1], p1[2], 5 + d1, o);
  if (in.hancock) // ComputeEmfAndStoreFunctor.cpp:297-301
  {
    B0 += in.sF[(long(o) * 3 + d0) * in.n1 + I.gcell(1, p0[0], p0[1], p0[2])];
    B1 += in.sF[(long(o) * 3 + d1) * in.n1 + I.gcell(1, p1[0], p1[1], p1[2])];
  }
  const double dB0d1 = qslope(in, p0[0], p0[1], p0[2], 5 + d0, o, d1);
  const double dB1d0 = qslope(in, p1[0], p1[1], p1[2], 5 + d1, o, d0);
  const double * S0 = in.S[d0];
  const double * S1 = in.S[d1];
  const double   f0 = double(s0), f1 = double(s1);
  for (int v = 0; v < 5; ++v)
    qo[v] += 0.5 * (f0 * S0[c0 + v * in.n1] + f1 * S1[c0 + v * in.n1]);
  auto tr = [&](int v) { return 0.5 * (f0 * S0[c0 + v * in.n1] + f1 * S1[c0 + v * in.n1]); };
  // the reference applies these four assignments in sequence (.cpp:324-362)
  if (d0 == 0)
    qo[kA] = B0 + 0.5 * (double(b0) * dB0d1);
  else
    qo[kA] += tr(kA);
  if (d0 == 1)
    qo[kB] = B0 + 0.5 * (double(b0) * dB0d1);
  else
    qo[kB] += tr(kB);
  if (d1 == 1)
    qo[kB] = B1 + 0.5 * (double(b1) * dB1d0);
  else
    qo[kB] += tr(kB);
  if (d1 == 2)
    qo[kC] = B1 + 0.5 * (double(b1) * dB1d0);
  else
    qo[kC] += tr(kC);
  qo[kD] = fmax(in.g.smallr, qo[kD]);
  qo[kP] = fmax(in.g.smallp * qo[kD], qo[kP]);
}

__global__ void __launch_bounds__(kThreads)
k_emf(DevMesh m, Gas<double> g, double scaling, double dt, const double * __restrict__ Qp,
      const double * __restrict__ Q2, const double * __restrict__ Sx, const double * __restrict__ Sy,
      const double * __restrict__ Sz, const double * __restrict__ sF, int hancock, double * __restrict__ emf)
{
  const Idx  I{ m.dim, m.b };
  const int  b = m.b, dim = m.dim;
  const long ne = I.ecells();
  const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= long(m.nowned) * ne)
    return;
  const int  o = int(t / ne);
  const long ci = t - long(o) * ne;
  const int  n = b + 1;
  const int  i = int(ci % n), j = int((ci / n) % n), k = dim == 3 ? int(ci / (long(n) * n)) : 0;
  EmfIn      in{ QView{ Qp, I, I.gcells(2) }, Q2, { Sx, Sy, Sz }, sF, I.gcells(1), dim == 3 ? 3 : 2, g, hancock };
  const double dtdx = dt / dx_of(m, scaling, o);
  double       qe[4][8];
  if (dim == 2)
  {
    edge_state_2d(in, I, i, j, o, 0, qe[kLB]);
    edge_state_2d(in, I, i - 1, j - 1, o, 1, qe[kRT]);
    edge_state_2d(in, I, i - 1, j, o, 2, qe[kRB]);
    edge_state_2d(in, I, i, j - 1, o, 3, qe[kLT]);
    emf[long(o) * ne + ci] = edge_emf(qe, 2, g) * dtdx;
    return;
  }
  if (k < b)
  {
    edge_state_3d(in, I, i, j, k, o, 0, 0, 1, qe[kLB]);
    edge_state_3d(in, I, i - 1, j - 1, k, o, 1, 0, 1, qe[kRT]);
    edge_state_3d(in, I, i - 1, j, k, o, 2, 0, 1, qe[kRB]);
    edge_state_3d(in, I, i, j - 1, k, o, 3, 0, 1, qe[kLT]);
    emf[(long(o) * 3 + 2) * ne + ci] = edge_emf(qe, 2, g) * dtdx;
  }
  if (j < b)
  {
    edge_state_3d(in, I, i, j, k, o, 0, 0, 2, qe[kLB]);
    edge_state_3d(in, I, i - 1, j, k - 1, o, 1, 0, 2, qe[kRT]);
    // RB and LT swap places for edges along y (ComputeEmfAndStoreFunctor.cpp:550-560)
    edge_state_3d(in, I, i - 1, j, k, o, 2, 0, 2, qe[kLT]);
    edge_state_3d(in, I, i, j, k - 1, o, 3, 0, 2, qe[kRB]);
    emf[(long(o) * 3 + 1) * ne + ci] = edge_emf(qe, 1, g) * dtdx;
  }
  if (i < b)
  {
    edge_state_3d(in, I, i, j, k, o, 0, 1, 2, qe[kLB]);
    edge_state_3d(in, I, i, j - 1, k - 1, o, 1, 1, 2, qe[kRT]);
    edge_state_3d(in, I, i, j - 1, k, o, 2, 1, 2, qe[kRB]);
    edge_state_3d(in, I, i, j, k - 1, o, 3, 1, 2, qe[kLT]);
    emf[(long(o) * 3 + 0) * ne + ci] = edge_emf(qe, 0, g) * dtdx;
  }
}

// ---------------------------------------------------------------------------
// K13: border-edge EMF correction from the precomputed gather list
// (CorrectEmfAtBlockBorderFunctor.cpp:94-360)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads)
k_emf_fix(int nfix, const EmfFix * __restrict__ fix, const int64_t * __restrict__ src,
          const double * __restrict__ emf, double * __restrict__ emf2)
{
  const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= nfix)
    return;
  const EmfFix f = fix[t];
  double       s = 0;
  for (int n = 0; n < f.nsrc; ++n)
    s += emf[src[f.first + n]];
  emf2[f.dst] = s / double(f.cnt) / f.ratio;
}

// ---------------------------------------------------------------------------
// K14: face-B update by the discrete curl of the EMF (ReadEmfAndUpdateFunctor.cpp:58-163)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads)
k_emf_update(DevMesh m, const double * __restrict__ emf2, double * __restrict__ B2)
{
  const Idx  I{ m.dim, m.b };
  const int  b = m.b, dim = m.dim;
  const long per = I.fper();
  const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= long(m.nowned) * per)
    return;
  const int  o = int(t / per);
  long       r = t - long(o) * per;
  int        c = 0;
  while (c < 2 && r >= I.fsize(c))
  {
    r -= I.fsize(c);
    ++c;
  }
  const int  nx = c == 0 ? b + 1 : b;
  const int  ny = c == 1 ? b + 1 : b;
  const int  i = int(r % nx), j = int((r / nx) % ny), k = dim == 3 ? int(r / (long(nx) * ny)) : 0;
  const long ne = I.ecells();
  double &   Bv = B2[t];
  if (dim == 2)
  {
    const double * e = emf2 + long(o) * ne;
    if (c == 0)
      Bv += (e[I.edge(i, j + 1, 0)] - e[I.edge(i, j, 0)]);
    else if (c == 1)
      Bv -= (e[I.edge(i + 1, j, 0)] - e[I.edge(i, j, 0)]);
    return;
  }
  const double * ex = emf2 + (long(o) * 3 + 0) * ne;
  const double * ey = emf2 + (long(o) * 3 + 1) * ne;
  const double * ez = emf2 + (long(o) * 3 + 2) * ne;
  double         dB = 0;
  if (c == 0)
  {
    dB += (ez[I.edge(i, j + 1, k)] - ez[I.edge(i, j, k)]);
    dB -= (ey[I.edge(i, j, k + 1)] - ey[I.edge(i, j, k)]);
  }
  else if (c == 1)
  {
    dB += (ex[I.edge(i, j, k + 1)] - ex[I.edge(i, j, k)]);
    dB -= (ez[I.edge(i + 1, j, k)] - ez[I.edge(i, j, k)]);
  }
  else
  {
    dB += (ey[I.edge(i + 1, j, k)] - ey[I.edge(i, j, k)]);
    dB -= (ex[I.edge(i, j + 1, k)] - ex[I.edge(i, j, k)]);
  }
  Bv += dB;
}

// ---------------------------------------------------------------------------
// K15: inverse dt reduction (ComputeDtMHDFunctor.cpp:86-151, Kokkos::Max): wave64 butterfly
// max, one atomic per workgroup on the bit pattern (inv_dt >= 0 so bits order like values).
// A NaN candidate is left out of the max and raises error bit 2.
// ---------------------------------------------------------------------------
// signal speed over dx of one cell (ComputeDtMHD); IX is the run-time or compile-time index type
template <class IX>
__device__ inline double
cell_inv_dt(const DevMesh & m, const IX & I, const Gas<double> & g, double scaling, const double * __restrict__ U,
            const double * __restrict__ B, int o, long ci, int i, int j, int k)
{
  const int      b = I.b, dim = I.dim;
  const long     ncell = I.cells();
  const double * Bo = B + long(o) * I.fper();
  double         u[8], q[8];
  for (int v = 0; v < 5; ++v)
    u[v] = U[(long(o) * 5 + v) * ncell + ci];
  if (dim == 2)
  {
    u[kA] = 0.5 * (Bo[I.face(0, i, j, 0)] + Bo[I.face(0, i + 1, j, 0)]);
    u[kB] = 0.5 * (Bo[I.face(1, i, j, 0)] + Bo[I.face(1, i, j + 1, 0)]);
    u[kC] = 0.5 * (Bo[I.face(2, i, j, 0)] + Bo[I.face(2, i, j, 0)]);
  }
  else
  {
    u[kA] = 0.5 * (Bo[I.face(0, i, j, k)] + Bo[I.face(0, i + 1, j, k)]);
    u[kB] = 0.5 * (Bo[I.face(1, i, j, k)] + Bo[I.face(1, i, j + 1, k)]);
    u[kC] = 0.5 * (Bo[I.face(2, i, j, k)] + Bo[I.face(2, i, j, k + 1)]);
  }
  to_primitive(u, q, g);
  const double b2 = q[kA] * q[kA] + q[kB] * q[kB] + q[kC] * q[kC];
  const double c2 = g.gamma0 * q[kP] / q[kD];
  const double d2 = 0.5 * (b2 / q[kD] + c2);
  const double dx = 1.0 / double(long(b) << m.level[o]) * scaling;
  double       v[3];
  v[0] = sqrt(d2 + sqrt(d2 * d2 - c2 * q[kA] * q[kA] / q[kD])) + fabs(q[kU]);
  v[1] = sqrt(d2 + sqrt(d2 * d2 - c2 * q[kB] * q[kB] / q[kD])) + fabs(q[kV]);
  if (dim == 2)
    return cell_rate<2>(v, dx, g);
  v[2] = sqrt(d2 + sqrt(d2 * d2 - c2 * q[kC] * q[kC] / q[kD])) + fabs(q[kW]);
  return cell_rate<3>(v, dx, g);
}

// block max of the candidates, one atomic per workgroup; NaN candidates raise flag bit 2
__device__ inline void
reduce_inv_dt(double cand, bool nan, unsigned long long * red)
{
  for (int off = 32; off > 0; off >>= 1)
    cand = fmax(cand, __shfl_xor(cand, off, 64));
  __shared__ double wmax[kThreads / 64];
  const int         lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
    wmax[wave] = cand;
  if (__any(nan) && lane == 0)
    atomicOr(red + 1, 2ull);
  __syncthreads();
  if (threadIdx.x == 0)
  {
    double mx = wmax[0];
    for (int w = 1; w < kThreads / 64; ++w)
      mx = fmax(mx, wmax[w]);
    atomicMax(red, (unsigned long long)__double_as_longlong(mx));
  }
}

__global__ void __launch_bounds__(kThreads)
k_inv_dt(DevMesh m, Gas<double> g, double scaling, const double * __restrict__ U, const double * __restrict__ B,
         unsigned long long * red)
{
  const Idx  I{ m.dim, m.b };
  const int  b = m.b, dim = m.dim;
  const long ncell = I.cells();
  const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
  double     cand = 0.0;
  bool       nan = false;
  if (t < long(m.nowned) * ncell)
  {
    const int  o = int(t / ncell);
    const long ci = t - long(o) * ncell;
    const int  i = int(ci % b), j = int((ci / b) % b), k = dim == 3 ? int(ci / (long(b) * b)) : 0;
    cand = cell_inv_dt(m, I, g, scaling, U, B, o, ci, i, j, k);
    if (cand != cand)
    {
      nan = true;
      cand = 0.0;
    }
  }
  reduce_inv_dt(cand, nan, red);
}

// The same with the block size fixed at compile time: a bounded grid of workgroups, each
// folding many cells before its one atomic, numbered XCD by XCD.
template <int DIM, int BS>
__global__ void __launch_bounds__(kThreads)
k_inv_dt_c(DevMesh m, Gas<double> g, double scaling, const double * __restrict__ U, const double * __restrict__ B,
           unsigned long long * red)
{
  using IX = IdxC<DIM, BS>;
  constexpr IX  I{};
  constexpr int ncell = IX::cells();
  const long    total = long(m.nowned) * ncell;
  double        cand = 0.0;
  bool          nan = false;
  for (long t = xcd_group() * kThreads + threadIdx.x; t < total; t += long(gridDim.x) * kThreads)
  {
    const int o = int(t / ncell);
    const int ci = int(t - long(o) * ncell);
    const int i = ci % BS, j = (ci / BS) % BS, k = DIM == 3 ? ci / (BS * BS) : 0;
    const double c = cell_inv_dt(m, I, g, scaling, U, B, o, ci, i, j, k);
    if (c != c)
      nan = true;
    else
      cand = fmax(cand, c);
  }
  reduce_inv_dt(cand, nan, red);
}

// ---------------------------------------------------------------------------
// K16: outside blocks (FillOutsideCellFunctor, border_conditions/FillOutside.cpp:97-395)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads)
k_fill_outside_hydro(DevMesh m, double * __restrict__ U)
{
  const Idx  I{ m.dim, m.b };
  const int  b = m.b, dim = m.dim;
  const long ncell = I.cells();
  const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= long(m.noutside) * ncell)
    return;
  const int  oo = int(t / ncell);
  const long ci = t - long(oo) * ncell;
  const int  o = m.nowned + oo;
  const int  src = m.out_src[oo];
  const int * od = m.out_dirs + 3 * oo;
  const int  cc[3] = { int(ci % b), int((ci / b) % b), dim == 3 ? int(ci / (long(b) * b)) : 0 };
  for (int face = 0; face < 2 * dim; ++face)
  {
    const int d = face / 2;
    const int want = (face % 2 == 0) ? -1 : 1;
    if (od[d] != want)
      continue;
    const int bct = m.bc[face];
    int       cin[3] = { cc[0], cc[1], cc[2] };
    if (bct == 2)
      cin[d] = (face % 2 == 0) ? 0 : b - 1;
    else if (bct == 3)
      cin[d] = b - 1 - cc[d];
    else
      continue;
    const long ii = I.cell(cin[0], cin[1], cin[2]);
    for (int v = 0; v < 5; ++v)
      U[(long(o) * 5 + v) * ncell + ci] = U[(long(src) * 5 + v) * ncell + ii];
    if (bct == 3)
      U[(long(o) * 5 + kU + d) * ncell + ci] *= -1;
  }
}

__global__ void __launch_bounds__(kThreads)
k_fill_outside_mag(DevMesh m, double * __restrict__ B)
{
  const Idx  I{ m.dim, m.b };
  const int  b = m.b, dim = m.dim;
  const long per = I.fper();
  const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= long(m.noutside) * per)
    return;
  const int  oo = int(t / per);
  long       r = t - long(oo) * per;
  const int  o = m.nowned + oo;
  const int  src = m.out_src[oo];
  const int * od = m.out_dirs + 3 * oo;
  int        c = 0;
  while (c < 2 && r >= I.fsize(c))
  {
    r -= I.fsize(c);
    ++c;
  }
  const int nx = c == 0 ? b + 1 : b;
  const int ny = c == 1 ? b + 1 : b;
  int       f[3] = { int(r % nx), int((r / nx) % ny), dim == 3 ? int(r / (long(nx) * ny)) : 0 };
  for (int d = 0; d < dim; ++d)
  {
    if (od[d] == 0)
      continue;
    const int face = 2 * d + (od[d] < 0 ? 0 : 1);
    if (m.bc[face] != 2)
      continue;
    const int ext = (d == c && !(dim == 2 && c == 2)) ? b + 1 : b;
    f[d] = od[d] < 0 ? 0 : ext - 1;
  }
  B[long(o) * per + (t - long(oo) * per)] = B[long(src) * per + I.face(c, f[0], f[1], f[2])];
}

} // namespace

Gas<double>
gas_of(const Settings & s)
{
  Gas<double> g;
  g.gamma0 = s.gamma0;
  g.smallr = s.smallr;
  g.smallc = s.smallc;
  g.smallp = s.smallc * s.smallc / s.gamma0;
  g.slope_type = s.slope_type;
  g.grav = s.gravity_enabled;
  for (int d = 0; d < 3; ++d)
    g.g[d] = s.gravity[d];
  return g;
}

// ---------------------------------------------------------------------------
// Uniform gravity source (AddGravitySourceTerm.cpp:42-100): momentum += dt g (rho_old +
// rho_new) / 2, total energy += the kinetic energy change, on the owned cells of U2.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads)
k_gravity_source(DevMesh m, Gas<double> g, double dt, const double * __restrict__ U, double * __restrict__ U2)
{
  const Idx  I{ m.dim, m.b };
  const long ncell = I.cells();
  const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= long(m.nowned) * ncell)
    return;
  const int  o = int(t / ncell);
  const long ci = t - long(o) * ncell;
  gravity_cell(m.dim, g, dt, U + long(o) * 5 * ncell + ci, U2 + long(o) * 5 * ncell + ci, ncell);
}

#define MHDCT_HIP(x)                                                                               \
  do                                                                                               \
  {                                                                                                \
    hipError_t e_ = (x);                                                                           \
    if (e_ != hipSuccess)                                                                          \
    {                                                                                              \
      c.err = std::string(#x) + ": " + hipGetErrorString(e_);                                      \
      return -1;                                                                                   \
    }                                                                                              \
  } while (0)

int
plain_alloc(Context & c)
{
  plain_free(c);
  const Shapes S{ c.dim, c.b };
  const long   nq = c.dm.nowned;
  const long   ntot = c.dm.ntotal;
  struct A
  {
    double ** p;
    long      n;
  } arrs[] = {
    { &c.ps.Q, nq * 11 * S.gcells(2) },
    { &c.ps.Q2, nq * 8 * S.gcells(1) },
    { &c.ps.S[0], nq * 8 * S.gcells(1) },
    { &c.ps.S[1], nq * 8 * S.gcells(1) },
    { &c.ps.S[2], c.dim == 3 ? nq * 8 * S.gcells(1) : 8 },
    { &c.ps.Fl, nq * 8 * S.flux_cells() },
    { &c.ps.E, nq * S.ncomp_e() * S.gcells(2) },
    { &c.ps.sF, nq * (c.dim == 3 ? 3 : 2) * S.gcells(1) },
    { &c.ps.emf, nq * S.ncomp_e() * S.edge_cells() },
    { &c.ps.emf2, nq * S.ncomp_e() * S.edge_cells() },
    { &c.ps.Bg, ntot * S.gfper() },
  };
  c.ps.bytes = 0;
  for (auto & a : arrs)
  {
    const size_t bytes = size_t(a.n > 0 ? a.n : 1) * sizeof(double);
    MHDCT_HIP(hipMalloc(a.p, bytes));
    MHDCT_HIP(hipMemset(*a.p, 0, bytes));
    c.ps.bytes += bytes;
  }
  return 0;
}

void
plain_free(Context & c)
{
  double ** ps[] = { &c.ps.Q, &c.ps.Q2, &c.ps.S[0], &c.ps.S[1], &c.ps.S[2], &c.ps.Fl,
                     &c.ps.E, &c.ps.sF, &c.ps.emf, &c.ps.emf2, &c.ps.Bg };
  for (auto p : ps)
  {
    if (*p)
      (void)hipFree(*p);
    *p = nullptr;
  }
  c.ps.bytes = 0;
}

void
launch_viscous(Context & c, const double * Qp, int qvars, double * Fl, double * U2, double * B2, double dt,
               hipStream_t s)
{
  const DevMesh & m = c.dm;
  const Shapes    S{ c.dim, c.b };
  const long      nq = m.nowned;
  if (nq == 0)
    return;
  for (int dir = 0; dir < c.dim; ++dir)
  {
    hipLaunchKernelGGL(k_viscous_fluxes, grid_of(nq * S.flux_cells()), kThreads, 0, s, m, c.st.scaling, dt, c.st.mu,
                       dir, Qp, qvars, Fl);
    hipLaunchKernelGGL(k_update, grid_of(nq * S.cells()), kThreads, 0, s, m, dir, Fl, U2, B2);
  }
}

int
plain_step(Context & c, const double * U, double * U2, const double * B, double * B2, double dt, hipStream_t s)
{
  const DevMesh & m = c.dm;
  const Shapes    S{ c.dim, c.b };
  const Gas<double> g = gas_of(c.st);
  const long      nq = m.nowned;
  if (nq == 0)
    return 0;
  const double sc = c.st.scaling;
  hipLaunchKernelGGL(k_fill_bface_ghosted, grid_of(nq * S.gfper()), kThreads, 0, s, m, B, c.ps.Bg);
  hipLaunchKernelGGL(k_convert_to_primitives, grid_of(nq * S.gcells(2)), kThreads, 0, s, m, g,
                     c.st.abort_when_negative_eint, U, B, c.ps.Bg, c.ps.Q, c.red + 1);
  hipLaunchKernelGGL(k_slopes, grid_of(nq * S.gcells(1)), kThreads, 0, s, m, g.slope_type, c.ps.Q, c.ps.S[0],
                     c.ps.S[1], c.ps.S[2]);
  const int rk2 = c.st.time_integrator == 1;
  const double visc_pred = (c.st.viscosity_enabled && c.st.viscosity_predictor) ? c.st.mu : -1.0;
  hipLaunchKernelGGL(k_predictor, grid_of(nq * S.gcells(1)), kThreads, 0, s, m, sc, g, dt, rk2, visc_pred, c.ps.Q,
                     c.ps.S[0], c.ps.S[1], c.ps.S[2], c.ps.Q2);
  hipLaunchKernelGGL(k_elec_field, grid_of(nq * S.gcells(2)), kThreads, 0, s, m, c.ps.Q, c.ps.E);
  hipLaunchKernelGGL(k_sfacemag, grid_of(nq * S.gcells(1)), kThreads, 0, s, m, sc, dt, c.ps.E, c.ps.sF);
  for (int dir = 0; dir < c.dim; ++dir)
  {
    hipLaunchKernelGGL(k_fluxes, grid_of(nq * S.flux_cells()), kThreads, 0, s, m, g, sc, dt, dir, c.ps.Q,
                       c.ps.Q2, c.ps.S[dir], c.ps.sF, c.ps.Fl);
    hipLaunchKernelGGL(k_update, grid_of(nq * S.cells()), kThreads, 0, s, m, dir, c.ps.Fl, U2, B2);
  }
  if (c.st.viscosity_enabled) // GodunovImplemV0.cpp:179-192
    launch_viscous(c, c.ps.Q, 11, c.ps.Fl, U2, B2, dt, s);
  hipLaunchKernelGGL(k_emf, grid_of(nq * S.edge_cells()), kThreads, 0, s, m, g, sc, dt, c.ps.Q, c.ps.Q2,
                     c.ps.S[0], c.ps.S[1], c.ps.S[2], c.ps.sF, rk2 ? 0 : 1, c.ps.emf);
  MHDCT_HIP(hipMemcpyAsync(c.ps.emf2, c.ps.emf, size_t(nq) * S.ncomp_e() * S.edge_cells() * sizeof(double),
                           hipMemcpyDeviceToDevice, s));
  if (m.nfix > 0)
    hipLaunchKernelGGL(k_emf_fix, grid_of(m.nfix), kThreads, 0, s, m.nfix, m.fix, m.fix_src, c.ps.emf, c.ps.emf2);
  hipLaunchKernelGGL(k_emf_update, grid_of(nq * S.fper()), kThreads, 0, s, m, c.ps.emf2, B2);
  if (g.grav)
    hipLaunchKernelGGL(k_gravity_source, grid_of(nq * S.cells()), kThreads, 0, s, m, g, dt, U, U2);
  MHDCT_HIP(hipGetLastError());
  return 0;
}

int
launch_inv_dt(Context & c, const double * U, const double * B, hipStream_t s)
{
  const Shapes S{ c.dim, c.b };
  const long   nq = c.dm.nowned;
  if (nq == 0)
    return 0;
  if (c.impl != 10 && (c.b == 8 || c.b == 16 || c.b == 32))
  {
    // the fused forms' reduction: compile-time geometry, bounded grid (a multiple of 8)
    const long groups = (nq * S.cells() + kThreads - 1) / kThreads;
    const unsigned grid = unsigned(std::min<long>((groups + 7) / 8 * 8, 2048));
    const Gas<double> g = gas_of(c.st);
    switch (c.dim * 100 + c.b)
    {
    case 208: hipLaunchKernelGGL((k_inv_dt_c<2, 8>), grid, kThreads, 0, s, c.dm, g, c.st.scaling, U, B, c.red); break;
    case 216: hipLaunchKernelGGL((k_inv_dt_c<2, 16>), grid, kThreads, 0, s, c.dm, g, c.st.scaling, U, B, c.red); break;
    case 232: hipLaunchKernelGGL((k_inv_dt_c<2, 32>), grid, kThreads, 0, s, c.dm, g, c.st.scaling, U, B, c.red); break;
    case 308: hipLaunchKernelGGL((k_inv_dt_c<3, 8>), grid, kThreads, 0, s, c.dm, g, c.st.scaling, U, B, c.red); break;
    case 316: hipLaunchKernelGGL((k_inv_dt_c<3, 16>), grid, kThreads, 0, s, c.dm, g, c.st.scaling, U, B, c.red); break;
    default: hipLaunchKernelGGL((k_inv_dt_c<3, 32>), grid, kThreads, 0, s, c.dm, g, c.st.scaling, U, B, c.red); break;
    }
  }
  else
    hipLaunchKernelGGL(k_inv_dt, grid_of(nq * S.cells()), kThreads, 0, s, c.dm, gas_of(c.st), c.st.scaling, U, B,
                       c.red);
  MHDCT_HIP(hipGetLastError());
  return 0;
}


// ---------------------------------------------------------------------------
// RK2-SSP stage combination: y = c0 * x + c1 * y over n values (SolverGodunovMHD.cpp:583-588,
// 617-622 call kalypsso::core's LinearCombination on every block of U and of Bface).  Two
// products and one sum, in this order, each one IEEE operation.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads)
k_linear_combination(long n, double c0, double c1, const double * __restrict__ x, double * __restrict__ y)
{
  for (long t = long(blockIdx.x) * blockDim.x + threadIdx.x; t < n; t += long(gridDim.x) * blockDim.x)
    y[t] = c0 * x[t] + c1 * y[t];
}

int
launch_linear_combination(Context & c, const double * U, double * U_RK, const double * B, double * B_RK, double c0,
                          double c1, hipStream_t s)
{
  const Shapes S{ c.dim, c.b };
  const long   nt = c.dm.ntotal;
  if (nt == 0)
    return 0;
  const long nu = nt * 5 * S.cells(), nb = nt * S.fper();
  const long cap = 8192; // a grid-stride loop over at most 8192 workgroups
  hipLaunchKernelGGL(k_linear_combination, dim3(unsigned(std::min(cap, (nu + kThreads - 1) / kThreads))), kThreads,
                     0, s, nu, c0, c1, U, U_RK);
  hipLaunchKernelGGL(k_linear_combination, dim3(unsigned(std::min(cap, (nb + kThreads - 1) / kThreads))), kThreads,
                     0, s, nb, c0, c1, B, B_RK);
  MHDCT_HIP(hipGetLastError());
  return 0;
}

int
launch_fill_outside(Context & c, double * U, double * B, hipStream_t s)
{
  const Shapes S{ c.dim, c.b };
  if (c.dm.noutside == 0)
    return 0;
  hipLaunchKernelGGL(k_fill_outside_hydro, grid_of(long(c.dm.noutside) * S.cells()), kThreads, 0, s, c.dm, U);
  hipLaunchKernelGGL(k_fill_outside_mag, grid_of(long(c.dm.noutside) * S.fper()), kThreads, 0, s, c.dm, B);
  MHDCT_HIP(hipGetLastError());
  return 0;
}

} // namespace mhdct
