// phys.hpp -- ideal-MHD point functions used by every kernel form.
//
// Same arithmetic, operation for operation, as the reference model functions so that the
// FP64 path reproduces its rounding (built with -ffp-contract=off):
//   primitives      models/mhd_utils.h:449-538   (compute_primitives)
//   fast speed      models/mhd_utils.h:93-122    (find_speed_fast)
//   info speed      models/mhd_utils.h:340-394   (find_speed_info)
//   slope limiter   scheme/ComputeLimitedSlopesFunctor.cpp:121-144
//   HLLD            models/RiemannSolvers_MHD.h:311-551
//   2D HLLD EMF     models/RiemannSolvers_MHD.h:629-869, 890-1097
// T is the arithmetic type of the bulk (double, or float for the mixed-precision form).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace mhdct
{

// type-preserving overloads for both double and float bulk types
using std::copysign;
using std::fabs;
using std::fmax;
using std::fmin;
using std::sqrt;

// cell state component order (MHD::VarId, models/MHD.h:35-62)
enum : int
{
  kD = 0,
  kP = 1,
  kU = 2,
  kV = 3,
  kW = 4,
  kA = 5,
  kB = 6,
  kC = 7
};

template <typename T>
struct Gas
{
  T   gamma0;
  T   smallr;
  T   smallc;
  T   smallp;
  T   slope_type;
  int grav;  // [gravity] enabled: uniform gravity field g
  T   g[3];  // [gravity] uniform_gx, uniform_gy, uniform_gz
};

// Uniform gravity source of one cell (AddGravitySourceTerm.cpp:42-100); u and u2 point at the
// cell's density in U and U2, variables `stride` apart.
__host__ __device__ inline void
gravity_cell(int dim, const Gas<double> & g, double dt, const double * u, double * u2, long stride)
{
  const double rho_old = u[0];
  const double rho_new = u2[0];
  double       rhou = u2[2 * stride], rhov = u2[3 * stride], rhow = 0.0;
  if (dim == 3)
    rhow = u2[4 * stride];
  double ekin_old = 0.5 * (rhou * rhou + rhov * rhov) / rho_new;
  if (dim == 3)
    ekin_old += 0.5 * (rhow * rhow) / rho_new;
  rhou += dt * g.g[0] * (rho_old + rho_new) / 2;
  rhov += dt * g.g[1] * (rho_old + rho_new) / 2;
  if (dim == 3)
    rhow += dt * g.g[2] * (rho_old + rho_new) / 2;
  u2[2 * stride] = rhou;
  u2[3 * stride] = rhov;
  if (dim == 3)
    u2[4 * stride] = rhow;
  double ekin_new = 0.5 * (rhou * rhou + rhov * rhov) / rho_new;
  if (dim == 3)
    ekin_new += 0.5 * (rhow * rhow) / rho_new;
  u2[1 * stride] += (ekin_new - ekin_old);
}

// The CFL rate of a cell from its fastest signal speeds v[d] = c_f,d + |u_d|
// (ComputeDtMHDFunctor.cpp:86-151), or, with gravity, the rate that also keeps
// g dt^2 / (2 dx) + |v| dt / dx <= cfl: each v[d] / dx scaled by k / (sqrt(1 + 2k) - 1),
// k = max(sum_d |g_d| dx / |v|^2, 1e-4) (compute_cfl_with_gravity, :108-249).
template <int DIM>
__host__ __device__ inline double
cell_rate(const double * v, double dx, const Gas<double> & g)
{
  if (!g.grav)
    return DIM == 2 ? v[0] / dx + v[1] / dx : v[0] / dx + v[1] / dx + v[2] / dx;
  double v2 = v[0] * v[0] + v[1] * v[1];
  if (DIM == 3)
    v2 += v[2] * v[2];
  double k = fabs(g.g[0]) + fabs(g.g[1]);
  if (DIM == 3)
    k += fabs(g.g[2]);
  k *= dx / v2;
  k = fmax(k, 1e-4);
  const double factor = k / (sqrt(1.0 + 2.0 * k) - 1.0);
  double       r = 0.0;
  r += (v[0] * factor) / dx;
  r += (v[1] * factor) / dx;
  if (DIM == 3)
    r += (v[2] * factor) / dx;
  return r;
}

template <typename T>
__host__ __device__ inline T
fast_speed(const T * q, int dir, const Gas<T> & g)
{
  const T d = q[kD], p = q[kP], a = q[kA], b = q[kB], c = q[kC];
  const T b2 = a * a + b * b + c * c;
  const T c2 = g.gamma0 * p / d;
  const T d2 = T(0.5) * (b2 / d + c2);
  const T bn = dir == 0 ? a : (dir == 1 ? b : c);
  return sqrt(d2 + sqrt(d2 * d2 - c2 * bn * bn / d));
}

// conservative (rho, E, mx, my, mz, Bx, By, Bz) -> primitive (rho, p, u, v, w, Bx, By, Bz).
// Returns the internal energy so a caller can flag eint < 0.
template <typename T>
__host__ __device__ inline T
to_primitive(const T * u, T * q, const Gas<T> & g)
{
  q[kD] = fmax(u[kD], g.smallr);
  q[kU] = u[kU] / q[kD];
  q[kV] = u[kV] / q[kD];
  q[kW] = u[kW] / q[kD];
  q[kA] = u[kA];
  q[kB] = u[kB];
  q[kC] = u[kC];
  const T eken = T(0.5) * (q[kU] * q[kU] + q[kV] * q[kV] + q[kW] * q[kW]);
  const T emag = T(0.5) * (q[kA] * q[kA] + q[kB] * q[kB] + q[kC] * q[kC]);
  const T eint = (u[kP] - emag) / q[kD] - eken;
  q[kP] = fmax((g.gamma0 - T(1)) * q[kD] * eint, q[kD] * g.smallp);
  return eint;
}

template <typename T>
__host__ __device__ inline T
limited_slope(T qc, T qp, T qm, T slope_type)
{
  T dq = 0;
  if (slope_type == T(1) || slope_type == T(2))
  {
    const T dl = slope_type * (qc - qm);
    const T dr = slope_type * (qp - qc);
    const T dc = T(0.5) * (qp - qm);
    const T sg = (dc >= T(0)) ? T(1) : T(-1);
    const T mn = fmin(fabs(dl), fabs(dr));
    const T lim = (dl * dr) <= T(0) ? T(0) : mn;
    dq = sg * fmin(lim, fabs(dc));
  }
  return dq;
}

// HLLD flux along x of the (already rotated) left/right states; ql[kA], qr[kA] are
// overwritten with the continuous normal field like the reference does.
template <typename T>
__host__ __device__ inline void
hlld(T * ql, T * qr, T * f, const Gas<T> & g)
{
  const T entho = T(1) / (g.gamma0 - T(1));
  const T a = T(0.5) * (ql[kA] + qr[kA]);
  const T sgnm = (a >= T(0)) ? T(1) : T(-1);
  ql[kA] = a;
  qr[kA] = a;

  const T rl = ql[kD], pl = ql[kP], ul = ql[kU], vl = ql[kV], wl = ql[kW], bl = ql[kB], cl = ql[kC];
  const T ecinl = T(0.5) * (ul * ul + vl * vl + wl * wl) * rl;
  const T emagl = T(0.5) * (a * a + bl * bl + cl * cl);
  const T etotl = pl * entho + ecinl + emagl;
  const T ptotl = pl + emagl;
  const T vdotbl = ul * a + vl * bl + wl * cl;

  const T rr = qr[kD], pr = qr[kP], ur = qr[kU], vr = qr[kV], wr = qr[kW], br = qr[kB], cr = qr[kC];
  const T ecinr = T(0.5) * (ur * ur + vr * vr + wr * wr) * rr;
  const T emagr = T(0.5) * (a * a + br * br + cr * cr);
  const T etotr = pr * entho + ecinr + emagr;
  const T ptotr = pr + emagr;
  const T vdotbr = ur * a + vr * br + wr * cr;

  const T cfl_ = fast_speed(ql, 0, g);
  const T cfr_ = fast_speed(qr, 0, g);
  const T sl = fmin(ul, ur) - fmax(cfl_, cfr_);
  const T sr = fmax(ul, ur) + fmax(cfl_, cfr_);

  const T rcl = rl * (ul - sl);
  const T rcr = rr * (sr - ur);
  const T ustar = (rcr * ur + rcl * ul + (ptotl - ptotr)) / (rcr + rcl);
  const T ptotstar = (rcr * ptotl + rcl * ptotr + rcl * rcr * (ul - ur)) / (rcr + rcl);

  // left star region
  const T rstarl = rl * (sl - ul) / (sl - ustar);
  T       estar = rl * (sl - ul) * (sl - ustar) - a * a;
  const T el = rl * (sl - ul) * (sl - ul) - a * a;
  T       vstarl, wstarl, bstarl, cstarl;
  if (fabs(estar) < T(1e-4) * (a * a))
  {
    vstarl = vl;
    bstarl = bl;
    wstarl = wl;
    cstarl = cl;
  }
  else
  {
    vstarl = vl - a * bl * (ustar - ul) / estar;
    bstarl = bl * el / estar;
    wstarl = wl - a * cl * (ustar - ul) / estar;
    cstarl = cl * el / estar;
  }
  const T vdotbstarl = ustar * a + vstarl * bstarl + wstarl * cstarl;
  const T etotstarl =
    ((sl - ul) * etotl - ptotl * ul + ptotstar * ustar + a * (vdotbl - vdotbstarl)) / (sl - ustar);
  const T sqrrstarl = sqrt(rstarl);
  const T sal = ustar - fabs(a) / sqrrstarl;

  // right star region
  const T rstarr = rr * (sr - ur) / (sr - ustar);
  estar = rr * (sr - ur) * (sr - ustar) - a * a;
  const T er = rr * (sr - ur) * (sr - ur) - a * a;
  T       vstarr, wstarr, bstarr, cstarr;
  if (fabs(estar) < T(1e-4) * (a * a))
  {
    vstarr = vr;
    bstarr = br;
    wstarr = wr;
    cstarr = cr;
  }
  else
  {
    vstarr = vr - a * br * (ustar - ur) / estar;
    bstarr = br * er / estar;
    wstarr = wr - a * cr * (ustar - ur) / estar;
    cstarr = cr * er / estar;
  }
  const T vdotbstarr = ustar * a + vstarr * bstarr + wstarr * cstarr;
  const T etotstarr =
    ((sr - ur) * etotr - ptotr * ur + ptotstar * ustar + a * (vdotbr - vdotbstarr)) / (sr - ustar);
  const T sqrrstarr = sqrt(rstarr);
  const T sar = ustar + fabs(a) / sqrrstarr;

  // double-star region
  const T den = sqrrstarl + sqrrstarr;
  const T vss = (sqrrstarl * vstarl + sqrrstarr * vstarr + sgnm * (bstarr - bstarl)) / den;
  const T wss = (sqrrstarl * wstarl + sqrrstarr * wstarr + sgnm * (cstarr - cstarl)) / den;
  const T bss = (sqrrstarl * bstarr + sqrrstarr * bstarl + sgnm * sqrrstarl * sqrrstarr * (vstarr - vstarl)) / den;
  const T css = (sqrrstarl * cstarr + sqrrstarr * cstarl + sgnm * sqrrstarl * sqrrstarr * (wstarr - wstarl)) / den;
  const T vdotbss = ustar * a + vss * bss + wss * css;

  T ro, uo, vo, wo, bo, co, ptoto, etoto, vdotbo;
  if (sl > T(0))
  {
    ro = rl; uo = ul; vo = vl; wo = wl; bo = bl; co = cl; ptoto = ptotl; etoto = etotl; vdotbo = vdotbl;
  }
  else if (sal > T(0))
  {
    ro = rstarl; uo = ustar; vo = vstarl; wo = wstarl; bo = bstarl; co = cstarl;
    ptoto = ptotstar; etoto = etotstarl; vdotbo = vdotbstarl;
  }
  else if (ustar > T(0))
  {
    ro = rstarl; uo = ustar; vo = vss; wo = wss; bo = bss; co = css;
    ptoto = ptotstar; etoto = etotstarl - sgnm * sqrrstarl * (vdotbstarl - vdotbss); vdotbo = vdotbss;
  }
  else if (sar > T(0))
  {
    ro = rstarr; uo = ustar; vo = vss; wo = wss; bo = bss; co = css;
    ptoto = ptotstar; etoto = etotstarr + sgnm * sqrrstarr * (vdotbstarr - vdotbss); vdotbo = vdotbss;
  }
  else if (sr > T(0))
  {
    ro = rstarr; uo = ustar; vo = vstarr; wo = wstarr; bo = bstarr; co = cstarr;
    ptoto = ptotstar; etoto = etotstarr; vdotbo = vdotbstarr;
  }
  else
  {
    ro = rr; uo = ur; vo = vr; wo = wr; bo = br; co = cr; ptoto = ptotr; etoto = etotr; vdotbo = vdotbr;
  }
  f[kD] = ro * uo;
  f[kP] = (etoto + ptoto) * uo - a * vdotbo;
  f[kU] = ro * uo * uo - a * a + ptoto;
  f[kV] = ro * uo * vo - a * bo;
  f[kW] = ro * uo * wo - a * co;
  f[kA] = T(0);
  f[kB] = bo * uo - a * vo;
  f[kC] = co * uo - a * wo;
}

template <typename T>
__host__ __device__ inline T
max4(T a0, T a1, T a2, T a3)
{
  T r = a0;
  r = (a1 > r) ? a1 : r;
  r = (a2 > r) ? a2 : r;
  r = (a3 > r) ? a3 : r;
  return r;
}
template <typename T>
__host__ __device__ inline T
min4(T a0, T a1, T a2, T a3)
{
  T r = a0;
  r = (a1 < r) ? a1 : r;
  r = (a2 < r) ? a2 : r;
  r = (a3 < r) ? a3 : r;
  return r;
}
template <typename T>
__host__ __device__ inline T
max5(T a0, T a1, T a2, T a3, T a4)
{
  T r = a0;
  r = (a1 > r) ? a1 : r;
  r = (a2 > r) ? a2 : r;
  r = (a3 > r) ? a3 : r;
  r = (a4 > r) ? a4 : r;
  return r;
}

// Edge state slots of the four reconstructed corner states.
enum : int
{
  kRT = 0,
  kRB = 1,
  kLT = 2,
  kLB = 3
};

// 2D HLLD electric field at an edge from the four corner states (RT, RB, LT, LB) given in
// the lab frame; `axis` is the edge direction (0 = x, 1 = y, 2 = z).
template <int AXIS, typename T>
__host__ __device__ inline T
edge_emf_ax(const T (&qe)[4][8], const Gas<T> & g)
{
  // in-plane (iu, iv), out-of-plane iw, and the same for B
  constexpr int iu = AXIS == 2 ? kU : (AXIS == 1 ? kW : kV);
  constexpr int iv = AXIS == 2 ? kV : (AXIS == 1 ? kU : kW);
  constexpr int iw = AXIS == 2 ? kW : (AXIS == 1 ? kV : kU);
  constexpr int ia = AXIS == 2 ? kA : (AXIS == 1 ? kC : kB);
  constexpr int ib = AXIS == 2 ? kB : (AXIS == 1 ? kA : kC);
  constexpr int ic = AXIS == 2 ? kC : (AXIS == 1 ? kB : kA);
  const T * RT = qe[kRT];
  const T * LT = qe[kLT];
  const T * RB = qe[kRB];
  const T * LB = qe[kLB];
  // the LL, RL, LR, RR states of the magnetic Riemann problem
  T s[4][8];
  const T *src[4] = { RT, LT, RB, LB };
  for (int n = 0; n < 4; ++n)
  {
    s[n][kD] = src[n][kD];
    s[n][kP] = src[n][kP];
    s[n][kU] = src[n][iu];
    s[n][kV] = src[n][iv];
    s[n][kW] = src[n][iw];
    s[n][kC] = src[n][ic];
  }
  enum
  {
    LL = 0,
    RL = 1,
    LR = 2,
    RR = 3
  };
  s[LL][kA] = T(0.5) * (RT[ia] + LT[ia]);
  s[RL][kA] = T(0.5) * (RT[ia] + LT[ia]);
  s[LR][kA] = T(0.5) * (RB[ia] + LB[ia]);
  s[RR][kA] = T(0.5) * (RB[ia] + LB[ia]);
  s[LL][kB] = T(0.5) * (RT[ib] + RB[ib]);
  s[RL][kB] = T(0.5) * (LT[ib] + LB[ib]);
  s[LR][kB] = T(0.5) * (RT[ib] + RB[ib]);
  s[RR][kB] = T(0.5) * (LT[ib] + LB[ib]);

  const T ELL = s[LL][kU] * s[LL][kB] - s[LL][kV] * s[LL][kA];
  const T ERL = s[RL][kU] * s[RL][kB] - s[RL][kV] * s[RL][kA];
  const T ELR = s[LR][kU] * s[LR][kB] - s[LR][kV] * s[LR][kA];
  const T ERR = s[RR][kU] * s[RR][kB] - s[RR][kV] * s[RR][kA];

  const T rLL = s[LL][kD], pLL = s[LL][kP], uLL = s[LL][kU], vLL = s[LL][kV];
  const T aLL = s[LL][kA], bLL = s[LL][kB], cLL = s[LL][kC];
  const T rLR = s[LR][kD], pLR = s[LR][kP], uLR = s[LR][kU], vLR = s[LR][kV];
  const T aLR = s[LR][kA], bLR = s[LR][kB], cLR = s[LR][kC];
  const T rRL = s[RL][kD], pRL = s[RL][kP], uRL = s[RL][kU], vRL = s[RL][kV];
  const T aRL = s[RL][kA], bRL = s[RL][kB], cRL = s[RL][kC];
  const T rRR = s[RR][kD], pRR = s[RR][kP], uRR = s[RR][kU], vRR = s[RR][kV];
  const T aRR = s[RR][kA], bRR = s[RR][kB], cRR = s[RR][kC];

  const T cLLx = fast_speed(s[LL], 0, g), cLRx = fast_speed(s[LR], 0, g);
  const T cRLx = fast_speed(s[RL], 0, g), cRRx = fast_speed(s[RR], 0, g);
  const T cLLy = fast_speed(s[LL], 1, g), cLRy = fast_speed(s[LR], 1, g);
  const T cRLy = fast_speed(s[RL], 1, g), cRRy = fast_speed(s[RR], 1, g);

  const T SL = min4(uLL, uLR, uRL, uRR) - max4(cLLx, cLRx, cRLx, cRRx);
  const T SR = max4(uLL, uLR, uRL, uRR) + max4(cLLx, cLRx, cRLx, cRRx);
  const T SB = min4(vLL, vLR, vRL, vRR) - max4(cLLy, cLRy, cRLy, cRRy);
  const T ST = max4(vLL, vLR, vRL, vRR) + max4(cLLy, cLRy, cRLy, cRRy);

  const T PtLL = pLL + T(0.5) * (aLL * aLL + bLL * bLL + cLL * cLL);
  const T PtLR = pLR + T(0.5) * (aLR * aLR + bLR * bLR + cLR * cLR);
  const T PtRL = pRL + T(0.5) * (aRL * aRL + bRL * bRL + cRL * cRL);
  const T PtRR = pRR + T(0.5) * (aRR * aRR + bRR * bRR + cRR * cRR);

  const T rcLLx = rLL * (uLL - SL), rcRLx = rRL * (SR - uRL);
  const T rcLRx = rLR * (uLR - SL), rcRRx = rRR * (SR - uRR);
  const T rcLLy = rLL * (vLL - SB), rcLRy = rLR * (ST - vLR);
  const T rcRLy = rRL * (vRL - SB), rcRRy = rRR * (ST - vRR);

  const T ustar = (rcLLx * uLL + rcLRx * uLR + rcRLx * uRL + rcRRx * uRR + (PtLL - PtRL + PtLR - PtRR)) /
                  (rcLLx + rcLRx + rcRLx + rcRRx);
  const T vstar = (rcLLy * vLL + rcLRy * vLR + rcRLy * vRL + rcRRy * vRR + (PtLL - PtLR + PtRL - PtRR)) /
                  (rcLLy + rcLRy + rcRLy + rcRRy);

  const T rsLLx = rLL * (SL - uLL) / (SL - ustar);
  const T BsLL = bLL * (SL - uLL) / (SL - ustar);
  const T rsLLy = rLL * (SB - vLL) / (SB - vstar);
  const T AsLL = aLL * (SB - vLL) / (SB - vstar);
  const T rsLL = rLL * (SL - uLL) / (SL - ustar) * (SB - vLL) / (SB - vstar);
  const T EsLLx = ustar * BsLL - vLL * aLL;
  const T EsLLy = uLL * bLL - vstar * AsLL;
  const T EsLL = ustar * BsLL - vstar * AsLL;

  const T rsLRx = rLR * (SL - uLR) / (SL - ustar);
  const T BsLR = bLR * (SL - uLR) / (SL - ustar);
  const T rsLRy = rLR * (ST - vLR) / (ST - vstar);
  const T AsLR = aLR * (ST - vLR) / (ST - vstar);
  const T rsLR = rLR * (SL - uLR) / (SL - ustar) * (ST - vLR) / (ST - vstar);
  const T EsLRx = ustar * BsLR - vLR * aLR;
  const T EsLRy = uLR * bLR - vstar * AsLR;
  const T EsLR = ustar * BsLR - vstar * AsLR;

  const T rsRLx = rRL * (SR - uRL) / (SR - ustar);
  const T BsRL = bRL * (SR - uRL) / (SR - ustar);
  const T rsRLy = rRL * (SB - vRL) / (SB - vstar);
  const T AsRL = aRL * (SB - vRL) / (SB - vstar);
  const T rsRL = rRL * (SR - uRL) / (SR - ustar) * (SB - vRL) / (SB - vstar);
  const T EsRLx = ustar * BsRL - vRL * aRL;
  const T EsRLy = uRL * bRL - vstar * AsRL;
  const T EsRL = ustar * BsRL - vstar * AsRL;

  const T rsRRx = rRR * (SR - uRR) / (SR - ustar);
  const T BsRR = bRR * (SR - uRR) / (SR - ustar);
  const T rsRRy = rRR * (ST - vRR) / (ST - vstar);
  const T AsRR = aRR * (ST - vRR) / (ST - vstar);
  const T rsRR = rRR * (SR - uRR) / (SR - ustar) * (ST - vRR) / (ST - vstar);
  const T EsRRx = ustar * BsRR - vRR * aRR;
  const T EsRRy = uRR * bRR - vstar * AsRR;
  const T EsRR = ustar * BsRR - vstar * AsRR;

  const T caL = max5(fabs(aLR) / sqrt(rsLRx), fabs(AsLR) / sqrt(rsLR), fabs(aLL) / sqrt(rsLLx),
                     fabs(AsLL) / sqrt(rsLL), g.smallc);
  const T caR = max5(fabs(aRR) / sqrt(rsRRx), fabs(AsRR) / sqrt(rsRR), fabs(aRL) / sqrt(rsRLx),
                     fabs(AsRL) / sqrt(rsRL), g.smallc);
  const T caB = max5(fabs(bLL) / sqrt(rsLLy), fabs(BsLL) / sqrt(rsLL), fabs(bRL) / sqrt(rsRLy),
                     fabs(BsRL) / sqrt(rsRL), g.smallc);
  const T caT = max5(fabs(bLR) / sqrt(rsLRy), fabs(BsLR) / sqrt(rsLR), fabs(bRR) / sqrt(rsRRy),
                     fabs(BsRR) / sqrt(rsRR), g.smallc);

  const T SAL = fmin(ustar - caL, T(0));
  const T SAR = fmax(ustar + caR, T(0));
  const T SAB = fmin(vstar - caB, T(0));
  const T SAT = fmax(vstar + caT, T(0));

  const T AsT = (SAR * AsRR - SAL * AsLR) / (SAR - SAL);
  const T AsB = (SAR * AsRL - SAL * AsLL) / (SAR - SAL);
  const T BsR = (SAT * BsRR - SAB * BsRL) / (SAT - SAB);
  const T BsL = (SAT * BsLR - SAB * BsLL) / (SAT - SAB);

  const T SBp = (T(1) + copysign(T(1), SB)) / T(2), SBn = T(1) - SBp;
  const T STp = (T(1) + copysign(T(1), ST)) / T(2), STn = T(1) - STp;
  const T SLp = (T(1) + copysign(T(1), SL)) / T(2), SLn = T(1) - SLp;
  const T SRp = (T(1) + copysign(T(1), SR)) / T(2), SRn = T(1) - SRp;

  T E = 0, t;
  t = (SAL * SAB * EsRR - SAL * SAT * EsRL - SAR * SAB * EsLR + SAR * SAT * EsLL) / (SAR - SAL) / (SAT - SAB) -
      SAT * SAB / (SAT - SAB) * (AsT - AsB) + SAR * SAL / (SAR - SAL) * (BsR - BsL);
  E += (SBn * STp * SLn * SRp) * t;

  t = (SAR * EsLLx - SAL * EsRLx + SAR * SAL * (bRL - bLL)) / (SAR - SAL);
  t = SLp * ELL + SLn * SRn * ERL + SLn * SRp * t;
  E += SBp * t;

  t = (SAR * EsLRx - SAL * EsRRx + SAR * SAL * (bRR - bLR)) / (SAR - SAL);
  t = SLp * ELR + SLn * SRn * ERR + SLn * SRp * t;
  E += (SBn * STn) * t;

  t = (SAT * EsLLy - SAB * EsLRy - SAT * SAB * (aLR - aLL)) / (SAT - SAB);
  E += (SBn * STp * SLp) * t;

  t = (SAT * EsRLy - SAB * EsRRy - SAT * SAB * (aRR - aRL)) / (SAT - SAB);
  E += (SBn * STp * SLn * SRn) * t;
  return E;
}

template <typename T>
__host__ __device__ inline T
edge_emf(const T (&qe)[4][8], int axis, const Gas<T> & g)
{
  return axis == 2 ? edge_emf_ax<2>(qe, g) : (axis == 1 ? edge_emf_ax<1>(qe, g) : edge_emf_ax<0>(qe, g));
}

} // namespace mhdct
