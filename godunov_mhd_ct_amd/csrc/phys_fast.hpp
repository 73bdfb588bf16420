// phys_fast.hpp -- the HLLD flux and the 2D-HLLD EMF of phys.hpp with their divisions and square
// roots shared: the same mathematics (models/RiemannSolvers_MHD.h:311-551, 629-869) with every
// repeated denominator inverted once and multiplied in, the Alfven-speed maxima taken over
// squares before one square root, and the largest of several fast speeds taken as the square
// root of the largest radicand (the square root is monotonic).  This cuts the 2D-HLLD EMF from 70
// divisions and 28 square roots to 20 reciprocals and 14 square roots, the HLLD flux from 27
// divisions and 6 square roots to 11 and 5.  In FP64 each reciprocal and square root is a
// Refinements of the hardware estimates (rcp_nr, sqrt_nr): 4 and 8 instructions, against the 11
// and 17 of the correctly rounded sequences.  The fused form uses these; its results differ from
// the reference's operation order by a few ulps per evaluation (its parity bar is the north
// star's 1e-12 relative L-inf per field, momentum as a vector).  The plain form keeps phys.hpp's
// order, bit for bit the reference's.
#pragma once

#include "phys.hpp"

namespace mhdct
{

// 1/x in FP64: the hardware estimate r = v_rcp_f64(x) (relative error below 2^-25 on MI355X)
// refined by one third-order step of the reciprocal iteration: with e = 1 - x r,
// 1/x = r / (1 - e) = r (1 + e + e^2 + ...), so r <- r + r (e + e^2) leaves a relative error of
// e^3 (below 2^-75) before the final rounding; three fused multiply-adds.  Equal to the correctly
// rounded 1/x on all of 4M random normal x over 2^-30 .. 2^31 on MI355X, though not guaranteed
// to be.  The FP32 form keeps 1/x (the hardware-approximate FP32 division of its build).
template <typename T>
__host__ __device__ __forceinline__ T
rcp_nr(T x)
{
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (sizeof(T) == 8)
  {
    const double r = __builtin_amdgcn_rcp(x);
    const double e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, __builtin_fma(e, e, e), r);
  }
  else
#endif
    return T(1) / x;
}

// sqrt(x) in FP64: the hardware estimate y = v_rsq_f64(x) of 1/sqrt(x), one step of Goldschmidt's
// coupled iteration for g = sqrt(x) and h = 1/(2 sqrt(x)) (r = 1/2 - g h; g <- g + g r;
// h <- h + h r; within a few tens of ulps), then one residual correction g <- g + (x - g^2) h.
// Equal to the correctly rounded sqrt(x) on all of 4M random normal x over 2^-30 .. 2^31 on
// MI355X, though not guaranteed to be; sqrt(+-0) = +-0, NaN below zero.
template <typename T>
__host__ __device__ __forceinline__ T
sqrt_nr(T x)
{
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (sizeof(T) == 8)
  {
    const double y = __builtin_amdgcn_rsq(x);
    double       g = x * y;
    double       h = 0.5 * y;
    const double r = __builtin_fma(-g, h, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    const double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    return x == 0.0 ? x : g;
  }
  else
#endif
    return sqrt(x);
}

// 1/a and 1/b (rcp2), 1/a .. 1/d (rcp4) from one reciprocal of the product: Montgomery's batch
// inversion, 1/a = b / (ab) and so on, 3 and 9 multiplies beside one rcp_nr.  In FP64 on the
// device only, for nonzero factors whose product stays a normal number (densities, wave-speed
// differences and their products here); a few ulps, not correctly rounded.  Elsewhere one
// reciprocal each.
template <typename T>
__host__ __device__ __forceinline__ void
rcp2(T a, T b, T & ia, T & ib)
{
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (sizeof(T) == 8)
  {
    const T i = rcp_nr(a * b);
    ia = i * b;
    ib = i * a;
    return;
  }
#endif
  ia = rcp_nr(a);
  ib = rcp_nr(b);
}

template <typename T>
__host__ __device__ __forceinline__ void
rcp4(T a, T b, T c, T d, T & ia, T & ib, T & ic, T & id)
{
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (sizeof(T) == 8)
  {
    const T ab = a * b, cd = c * d;
    const T i = rcp_nr(ab * cd);
    const T iab = i * cd, icd = i * ab;
    ia = iab * b;
    ib = iab * a;
    ic = icd * d;
    id = icd * c;
    return;
  }
#endif
  ia = rcp_nr(a);
  ib = rcp_nr(b);
  ic = rcp_nr(c);
  id = rcp_nr(d);
}

// HLLD flux along x (riemann_hlld, models/RiemannSolvers_MHD.h:311-551), shared denominators:
// 1/(rcr + rcl), 1/(sl - ustar), 1/(sr - ustar), 1/e*_L, 1/e*_R, 1/(sqrt rho*_L + sqrt rho*_R),
// and one reciprocal density per state in the fast speeds.
template <typename T>
__host__ __device__ inline void
hlld_fast(T * ql, T * qr, T * f, const Gas<T> & g)
{
  const T entho = rcp_nr(g.gamma0 - T(1));
  const T a = T(0.5) * (ql[kA] + qr[kA]);
  const T sgnm = (a >= T(0)) ? T(1) : T(-1);
  ql[kA] = a;
  qr[kA] = a;
  const T a2 = a * a;

  const T rl = ql[kD], pl = ql[kP], ul = ql[kU], vl = ql[kV], wl = ql[kW], bl = ql[kB], cl = ql[kC];
  const T ecinl = T(0.5) * (ul * ul + vl * vl + wl * wl) * rl;
  const T emagl = T(0.5) * (a2 + bl * bl + cl * cl);
  const T etotl = pl * entho + ecinl + emagl;
  const T ptotl = pl + emagl;
  const T vdotbl = ul * a + vl * bl + wl * cl;

  const T rr = qr[kD], pr = qr[kP], ur = qr[kU], vr = qr[kV], wr = qr[kW], br = qr[kB], cr = qr[kC];
  const T ecinr = T(0.5) * (ur * ur + vr * vr + wr * wr) * rr;
  const T emagr = T(0.5) * (a2 + br * br + cr * cr);
  const T etotr = pr * entho + ecinr + emagr;
  const T ptotr = pr + emagr;
  const T vdotbr = ur * a + vr * br + wr * cr;

  // fast speeds along x: d2 = (B^2/rho + c^2)/2, c_f = sqrt(d2 + sqrt(d2^2 - c^2 Bx^2 / rho)); the
  // larger of the two is the square root of the larger radicand
  auto cfast2 = [&](T r, T p, T bb) {
    const T ir = rcp_nr(r);
    const T c2 = g.gamma0 * p * ir;
    const T d2 = T(0.5) * (bb * ir + c2);
    return d2 + sqrt_nr(d2 * d2 - c2 * a2 * ir);
  };
  const T cmax = sqrt_nr(fmax(cfast2(rl, pl, T(2) * emagl), cfast2(rr, pr, T(2) * emagr)));
  const T sl = fmin(ul, ur) - cmax;
  const T sr = fmax(ul, ur) + cmax;

  const T rcl = rl * (ul - sl);
  const T rcr = rr * (sr - ur);
  const T irc = rcp_nr(rcr + rcl);
  const T ustar = (rcr * ur + rcl * ul + (ptotl - ptotr)) * irc;
  const T ptotstar = (rcr * ptotl + rcl * ptotr + rcl * rcr * (ul - ur)) * irc;

  // left star region
  const T isl = rcp_nr(sl - ustar);
  const T rstarl = rl * (sl - ul) * isl;
  const T estarl = rl * (sl - ul) * (sl - ustar) - a2;
  const T el = rl * (sl - ul) * (sl - ul) - a2;
  T       vstarl, wstarl, bstarl, cstarl;
  if (fabs(estarl) < T(1e-4) * a2)
  {
    vstarl = vl;
    bstarl = bl;
    wstarl = wl;
    cstarl = cl;
  }
  else
  {
    const T ie = rcp_nr(estarl);
    const T aus = a * (ustar - ul) * ie;
    const T elr = el * ie;
    vstarl = vl - bl * aus;
    bstarl = bl * elr;
    wstarl = wl - cl * aus;
    cstarl = cl * elr;
  }
  const T vdotbstarl = ustar * a + vstarl * bstarl + wstarl * cstarl;
  const T etotstarl = ((sl - ul) * etotl - ptotl * ul + ptotstar * ustar + a * (vdotbl - vdotbstarl)) * isl;
  const T sqrrstarl = sqrt_nr(rstarl);

  // right star region
  const T isr = rcp_nr(sr - ustar);
  const T rstarr = rr * (sr - ur) * isr;
  const T estarr = rr * (sr - ur) * (sr - ustar) - a2;
  const T er = rr * (sr - ur) * (sr - ur) - a2;
  T       vstarr, wstarr, bstarr, cstarr;
  if (fabs(estarr) < T(1e-4) * a2)
  {
    vstarr = vr;
    bstarr = br;
    wstarr = wr;
    cstarr = cr;
  }
  else
  {
    const T ie = rcp_nr(estarr);
    const T aus = a * (ustar - ur) * ie;
    const T err = er * ie;
    vstarr = vr - br * aus;
    bstarr = br * err;
    wstarr = wr - cr * aus;
    cstarr = cr * err;
  }
  const T vdotbstarr = ustar * a + vstarr * bstarr + wstarr * cstarr;
  const T etotstarr = ((sr - ur) * etotr - ptotr * ur + ptotstar * ustar + a * (vdotbr - vdotbstarr)) * isr;
  const T sqrrstarr = sqrt_nr(rstarr);

  // Alfven waves: u* -+ |Bx| / sqrt(rho*), compared with 0 as |Bx| / sqrt(rho*) against u*
  const T sal = ustar - fabs(a) * rcp_nr(sqrrstarl);
  const T sar = ustar + fabs(a) * rcp_nr(sqrrstarr);

  // double-star region
  const T iden = rcp_nr(sqrrstarl + sqrrstarr);
  const T sq2 = sqrrstarl * sqrrstarr;
  const T vss = (sqrrstarl * vstarl + sqrrstarr * vstarr + sgnm * (bstarr - bstarl)) * iden;
  const T wss = (sqrrstarl * wstarl + sqrrstarr * wstarr + sgnm * (cstarr - cstarl)) * iden;
  const T bss = (sqrrstarl * bstarr + sqrrstarr * bstarl + sgnm * sq2 * (vstarr - vstarl)) * iden;
  const T css = (sqrrstarl * cstarr + sqrrstarr * cstarl + sgnm * sq2 * (wstarr - wstarl)) * iden;
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
  f[kU] = ro * uo * uo - a2 + ptoto;
  f[kV] = ro * uo * vo - a * bo;
  f[kW] = ro * uo * wo - a * co;
  f[kA] = T(0);
  f[kB] = bo * uo - a * vo;
  f[kC] = co * uo - a * wo;
}

// 2D HLLD electric field at an edge (mag_riemann2d_hlld and compute_emf,
// models/RiemannSolvers_MHD.h:629-869, 890-1097) from the four corner states (RT, RB, LT, LB) in
// the lab frame; AXIS the edge direction.  Shared: one reciprocal density per state (fast speeds
// in x and y), 1/(SL - u*), 1/(SR - u*), 1/(SB - v*), 1/(ST - v*) for the star states,
// 1/(SAR - SAL), 1/(SAT - SAB) for the averages and the result; each Alfven speed is the square
// root of the largest of B_n^2 / rho* over its four terms (rho* B_n^2 scaled by the same ratios
// as in the reference).  The 20 reciprocals come from 6 (rcp4, rcp2: batch inversion in FP64).
template <int AXIS, typename T>
__host__ __device__ inline T
edge_emf_fast(const T (&qe)[4][8], const Gas<T> & g)
{
  constexpr int iu = AXIS == 2 ? kU : (AXIS == 1 ? kW : kV);
  constexpr int iv = AXIS == 2 ? kV : (AXIS == 1 ? kU : kW);
  constexpr int ia = AXIS == 2 ? kA : (AXIS == 1 ? kC : kB);
  constexpr int ib = AXIS == 2 ? kB : (AXIS == 1 ? kA : kC);
  constexpr int ic = AXIS == 2 ? kC : (AXIS == 1 ? kB : kA);
  const T * RT = qe[kRT];
  const T * LT = qe[kLT];
  const T * RB = qe[kRB];
  const T * LB = qe[kLB];
  // the LL, RL, LR, RR states of the magnetic Riemann problem (compute_emf)
  const T aB = T(0.5) * (RT[ia] + LT[ia]), aT = T(0.5) * (RB[ia] + LB[ia]);
  const T bL = T(0.5) * (RT[ib] + RB[ib]), bR = T(0.5) * (LT[ib] + LB[ib]);
  const T rLL = RT[kD], pLL = RT[kP], uLL = RT[iu], vLL = RT[iv], cLL = RT[ic];
  const T rRL = LT[kD], pRL = LT[kP], uRL = LT[iu], vRL = LT[iv], cRL = LT[ic];
  const T rLR = RB[kD], pLR = RB[kP], uLR = RB[iu], vLR = RB[iv], cLR = RB[ic];
  const T rRR = LB[kD], pRR = LB[kP], uRR = LB[iu], vRR = LB[iv], cRR = LB[ic];
  const T aLL = aB, aRL = aB, aLR = aT, aRR = aT;
  const T bLL = bL, bLR = bL, bRL = bR, bRR = bR;

  const T ELL = uLL * bLL - vLL * aLL;
  const T ERL = uRL * bRL - vRL * aRL;
  const T ELR = uLR * bLR - vLR * aLR;
  const T ERR = uRR * bRR - vRR * aRR;

  // fast speeds along x and y with one reciprocal density per state: the radicands
  // d2 + sqrt(d2^2 - c^2 B_n^2 / rho); the largest speed is the square root of the largest
  T irLL, irLR, irRL, irRR;
  rcp4(rLL, rLR, rRL, rRR, irLL, irLR, irRL, irRR);
  auto cfast2 = [&](T ir, T p, T a, T b, T c, T & cx, T & cy) {
    const T c2 = g.gamma0 * p * ir;
    const T d2 = T(0.5) * ((a * a + b * b + c * c) * ir + c2);
    const T dd = d2 * d2;
    const T ci = c2 * ir;
    cx = d2 + sqrt_nr(dd - ci * a * a);
    cy = d2 + sqrt_nr(dd - ci * b * b);
  };
  T cLLx, cLLy, cLRx, cLRy, cRLx, cRLy, cRRx, cRRy;
  cfast2(irLL, pLL, aLL, bLL, cLL, cLLx, cLLy);
  cfast2(irLR, pLR, aLR, bLR, cLR, cLRx, cLRy);
  cfast2(irRL, pRL, aRL, bRL, cRL, cRLx, cRLy);
  cfast2(irRR, pRR, aRR, bRR, cRR, cRRx, cRRy);
  const T cmx = sqrt_nr(max4(cLLx, cLRx, cRLx, cRRx));
  const T cmy = sqrt_nr(max4(cLLy, cLRy, cRLy, cRRy));

  const T SL = min4(uLL, uLR, uRL, uRR) - cmx;
  const T SR = max4(uLL, uLR, uRL, uRR) + cmx;
  const T SB = min4(vLL, vLR, vRL, vRR) - cmy;
  const T ST = max4(vLL, vLR, vRL, vRR) + cmy;

  const T PtLL = pLL + T(0.5) * (aLL * aLL + bLL * bLL + cLL * cLL);
  const T PtLR = pLR + T(0.5) * (aLR * aLR + bLR * bLR + cLR * cLR);
  const T PtRL = pRL + T(0.5) * (aRL * aRL + bRL * bRL + cRL * cRL);
  const T PtRR = pRR + T(0.5) * (aRR * aRR + bRR * bRR + cRR * cRR);

  const T rcLLx = rLL * (uLL - SL), rcRLx = rRL * (SR - uRL);
  const T rcLRx = rLR * (uLR - SL), rcRRx = rRR * (SR - uRR);
  const T rcLLy = rLL * (vLL - SB), rcLRy = rLR * (ST - vLR);
  const T rcRLy = rRL * (vRL - SB), rcRRy = rRR * (ST - vRR);

  T irx, iry;
  rcp2(rcLLx + rcLRx + rcRLx + rcRRx, rcLLy + rcLRy + rcRLy + rcRRy, irx, iry);
  const T ustar = (rcLLx * uLL + rcLRx * uLR + rcRLx * uRL + rcRRx * uRR + (PtLL - PtRL + PtLR - PtRR)) * irx;
  const T vstar = (rcLLy * vLL + rcLRy * vLR + rcRLy * vRL + rcRRy * vRR + (PtLL - PtLR + PtRL - PtRR)) * iry;

  // compression ratios across the outer waves: f = (S - u)/(S - u*)
  T iSL, iSR, iSB, iST;
  rcp4(SL - ustar, SR - ustar, SB - vstar, ST - vstar, iSL, iSR, iSB, iST);
  const T fLLx = (SL - uLL) * iSL, fLRx = (SL - uLR) * iSL, fRLx = (SR - uRL) * iSR, fRRx = (SR - uRR) * iSR;
  const T fLLy = (SB - vLL) * iSB, fRLy = (SB - vRL) * iSB, fLRy = (ST - vLR) * iST, fRRy = (ST - vRR) * iST;

  const T BsLL = bLL * fLLx, AsLL = aLL * fLLy;
  const T BsLR = bLR * fLRx, AsLR = aLR * fLRy;
  const T BsRL = bRL * fRLx, AsRL = aRL * fRLy;
  const T BsRR = bRR * fRRx, AsRR = aRR * fRRy;
  const T EsLLx = ustar * BsLL - vLL * aLL, EsLLy = uLL * bLL - vstar * AsLL, EsLL = ustar * BsLL - vstar * AsLL;
  const T EsLRx = ustar * BsLR - vLR * aLR, EsLRy = uLR * bLR - vstar * AsLR, EsLR = ustar * BsLR - vstar * AsLR;
  const T EsRLx = ustar * BsRL - vRL * aRL, EsRLy = uRL * bRL - vstar * AsRL, EsRL = ustar * BsRL - vstar * AsRL;
  const T EsRRx = ustar * BsRR - vRR * aRR, EsRRy = uRR * bRR - vstar * AsRR, EsRR = ustar * BsRR - vstar * AsRR;

  // Alfven speeds: calfvenL = max(|aLR|/sqrt(rho*LRx), |A*LR|/sqrt(rho*LR), |aLL|/sqrt(rho*LLx),
  // |A*LL|/sqrt(rho*LL), smallc) with rho*x = rho f_x, rho* = rho f_x f_y, A* = a f_y: the squares
  // are a^2 / (rho f_x) times 1 and f_y.  Likewise calfvenB with b, f_y and f_x.
  auto sq = [](T x) { return x * x; };
  T ixLL, ixLR, ixRL, ixRR, iyLL, iyLR, iyRL, iyRR;
  rcp4(rLL * fLLx, rLR * fLRx, rRL * fRLx, rRR * fRRx, ixLL, ixLR, ixRL, ixRR);
  rcp4(rLL * fLLy, rLR * fLRy, rRL * fRLy, rRR * fRRy, iyLL, iyLR, iyRL, iyRR);
  const T sc2 = g.smallc * g.smallc;
  const T caL = sqrt_nr(max5(sq(aLR) * ixLR, sq(aLR) * ixLR * fLRy, sq(aLL) * ixLL, sq(aLL) * ixLL * fLLy, sc2));
  const T caR = sqrt_nr(max5(sq(aRR) * ixRR, sq(aRR) * ixRR * fRRy, sq(aRL) * ixRL, sq(aRL) * ixRL * fRLy, sc2));
  const T caB = sqrt_nr(max5(sq(bLL) * iyLL, sq(bLL) * iyLL * fLLx, sq(bRL) * iyRL, sq(bRL) * iyRL * fRLx, sc2));
  const T caT = sqrt_nr(max5(sq(bLR) * iyLR, sq(bLR) * iyLR * fLRx, sq(bRR) * iyRR, sq(bRR) * iyRR * fRRx, sc2));

  const T SAL = fmin(ustar - caL, T(0));
  const T SAR = fmax(ustar + caR, T(0));
  const T SAB = fmin(vstar - caB, T(0));
  const T SAT = fmax(vstar + caT, T(0));
  T iLR, iBT;
  rcp2(SAR - SAL, SAT - SAB, iLR, iBT);

  const T AsT = (SAR * AsRR - SAL * AsLR) * iLR;
  const T AsB = (SAR * AsRL - SAL * AsLL) * iLR;
  const T BsR = (SAT * BsRR - SAB * BsRL) * iBT;
  const T BsL = (SAT * BsLR - SAB * BsLL) * iBT;

  const T SBp = (T(1) + copysign(T(1), SB)) * T(0.5), SBn = T(1) - SBp;
  const T STp = (T(1) + copysign(T(1), ST)) * T(0.5), STn = T(1) - STp;
  const T SLp = (T(1) + copysign(T(1), SL)) * T(0.5), SLn = T(1) - SLp;
  const T SRp = (T(1) + copysign(T(1), SR)) * T(0.5), SRn = T(1) - SRp;

  T E = 0, t;
  t = (SAL * SAB * EsRR - SAL * SAT * EsRL - SAR * SAB * EsLR + SAR * SAT * EsLL) * iLR * iBT -
      SAT * SAB * iBT * (AsT - AsB) + SAR * SAL * iLR * (BsR - BsL);
  E += (SBn * STp * SLn * SRp) * t;

  t = (SAR * EsLLx - SAL * EsRLx + SAR * SAL * (bRL - bLL)) * iLR;
  t = SLp * ELL + SLn * SRn * ERL + SLn * SRp * t;
  E += SBp * t;

  t = (SAR * EsLRx - SAL * EsRRx + SAR * SAL * (bRR - bLR)) * iLR;
  t = SLp * ELR + SLn * SRn * ERR + SLn * SRp * t;
  E += (SBn * STn) * t;

  t = (SAT * EsLLy - SAB * EsLRy - SAT * SAB * (aLR - aLL)) * iBT;
  E += (SBn * STp * SLp) * t;

  t = (SAT * EsRLy - SAB * EsRRy - SAT * SAB * (aRR - aRL)) * iBT;
  E += (SBn * STp * SLn * SRn) * t;
  return E;
}

} // namespace mhdct
