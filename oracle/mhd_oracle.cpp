// mhd_oracle.cpp -- TEST INFRASTRUCTURE ONLY (the parity oracle).
//
// A plain C++ FP64 CPU restatement of the reference's per-block MHD-CT time step
// (GodunovImplemV0::do_time_step, scheme/GodunovImplemV0.cpp:101-207), of its dt
// reduction (SolverGodunovMHD.cpp:377-403, scheme/ComputeDtMHDFunctor.cpp:86-151) and of
// its outside-block fill (border_conditions/FillOutside.cpp:97-395).
//
// Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may load this
// library, and only as the checker.  Nothing in the shipped package links or calls it.
//
// It keeps the reference's stage order, flat loops (octant x cell) and expression order,
// and is compiled with -ffp-contract=off and without fast-math, so that every +,-,*,/,sqrt
// is one IEEE operation in the order the reference writes it.
//
// The reference delegates neighbour search, face-B ghost fill, restriction/prolongation
// and edge-sibling lookup to kalypsso::core, which is not in the reference tree.  Those
// are restated here from first principles (see the comments on Mesh::locate,
// fill_Bface_ghosted and correct_emf).
//
// Parity status per function (see tests/test_oracle_*.py for the checks):
//   compute_primitives, slopes, predictor, E, sFaceMag, HLLD, 2D-HLLD EMF, update
//   stages, dt          : pinned by closed forms / invariants / brute force in tests.
//   RK2_SSP stage (no predictor, no face-B source at the edges) and the stage combination:
//   pinned by the closed-form forward-Euler MUSCL stage of a 1D advected contact and by the
//   second-order time convergence of the two-stage scheme (tests/test_oracle_rk2.py).
//   AMR ghost fill, flux and EMF correction (core semantics): parity unpinned against
//   the true reference (kalypsso::core absent); pinned by invariants only (div B,
//   conservation, shared-face equality, uniform-state preservation).

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace oracle
{

// MHD::VarId (models/MHD.h:35-62)
enum
{
  ID = 0,
  IP = 1,
  IE = 1,
  IU = 2,
  IV = 3,
  IW = 4,
  IA = 5,
  IB = 6,
  IC = 7,
  IAL = 5,
  IBL = 6,
  ICL = 7,
  IAR = 8,
  IBR = 9,
  ICR = 10
};
// FaceMagState ids (models/MHD.h:84-92)
enum
{
  AL = 0,
  BL = 1,
  CL = 2,
  AR = 3,
  BR = 4,
  CR = 5
};
enum
{
  IX = 0,
  IY = 1,
  IZ = 2
};

using State = double[8];

struct Settings
{
  // HydroSettings (models/HydroSettings.cpp:19-30)
  double gamma0 = 1.4;
  double cfl = 0.5;
  double slope_type = 1.0;
  double smallr = 1e-10;
  double smallc = 1e-10;
  double smallp = 0; // smallc^2/gamma0
  double scaling = 1.0;
  int    abort_when_negative_eint = 0;
  // [run] time_integrator: 0 HANCOCK, 1 RK2_SSP (kalypsso::core's TimeIntegrator; read by the
  // functors through TimeIntegratorConfig::get_time_integrator)
  int    time_integrator = 0;
  // [gravity] enabled, uniform_gx, uniform_gy, uniform_gz (get_uniform_gravity_vector, core)
  int    gravity_enabled = 0;
  double gravity[3] = { 0, 0, 0 };
  // [viscosity] enabled, mu, hancock_predictor_enabled (ViscosityParams, core)
  int    viscosity_enabled = 0;
  double mu = 0.0;
  int    viscosity_predictor = 0;
  // [amr] cell_centered_prolongation_type of the ghost cells beside a coarser block
  // (ConvertToPrimitivesVariablesFunctor.cpp:697-710): 0 SIMPLE_COPY, 1 EXTRAPOLATE_LINEAR_MINMOD
  int    prolongation = 0;
};

// ---------------------------------------------------------------------------
// models/mhd_utils.h restated
// ---------------------------------------------------------------------------

// find_speed_fast<dir> (models/mhd_utils.h:93-122)
static inline double
find_speed_fast(const double * q, int dir, const Settings & s)
{
  const double gamma0 = s.gamma0;
  const double d = q[ID];
  const double p = q[IP];
  const double a = q[IA];
  const double b = q[IB];
  const double c = q[IC];
  const double b2 = a * a + b * b + c * c;
  const double c2 = gamma0 * p / d;
  const double d2 = 0.5 * (b2 / d + c2);
  double       cf = 0;
  if (dir == IX)
    cf = std::sqrt(d2 + std::sqrt(d2 * d2 - c2 * a * a / d));
  else if (dir == IY)
    cf = std::sqrt(d2 + std::sqrt(d2 * d2 - c2 * b * b / d));
  else
    cf = std::sqrt(d2 + std::sqrt(d2 * d2 - c2 * c * c / d));
  return cf;
}

// find_speed_info<dim> (models/mhd_utils.h:340-394)
static inline void
find_speed_info(const double * q, int dim, const Settings & s, double out[3])
{
  const double gamma0 = s.gamma0;
  const double d = q[ID];
  const double p = q[IP];
  const double u = q[IU];
  const double v = q[IV];
  const double w = q[IW];
  const double a = q[IA];
  const double b = q[IB];
  const double c = q[IC];
  const double b2 = a * a + b * b + c * c;
  const double c2 = gamma0 * p / d;
  const double d2 = 0.5 * (b2 / d + c2);
  double       cf = std::sqrt(d2 + std::sqrt(d2 * d2 - c2 * a * a / d));
  out[IX] = cf + std::fabs(u);
  cf = std::sqrt(d2 + std::sqrt(d2 * d2 - c2 * b * b / d));
  out[IY] = cf + std::fabs(v);
  if (dim == 3)
  {
    cf = std::sqrt(d2 + std::sqrt(d2 * d2 - c2 * c * c / d));
    out[IZ] = cf + std::fabs(w);
  }
}

// compute_primitives (models/mhd_utils.h:449-538); cIso == 0 (perfect gas).
// Returns false when abort_when_negative_eint is set and eint < 0 (the reference aborts).
static inline bool
compute_primitives(const double * u, double * q, const Settings & s)
{
  q[ID] = std::fmax(u[ID], s.smallr);
  q[IU] = u[IU] / q[ID];
  q[IV] = u[IV] / q[ID];
  q[IW] = u[IW] / q[ID];
  q[IA] = u[IA];
  q[IB] = u[IB];
  q[IC] = u[IC];
  const double eken = 0.5 * (q[IU] * q[IU] + q[IV] * q[IV] + q[IW] * q[IW]);
  const double emag = 0.5 * (q[IA] * q[IA] + q[IB] * q[IB] + q[IC] * q[IC]);
  const double eint = (u[IE] - emag) / q[ID] - eken;
  q[IP] = std::fmax((s.gamma0 - 1.0) * q[ID] * eint, q[ID] * s.smallp);
  return !(s.abort_when_negative_eint && eint < 0);
}

// ---------------------------------------------------------------------------
// models/RiemannSolvers_MHD.h restated
// ---------------------------------------------------------------------------

// riemann_hlld (models/RiemannSolvers_MHD.h:311-551); cIso == 0.
static void
riemann_hlld(double * qleft, double * qright, double * flux, const Settings & st)
{
  const double gamma0 = st.gamma0;
  const double entho = 1.0 / (gamma0 - 1.0);

  double a = 0.5 * (qleft[IA] + qright[IA]);
  double sgnm = (a >= 0) ? 1.0 : -1.0;
  qleft[IA] = a;
  qright[IA] = a;

  double rl = qleft[ID], pl = qleft[IP], ul = qleft[IU], vl = qleft[IV], wl = qleft[IW];
  double bl = qleft[IB], cl = qleft[IC];
  double ecinl = 0.5 * (ul * ul + vl * vl + wl * wl) * rl;
  double emagl = 0.5 * (a * a + bl * bl + cl * cl);
  double etotl = pl * entho + ecinl + emagl;
  double ptotl = pl + emagl;
  double vdotbl = ul * a + vl * bl + wl * cl;

  double rr = qright[ID], pr = qright[IP], ur = qright[IU], vr = qright[IV], wr = qright[IW];
  double br = qright[IB], cr = qright[IC];
  double ecinr = 0.5 * (ur * ur + vr * vr + wr * wr) * rr;
  double emagr = 0.5 * (a * a + br * br + cr * cr);
  double etotr = pr * entho + ecinr + emagr;
  double ptotr = pr + emagr;
  double vdotbr = ur * a + vr * br + wr * cr;

  double cfastl = find_speed_fast(qleft, IX, st);
  double cfastr = find_speed_fast(qright, IX, st);

  double sl = std::fmin(ul, ur) - std::fmax(cfastl, cfastr);
  double sr = std::fmax(ul, ur) + std::fmax(cfastl, cfastr);

  double rcl = rl * (ul - sl);
  double rcr = rr * (sr - ur);

  double ustar = (rcr * ur + rcl * ul + (ptotl - ptotr)) / (rcr + rcl);
  double ptotstar = (rcr * ptotl + rcl * ptotr + rcl * rcr * (ul - ur)) / (rcr + rcl);

  double estar, rstarl, el;
  rstarl = rl * (sl - ul) / (sl - ustar);
  estar = rl * (sl - ul) * (sl - ustar) - a * a;
  el = rl * (sl - ul) * (sl - ul) - a * a;
  double vstarl, wstarl, bstarl, cstarl;
  if (std::fabs(estar) < 1e-4 * (a * a))
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
  double vdotbstarl = ustar * a + vstarl * bstarl + wstarl * cstarl;
  double etotstarl =
    ((sl - ul) * etotl - ptotl * ul + ptotstar * ustar + a * (vdotbl - vdotbstarl)) / (sl - ustar);
  double sqrrstarl = std::sqrt(rstarl);
  double calfvenl = std::fabs(a) / sqrrstarl;
  double sal = ustar - calfvenl;

  double rstarr, er;
  rstarr = rr * (sr - ur) / (sr - ustar);
  estar = rr * (sr - ur) * (sr - ustar) - a * a;
  er = rr * (sr - ur) * (sr - ur) - a * a;
  double vstarr, wstarr, bstarr, cstarr;
  if (std::fabs(estar) < 1e-4 * (a * a))
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
  double vdotbstarr = ustar * a + vstarr * bstarr + wstarr * cstarr;
  double etotstarr =
    ((sr - ur) * etotr - ptotr * ur + ptotstar * ustar + a * (vdotbr - vdotbstarr)) / (sr - ustar);
  double sqrrstarr = std::sqrt(rstarr);
  double calfvenr = std::fabs(a) / sqrrstarr;
  double sar = ustar + calfvenr;

  double vstarstar =
    (sqrrstarl * vstarl + sqrrstarr * vstarr + sgnm * (bstarr - bstarl)) / (sqrrstarl + sqrrstarr);
  double wstarstar =
    (sqrrstarl * wstarl + sqrrstarr * wstarr + sgnm * (cstarr - cstarl)) / (sqrrstarl + sqrrstarr);
  double bstarstar =
    (sqrrstarl * bstarr + sqrrstarr * bstarl + sgnm * sqrrstarl * sqrrstarr * (vstarr - vstarl)) /
    (sqrrstarl + sqrrstarr);
  double cstarstar =
    (sqrrstarl * cstarr + sqrrstarr * cstarl + sgnm * sqrrstarl * sqrrstarr * (wstarr - wstarl)) /
    (sqrrstarl + sqrrstarr);
  double vdotbstarstar = ustar * a + vstarstar * bstarstar + wstarstar * cstarstar;
  double etotstarstarl = etotstarl - sgnm * sqrrstarl * (vdotbstarl - vdotbstarstar);
  double etotstarstarr = etotstarr + sgnm * sqrrstarr * (vdotbstarr - vdotbstarstar);

  double ro, uo, vo, wo, bo, co, ptoto, etoto, vdotbo;
  if (sl > 0)
  {
    ro = rl; uo = ul; vo = vl; wo = wl; bo = bl; co = cl;
    ptoto = ptotl; etoto = etotl; vdotbo = vdotbl;
  }
  else if (sal > 0)
  {
    ro = rstarl; uo = ustar; vo = vstarl; wo = wstarl; bo = bstarl; co = cstarl;
    ptoto = ptotstar; etoto = etotstarl; vdotbo = vdotbstarl;
  }
  else if (ustar > 0)
  {
    ro = rstarl; uo = ustar; vo = vstarstar; wo = wstarstar; bo = bstarstar; co = cstarstar;
    ptoto = ptotstar; etoto = etotstarstarl; vdotbo = vdotbstarstar;
  }
  else if (sar > 0)
  {
    ro = rstarr; uo = ustar; vo = vstarstar; wo = wstarstar; bo = bstarstar; co = cstarstar;
    ptoto = ptotstar; etoto = etotstarstarr; vdotbo = vdotbstarstar;
  }
  else if (sr > 0)
  {
    ro = rstarr; uo = ustar; vo = vstarr; wo = wstarr; bo = bstarr; co = cstarr;
    ptoto = ptotstar; etoto = etotstarr; vdotbo = vdotbstarr;
  }
  else
  {
    ro = rr; uo = ur; vo = vr; wo = wr; bo = br; co = cr;
    ptoto = ptotr; etoto = etotr; vdotbo = vdotbr;
  }

  flux[ID] = ro * uo;
  flux[IP] = (etoto + ptoto) * uo - a * vdotbo;
  flux[IU] = ro * uo * uo - a * a + ptoto;
  flux[IV] = ro * uo * vo - a * bo;
  flux[IW] = ro * uo * wo - a * co;
  flux[IA] = 0.0;
  flux[IB] = bo * uo - a * vo;
  flux[IC] = co * uo - a * wo;
}

static inline double
FMAX4(double a0, double a1, double a2, double a3)
{
  double r = a0;
  r = (a1 > r) ? a1 : r;
  r = (a2 > r) ? a2 : r;
  r = (a3 > r) ? a3 : r;
  return r;
}
static inline double
FMIN4(double a0, double a1, double a2, double a3)
{
  double r = a0;
  r = (a1 < r) ? a1 : r;
  r = (a2 < r) ? a2 : r;
  r = (a3 < r) ? a3 : r;
  return r;
}
static inline double
FMAX5(double a0, double a1, double a2, double a3, double a4)
{
  double r = a0;
  r = (a1 > r) ? a1 : r;
  r = (a2 > r) ? a2 : r;
  r = (a3 > r) ? a3 : r;
  r = (a4 > r) ? a4 : r;
  return r;
}

// index of the four states of the 2D magnetic Riemann problem
enum
{
  ILL = 0,
  IRL = 1,
  ILR = 2,
  IRR = 3
};
// index of the four reconstructed edge states
enum
{
  IRT = 0,
  IRB = 1,
  ILT = 2,
  ILB = 3
};

// mag_riemann2d_hlld (models/RiemannSolvers_MHD.h:629-869)
static double
mag_riemann2d_hlld(double qLLRR[4][8], double eLLRR[4], const Settings & st)
{
  const double * qLL = qLLRR[ILL];
  const double * qRL = qLLRR[IRL];
  const double * qLR = qLLRR[ILR];
  const double * qRR = qLLRR[IRR];
  const double   ELL = eLLRR[ILL];
  const double   ERL = eLLRR[IRL];
  const double   ELR = eLLRR[ILR];
  const double   ERR = eLLRR[IRR];

  const double rLL = qLL[ID], pLL = qLL[IP], uLL = qLL[IU], vLL = qLL[IV];
  const double aLL = qLL[IA], bLL = qLL[IB], cLL = qLL[IC];
  const double rLR = qLR[ID], pLR = qLR[IP], uLR = qLR[IU], vLR = qLR[IV];
  const double aLR = qLR[IA], bLR = qLR[IB], cLR = qLR[IC];
  const double rRL = qRL[ID], pRL = qRL[IP], uRL = qRL[IU], vRL = qRL[IV];
  const double aRL = qRL[IA], bRL = qRL[IB], cRL = qRL[IC];
  const double rRR = qRR[ID], pRR = qRR[IP], uRR = qRR[IU], vRR = qRR[IV];
  const double aRR = qRR[IA], bRR = qRR[IB], cRR = qRR[IC];

  double cFastLLx = find_speed_fast(qLL, IX, st);
  double cFastLRx = find_speed_fast(qLR, IX, st);
  double cFastRLx = find_speed_fast(qRL, IX, st);
  double cFastRRx = find_speed_fast(qRR, IX, st);
  double cFastLLy = find_speed_fast(qLL, IY, st);
  double cFastLRy = find_speed_fast(qLR, IY, st);
  double cFastRLy = find_speed_fast(qRL, IY, st);
  double cFastRRy = find_speed_fast(qRR, IY, st);

  double SL = FMIN4(uLL, uLR, uRL, uRR) - FMAX4(cFastLLx, cFastLRx, cFastRLx, cFastRRx);
  double SR = FMAX4(uLL, uLR, uRL, uRR) + FMAX4(cFastLLx, cFastLRx, cFastRLx, cFastRRx);
  double SB = FMIN4(vLL, vLR, vRL, vRR) - FMAX4(cFastLLy, cFastLRy, cFastRLy, cFastRRy);
  double ST = FMAX4(vLL, vLR, vRL, vRR) + FMAX4(cFastLLy, cFastLRy, cFastRLy, cFastRRy);

  double PtotLL = pLL + 0.5 * (aLL * aLL + bLL * bLL + cLL * cLL);
  double PtotLR = pLR + 0.5 * (aLR * aLR + bLR * bLR + cLR * cLR);
  double PtotRL = pRL + 0.5 * (aRL * aRL + bRL * bRL + cRL * cRL);
  double PtotRR = pRR + 0.5 * (aRR * aRR + bRR * bRR + cRR * cRR);

  double rcLLx = rLL * (uLL - SL);
  double rcRLx = rRL * (SR - uRL);
  double rcLRx = rLR * (uLR - SL);
  double rcRRx = rRR * (SR - uRR);
  double rcLLy = rLL * (vLL - SB);
  double rcLRy = rLR * (ST - vLR);
  double rcRLy = rRL * (vRL - SB);
  double rcRRy = rRR * (ST - vRR);

  double ustar =
    (rcLLx * uLL + rcLRx * uLR + rcRLx * uRL + rcRRx * uRR + (PtotLL - PtotRL + PtotLR - PtotRR)) /
    (rcLLx + rcLRx + rcRLx + rcRRx);
  double vstar =
    (rcLLy * vLL + rcLRy * vLR + rcRLy * vRL + rcRRy * vRR + (PtotLL - PtotLR + PtotRL - PtotRR)) /
    (rcLLy + rcLRy + rcRLy + rcRRy);

  double rstarLLx = rLL * (SL - uLL) / (SL - ustar);
  double BstarLL = bLL * (SL - uLL) / (SL - ustar);
  double rstarLLy = rLL * (SB - vLL) / (SB - vstar);
  double AstarLL = aLL * (SB - vLL) / (SB - vstar);
  double rstarLL = rLL * (SL - uLL) / (SL - ustar) * (SB - vLL) / (SB - vstar);
  double EstarLLx = ustar * BstarLL - vLL * aLL;
  double EstarLLy = uLL * bLL - vstar * AstarLL;
  double EstarLL = ustar * BstarLL - vstar * AstarLL;

  double rstarLRx = rLR * (SL - uLR) / (SL - ustar);
  double BstarLR = bLR * (SL - uLR) / (SL - ustar);
  double rstarLRy = rLR * (ST - vLR) / (ST - vstar);
  double AstarLR = aLR * (ST - vLR) / (ST - vstar);
  double rstarLR = rLR * (SL - uLR) / (SL - ustar) * (ST - vLR) / (ST - vstar);
  double EstarLRx = ustar * BstarLR - vLR * aLR;
  double EstarLRy = uLR * bLR - vstar * AstarLR;
  double EstarLR = ustar * BstarLR - vstar * AstarLR;

  double rstarRLx = rRL * (SR - uRL) / (SR - ustar);
  double BstarRL = bRL * (SR - uRL) / (SR - ustar);
  double rstarRLy = rRL * (SB - vRL) / (SB - vstar);
  double AstarRL = aRL * (SB - vRL) / (SB - vstar);
  double rstarRL = rRL * (SR - uRL) / (SR - ustar) * (SB - vRL) / (SB - vstar);
  double EstarRLx = ustar * BstarRL - vRL * aRL;
  double EstarRLy = uRL * bRL - vstar * AstarRL;
  double EstarRL = ustar * BstarRL - vstar * AstarRL;

  double rstarRRx = rRR * (SR - uRR) / (SR - ustar);
  double BstarRR = bRR * (SR - uRR) / (SR - ustar);
  double rstarRRy = rRR * (ST - vRR) / (ST - vstar);
  double AstarRR = aRR * (ST - vRR) / (ST - vstar);
  double rstarRR = rRR * (SR - uRR) / (SR - ustar) * (ST - vRR) / (ST - vstar);
  double EstarRRx = ustar * BstarRR - vRR * aRR;
  double EstarRRy = uRR * bRR - vstar * AstarRR;
  double EstarRR = ustar * BstarRR - vstar * AstarRR;

  double calfvenL = FMAX5(std::fabs(aLR) / std::sqrt(rstarLRx),
                          std::fabs(AstarLR) / std::sqrt(rstarLR),
                          std::fabs(aLL) / std::sqrt(rstarLLx),
                          std::fabs(AstarLL) / std::sqrt(rstarLL),
                          st.smallc);
  double calfvenR = FMAX5(std::fabs(aRR) / std::sqrt(rstarRRx),
                          std::fabs(AstarRR) / std::sqrt(rstarRR),
                          std::fabs(aRL) / std::sqrt(rstarRLx),
                          std::fabs(AstarRL) / std::sqrt(rstarRL),
                          st.smallc);
  double calfvenB = FMAX5(std::fabs(bLL) / std::sqrt(rstarLLy),
                          std::fabs(BstarLL) / std::sqrt(rstarLL),
                          std::fabs(bRL) / std::sqrt(rstarRLy),
                          std::fabs(BstarRL) / std::sqrt(rstarRL),
                          st.smallc);
  double calfvenT = FMAX5(std::fabs(bLR) / std::sqrt(rstarLRy),
                          std::fabs(BstarLR) / std::sqrt(rstarLR),
                          std::fabs(bRR) / std::sqrt(rstarRRy),
                          std::fabs(BstarRR) / std::sqrt(rstarRR),
                          st.smallc);

  double SAL = std::fmin(ustar - calfvenL, 0.0);
  double SAR = std::fmax(ustar + calfvenR, 0.0);
  double SAB = std::fmin(vstar - calfvenB, 0.0);
  double SAT = std::fmax(vstar + calfvenT, 0.0);

  double AstarT = (SAR * AstarRR - SAL * AstarLR) / (SAR - SAL);
  double AstarB = (SAR * AstarRL - SAL * AstarLL) / (SAR - SAL);
  double BstarR = (SAT * BstarRR - SAB * BstarRL) / (SAT - SAB);
  double BstarL = (SAT * BstarLR - SAB * BstarLL) / (SAT - SAB);

  double E = 0, tmpE = 0;
  double SB_pos = (1 + std::copysign(1.0, SB)) / 2, SB_neg = 1 - SB_pos;
  double ST_pos = (1 + std::copysign(1.0, ST)) / 2, ST_neg = 1 - ST_pos;
  double SL_pos = (1 + std::copysign(1.0, SL)) / 2, SL_neg = 1 - SL_pos;
  double SR_pos = (1 + std::copysign(1.0, SR)) / 2, SR_neg = 1 - SR_pos;

  tmpE = (SAL * SAB * EstarRR - SAL * SAT * EstarRL - SAR * SAB * EstarLR + SAR * SAT * EstarLL) /
           (SAR - SAL) / (SAT - SAB) -
         SAT * SAB / (SAT - SAB) * (AstarT - AstarB) + SAR * SAL / (SAR - SAL) * (BstarR - BstarL);
  E += (SB_neg * ST_pos * SL_neg * SR_pos) * tmpE;

  tmpE = (SAR * EstarLLx - SAL * EstarRLx + SAR * SAL * (bRL - bLL)) / (SAR - SAL);
  tmpE = SL_pos * ELL + SL_neg * SR_neg * ERL + SL_neg * SR_pos * tmpE;
  E += SB_pos * tmpE;

  tmpE = (SAR * EstarLRx - SAL * EstarRRx + SAR * SAL * (bRR - bLR)) / (SAR - SAL);
  tmpE = SL_pos * ELR + SL_neg * SR_neg * ERR + SL_neg * SR_pos * tmpE;
  E += (SB_neg * ST_neg) * tmpE;

  tmpE = (SAT * EstarLLy - SAB * EstarLRy - SAT * SAB * (aLR - aLL)) / (SAT - SAB);
  E += (SB_neg * ST_pos * SL_pos) * tmpE;

  tmpE = (SAT * EstarRLy - SAB * EstarRRy - SAT * SAB * (aRR - aRL)) / (SAT - SAB);
  E += (SB_neg * ST_pos * SL_neg * SR_neg) * tmpE;

  return E;
}

enum
{
  EMFX = 0,
  EMFY = 1,
  EMFZ = 2
};

// compute_emf<emfDir> (models/RiemannSolvers_MHD.h:890-1097)
static double
compute_emf(double qEdge[4][8], int emfDir, const Settings & st)
{
  const double * qRT = qEdge[IRT];
  const double * qLT = qEdge[ILT];
  const double * qRB = qEdge[IRB];
  const double * qLB = qEdge[ILB];
  double         qLLRR[4][8];
  double *       qLL = qLLRR[ILL];
  double *       qRL = qLLRR[IRL];
  double *       qLR = qLLRR[ILR];
  double *       qRR = qLLRR[IRR];

  qLL[ID] = qRT[ID];
  qRL[ID] = qLT[ID];
  qLR[ID] = qRB[ID];
  qRR[ID] = qLB[ID];
  qLL[IP] = qRT[IP];
  qRL[IP] = qLT[IP];
  qLR[IP] = qRB[IP];
  qRR[IP] = qLB[IP];

  // (iu, iv, iw, ia, ib, ic) of the rotated frame
  int iu, iv, iw, ia, ib, ic;
  if (emfDir == EMFZ)
  {
    iu = IU; iv = IV; iw = IW; ia = IA; ib = IB; ic = IC;
  }
  else if (emfDir == EMFY)
  {
    iu = IW; iv = IU; iw = IV; ia = IC; ib = IA; ic = IB;
  }
  else
  {
    iu = IV; iv = IW; iw = IU; ia = IB; ib = IC; ic = IA;
  }
  qLL[IU] = qRT[iu];
  qRL[IU] = qLT[iu];
  qLR[IU] = qRB[iu];
  qRR[IU] = qLB[iu];
  qLL[IV] = qRT[iv];
  qRL[IV] = qLT[iv];
  qLR[IV] = qRB[iv];
  qRR[IV] = qLB[iv];
  qLL[IA] = 0.5 * (qRT[ia] + qLT[ia]);
  qRL[IA] = 0.5 * (qRT[ia] + qLT[ia]);
  qLR[IA] = 0.5 * (qRB[ia] + qLB[ia]);
  qRR[IA] = 0.5 * (qRB[ia] + qLB[ia]);
  qLL[IB] = 0.5 * (qRT[ib] + qRB[ib]);
  qRL[IB] = 0.5 * (qLT[ib] + qLB[ib]);
  qLR[IB] = 0.5 * (qRT[ib] + qRB[ib]);
  qRR[IB] = 0.5 * (qLT[ib] + qLB[ib]);
  qLL[IW] = qRT[iw];
  qRL[IW] = qLT[iw];
  qLR[IW] = qRB[iw];
  qRR[IW] = qLB[iw];
  qLL[IC] = qRT[ic];
  qRL[IC] = qLT[ic];
  qLR[IC] = qRB[ic];
  qRR[IC] = qLB[ic];

  double eLLRR[4];
  eLLRR[ILL] = qLL[IU] * qLL[IB] - qLL[IV] * qLL[IA];
  eLLRR[IRL] = qRL[IU] * qRL[IB] - qRL[IV] * qRL[IA];
  eLLRR[ILR] = qLR[IU] * qLR[IB] - qLR[IV] * qLR[IA];
  eLLRR[IRR] = qRR[IU] * qRR[IB] - qRR[IV] * qRR[IA];

  return mag_riemann2d_hlld(qLLRR, eLLRR, st);
}

// slope_unsplit_scalar (scheme/ComputeLimitedSlopesFunctor.cpp:121-144)
static inline double
slope_unsplit_scalar(double q, double qPlus, double qMinus, double slope_type)
{
  double dq = 0;
  if (slope_type == 1 || slope_type == 2)
  {
    const double dlft = slope_type * (q - qMinus);
    const double drgt = slope_type * (qPlus - q);
    const double dcen = 0.5 * (qPlus - qMinus);
    const double dsgn = (dcen >= 0.0) ? 1.0 : -1.0;
    const double slop = std::fmin(std::fabs(dlft), std::fabs(drgt));
    const double dlim = (dlft * drgt) <= 0.0 ? 0.0 : slop;
    dq = dsgn * std::fmin(dlim, std::fabs(dcen));
  }
  return dq;
}

// ---------------------------------------------------------------------------
// Mesh: leaves of a brick forest (owned blocks first, then outside blocks).
// ---------------------------------------------------------------------------

struct Loc
{
  int oct = -1;   // block index
  int rel = 0;    // level(neigh) - level(cur): -1, 0, +1
  int ijk[3] = { 0, 0, 0 }; // local cell coords in the found block (lower fine cell if rel=+1)
};

struct Mesh
{
  int              dim = 2;
  int              b = 8;
  int              nowned = 0;
  int              noutside = 0;
  int              nbrick[3] = { 1, 1, 1 };
  int              periodic[3] = { 1, 1, 1 };
  int              bc[6] = { 0, 0, 0, 0, 0, 0 }; // 2 = ZERO_GRADIENT, 3 = WALL
  std::vector<int> level;
  std::vector<int> coord; // 3 per block, in units of blocks at that block's level
  std::unordered_map<uint64_t, int> hash;

  static uint64_t
  key(int l, int x, int y, int z)
  {
    const uint64_t off = 1u << 20;
    return (uint64_t(l) << 60) | ((uint64_t(x + off) & 0xFFFFF) << 40) |
           ((uint64_t(y + off) & 0xFFFFF) << 20) | (uint64_t(z + off) & 0xFFFFF);
  }
  int
  find(int l, int x, int y, int z) const
  {
    if (l < 0)
      return -1;
    auto it = hash.find(key(l, x, y, z));
    return it == hash.end() ? -1 : it->second;
  }
  int
  ntotal() const
  {
    return nowned + noutside;
  }
  // number of cells per dim over the whole brick at level l
  long
  ncells(int l, int d) const
  {
    return long(nbrick[d]) * (long(1) << l) * b;
  }
  static long
  fdiv(long a, long q)
  {
    return a >= 0 ? a / q : -((-a + q - 1) / q);
  }

  // Restates kalypsso::core StencilHelper::getNeighLoc (used by
  // ConvertToPrimitivesVariablesFunctor.cpp:682-683 and friends): the leaf containing the
  // cell (i,j,k) given in the frame of block `oct` (possibly outside [0,b)), at the same,
  // the next coarser or the next finer level.  Periodic directions wrap; beyond a bounded
  // border the outside blocks (the outward images of the border leaves, each at its leaf's
  // level) answer in the same way.
  Loc
  locate(int oct, int i, int j, int k) const
  {
    const int L = level[oct];
    long      g[3] = { long(coord[3 * oct + 0]) * b + i,
                       long(coord[3 * oct + 1]) * b + j,
                       long(coord[3 * oct + 2]) * b + (dim == 3 ? k : 0) };
    for (int d = 0; d < dim; ++d)
    {
      const long n = ncells(L, d);
      if ((g[d] < 0 || g[d] >= n) && periodic[d])
        g[d] = ((g[d] % n) + n) % n;
    }
    Loc r;
    // same level
    {
      int x = int(fdiv(g[0], b)), y = int(fdiv(g[1], b)), z = dim == 3 ? int(fdiv(g[2], b)) : 0;
      int o = find(L, x, y, z);
      if (o >= 0)
      {
        r.oct = o;
        r.rel = 0;
        r.ijk[0] = int(g[0] - long(x) * b);
        r.ijk[1] = int(g[1] - long(y) * b);
        r.ijk[2] = dim == 3 ? int(g[2] - long(z) * b) : 0;
        return r;
      }
    }
    // coarser
    {
      long gc[3] = { fdiv(g[0], 2), fdiv(g[1], 2), fdiv(g[2], 2) };
      int  x = int(fdiv(gc[0], b)), y = int(fdiv(gc[1], b)), z = dim == 3 ? int(fdiv(gc[2], b)) : 0;
      int  o = find(L - 1, x, y, z);
      if (o >= 0)
      {
        r.oct = o;
        r.rel = -1;
        r.ijk[0] = int(gc[0] - long(x) * b);
        r.ijk[1] = int(gc[1] - long(y) * b);
        r.ijk[2] = dim == 3 ? int(gc[2] - long(z) * b) : 0;
        return r;
      }
    }
    // finer
    {
      long gf[3] = { 2 * g[0], 2 * g[1], 2 * g[2] };
      int  x = int(fdiv(gf[0], b)), y = int(fdiv(gf[1], b)), z = dim == 3 ? int(fdiv(gf[2], b)) : 0;
      int  o = find(L + 1, x, y, z);
      if (o >= 0)
      {
        r.oct = o;
        r.rel = +1;
        r.ijk[0] = int(gf[0] - long(x) * b);
        r.ijk[1] = int(gf[1] - long(y) * b);
        r.ijk[2] = dim == 3 ? int(gf[2] - long(z) * b) : 0;
        return r;
      }
    }
    return r;
  }
};

// ---------------------------------------------------------------------------
// Reference containers (layout A0: octant slowest, then variable, then cell)
// ---------------------------------------------------------------------------

struct Layout
{
  int dim, b;
  // ghosted cell arrays with ghost width g
  int
  gn(int g) const
  {
    return b + 2 * g;
  }
  long
  gcells(int g) const
  {
    const long n = gn(g);
    return dim == 3 ? n * n * n : n * n;
  }
  long
  gidx(int g, int i, int j, int k) const
  {
    const long n = gn(g);
    return (i + g) + n * ((j + g) + (dim == 3 ? n * (k + g) : 0));
  }
  long
  cells() const
  {
    return dim == 3 ? long(b) * b * b : long(b) * b;
  }
  long
  cidx(int i, int j, int k) const
  {
    return i + long(b) * (j + (dim == 3 ? long(b) * k : 0));
  }
  // face arrays (no ghost): component c has extent b+1 along c (in 2D component Z is b x b)
  long
  fsize(int c) const
  {
    if (dim == 2)
      return c == IZ ? long(b) * b : long(b + 1) * b;
    return long(b + 1) * b * b;
  }
  long
  foff(int c) const
  {
    long o = 0;
    for (int d = 0; d < c; ++d)
      o += fsize(d);
    return o;
  }
  long
  fper() const
  {
    return fsize(0) + fsize(1) + fsize(2);
  }
  long
  fidx(int c, int i, int j, int k) const
  {
    const long nx = (c == IX && !(dim == 2 && c == IZ)) ? b + 1 : b;
    const long ny = (c == IY) ? b + 1 : b;
    return foff(c) + i + nx * (j + (dim == 3 ? ny * k : 0));
  }
  // ghosted face arrays (ghost width 2)
  long
  gfsize(int c) const
  {
    const long n = b + 4;
    if (dim == 2)
      return c == IZ ? n * n : (n + 1) * n;
    return (n + 1) * n * n;
  }
  long
  gfoff(int c) const
  {
    long o = 0;
    for (int d = 0; d < c; ++d)
      o += gfsize(d);
    return o;
  }
  long
  gfper() const
  {
    return gfsize(0) + gfsize(1) + gfsize(2);
  }
  long
  gfidx(int c, int i, int j, int k) const
  {
    const long n = b + 4;
    const long nx = (c == IX) ? n + 1 : n;
    const long ny = (c == IY) ? n + 1 : n;
    return gfoff(c) + (i + 2) + nx * ((j + 2) + (dim == 3 ? ny * (k + 2) : 0));
  }
};

class Solver
{
public:
  Settings s;
  Mesh     m;
  Layout   lay;
  int      nthreads = 1;

  // scratch (GodunovImplemV0.h:52-103)
  std::vector<double> Q, Q2, Sx, Sy, Sz, Fl, E, sF, emf, emf2, Bg;
  int                 flux_shape[3] = { 0, 0, 0 };
  std::vector<int>    conformal; // per owned block, per face (2*dim): 1 if neighbour finer
  std::string         err;

  double
  dx_of(int oct) const
  {
    // compute_cell_length(level, bx) * scaling_factor
    return 1.0 / double(long(lay.b) << m.level[oct]) * s.scaling;
  }

  void
  setup()
  {
    lay.dim = m.dim;
    lay.b = m.b;
    s.smallp = s.smallc * s.smallc / s.gamma0;
    const int nq = m.nowned;
    Q.assign(size_t(nq) * 11 * lay.gcells(2), 0.0);
    Q2.assign(size_t(nq) * 8 * lay.gcells(1), 0.0);
    Sx.assign(size_t(nq) * 8 * lay.gcells(1), 0.0);
    Sy.assign(size_t(nq) * 8 * lay.gcells(1), 0.0);
    Sz.assign(m.dim == 3 ? size_t(nq) * 8 * lay.gcells(1) : 0, 0.0);
    const long nf = m.dim == 3 ? long(lay.b + 1) * lay.b * lay.b : long(lay.b + 1) * lay.b;
    Fl.assign(size_t(nq) * 8 * nf, 0.0);
    E.assign(size_t(nq) * (m.dim == 3 ? 3 : 1) * lay.gcells(2), 0.0);
    sF.assign(size_t(nq) * (m.dim == 3 ? 3 : 2) * lay.gcells(1), 0.0);
    const long ne = m.dim == 3 ? long(lay.b + 1) * (lay.b + 1) * (lay.b + 1)
                               : long(lay.b + 1) * (lay.b + 1);
    emf.assign(size_t(nq) * (m.dim == 3 ? 3 : 1) * ne, 0.0);
    emf2.assign(emf.size(), 0.0);
    Bg.assign(size_t(m.ntotal()) * lay.gfper(), 0.0);
    // conformal status
    conformal.assign(size_t(nq) * 2 * m.dim, 0);
    for (int o = 0; o < nq; ++o)
      for (int d = 0; d < m.dim; ++d)
        for (int side = 0; side < 2; ++side)
        {
          int c[3] = { lay.b / 2, lay.b / 2, lay.b / 2 };
          c[d] = side == 0 ? -1 : lay.b;
          Loc l = m.locate(o, c[0], c[1], c[2]);
          conformal[size_t(o) * 2 * m.dim + 2 * d + side] = (l.oct >= 0 && l.rel == 1) ? 1 : 0;
        }
  }

  // ------------------- accessors -------------------
  double &
  q(int i, int j, int k, int v, int o)
  {
    return Q[size_t(o * 11 + v) * lay.gcells(2) + lay.gidx(2, i, j, k)];
  }
  double &
  q2(int i, int j, int k, int v, int o)
  {
    return Q2[size_t(o * 8 + v) * lay.gcells(1) + lay.gidx(1, i, j, k)];
  }
  double &
  sl(int dir, int i, int j, int k, int v, int o)
  {
    std::vector<double> & S = dir == IX ? Sx : (dir == IY ? Sy : Sz);
    return S[size_t(o * 8 + v) * lay.gcells(1) + lay.gidx(1, i, j, k)];
  }
  double &
  ef(int i, int j, int k, int c, int o)
  {
    const int nc = m.dim == 3 ? 3 : 1;
    return E[size_t(o * nc + c) * lay.gcells(2) + lay.gidx(2, i, j, k)];
  }
  double &
  sfm(int i, int j, int k, int c, int o)
  {
    const int nc = m.dim == 3 ? 3 : 2;
    return sF[size_t(o * nc + c) * lay.gcells(1) + lay.gidx(1, i, j, k)];
  }
  long
  fl_cells() const
  {
    return long(flux_shape[0]) * flux_shape[1] * (m.dim == 3 ? flux_shape[2] : 1);
  }
  double &
  fl(int i, int j, int k, int v, int o)
  {
    const long idx = i + long(flux_shape[0]) * (j + long(flux_shape[1]) * k);
    return Fl[size_t(o * 8 + v) * fl_cells() + idx];
  }
  long
  emf_cells() const
  {
    const long n = lay.b + 1;
    return m.dim == 3 ? n * n * n : n * n;
  }
  double &
  em(std::vector<double> & A, int i, int j, int k, int c, int o)
  {
    const int  nc = m.dim == 3 ? 3 : 1;
    const long n = lay.b + 1;
    return A[size_t(o * nc + c) * emf_cells() + i + n * (j + n * k)];
  }
  double &
  bg(int c, int i, int j, int k, int o)
  {
    return Bg[size_t(o) * lay.gfper() + lay.gfidx(c, i, j, k)];
  }

  // -------------------------------------------------------------------
  // FillBlockGhostFacesFunctor (kalypsso::core, called at GodunovImplemV0.cpp:224-235),
  // restated: face B with 2 ghost layers.  Interior faces are the block's own; a ghost face
  // is read from the leaf that holds one of the two cells it separates (see ghost_face for the
  // choice when the face lies on a leaf boundary): same level -> copy; coarser -> the coarse
  // face when the fine face lies on a
  // coarse face plane, else the mean of the two coarse faces bracketing it (piecewise
  // constant across, linear along: keeps the ghost cells divergence free); finer -> mean of
  // the 2^(dim-1) fine faces covering it.  2D Bz is cell centred: copy / coarse value / mean
  // of the 4 fine cells.
  // -------------------------------------------------------------------
  double
  face_value(const double * B, int o, int c, int i, int j, int k) const
  {
    return B[size_t(o) * lay.fper() + lay.fidx(c, i, j, k)];
  }

  double
  ghost_face(const double * B, int o, int c, int f[3])
  {
    const int b = lay.b;
    const int dim = m.dim;
    if (dim == 2 && c == IZ)
    {
      // cell-centred component
      Loc l = m.locate(o, f[0], f[1], 0);
      if (l.oct < 0)
        throw std::runtime_error("face ghost: neighbour not found");
      if (l.rel == 0 || l.rel == -1)
        return face_value(B, l.oct, IZ, l.ijk[0], l.ijk[1], 0);
      double sum = 0;
      for (int dj = 0; dj < 2; ++dj)
        for (int di = 0; di < 2; ++di)
          sum += face_value(B, l.oct, IZ, l.ijk[0] + di, l.ijk[1] + dj, 0);
      return sum / 4.0;
    }
    // The face separates cell lo (index f[c]-1 along c) from cell hi (index f[c]).  Beyond the
    // block along c the nearer of the two is used (both lie in one leaf, or the nearer one is
    // the finer); beside the block (0 <= f[c] <= b) the face may lie on the boundary between two
    // leaves: the value of the leaf at this block's level is taken, else the finer leaf's (ties:
    // hi).  Same-level neighbours then read identical face values, so the fluxes and EMFs they
    // both compute at a shared face or edge are bitwise equal.
    int  cc[3] = { f[0], f[1], f[2] };
    bool high_side = true; // face is the low face of cell cc
    Loc  l;
    if (f[c] > b)
    {
      cc[c] = f[c] - 1;
      high_side = false;
      l = m.locate(o, cc[0], cc[1], cc[2]);
    }
    else if (f[c] < 0)
    {
      l = m.locate(o, cc[0], cc[1], cc[2]);
    }
    else
    {
      Loc lh = m.locate(o, cc[0], cc[1], cc[2]);
      int lo[3] = { f[0], f[1], f[2] };
      lo[c] -= 1;
      Loc ll = m.locate(o, lo[0], lo[1], lo[2]);
      // preference: same level as this block, then finer, then coarser
      auto rank = [](int rel) { return rel == 0 ? 2 : (rel > 0 ? 1 : 0); };
      if (ll.oct >= 0 && lh.oct >= 0 && rank(ll.rel) > rank(lh.rel))
      {
        l = ll;
        cc[c] = lo[c];
        high_side = false;
      }
      else
        l = lh;
    }
    if (l.oct < 0)
      throw std::runtime_error("face ghost: neighbour not found");
    int fl_[3] = { l.ijk[0], l.ijk[1], l.ijk[2] };
    if (l.rel == 0)
    {
      if (!high_side)
        fl_[c] += 1;
      return face_value(B, l.oct, c, fl_[0], fl_[1], fl_[2]);
    }
    if (l.rel == -1)
    {
      // fine face index along c (global, level L) and its coarse position
      const int L = m.level[o];
      long      F = long(m.coord[3 * o + c]) * b + f[c];
      const long n = m.ncells(L, c);
      if (m.periodic[c])
        F = ((F % n) + n) % n;
      const long Sc = long(m.coord[3 * l.oct + c]) * b; // coarse block origin (coarse units)
      if (F % 2 == 0)
      {
        int fc[3] = { l.ijk[0], l.ijk[1], l.ijk[2] };
        long lc = F / 2 - Sc;
        if (lc < 0)
          lc += m.ncells(L - 1, c);
        if (lc > b)
          lc -= m.ncells(L - 1, c);
        fc[c] = int(lc);
        return face_value(B, l.oct, c, fc[0], fc[1], fc[2]);
      }
      int fc0[3] = { l.ijk[0], l.ijk[1], l.ijk[2] };
      int fc1[3] = { l.ijk[0], l.ijk[1], l.ijk[2] };
      long lc = (F - 1) / 2 - Sc;
      if (lc < 0)
        lc += m.ncells(L - 1, c);
      if (lc > b)
        lc -= m.ncells(L - 1, c);
      fc0[c] = int(lc);
      fc1[c] = int(lc) + 1;
      return 0.5 * (face_value(B, l.oct, c, fc0[0], fc0[1], fc0[2]) +
                    face_value(B, l.oct, c, fc1[0], fc1[1], fc1[2]));
    }
    // finer: the coarse face is fine face plane 2F; l.ijk is the lower fine cell of cell cc
    int base[3] = { l.ijk[0], l.ijk[1], l.ijk[2] };
    if (!high_side)
      base[c] += 2; // high face of the coarse cell = high face of fine cell +1
    int t1 = (c + 1) % 3, t2 = (c + 2) % 3;
    if (t1 > t2)
    {
      int tmp = t1;
      t1 = t2;
      t2 = tmp;
    }
    if (dim == 2)
    {
      const int t = c == IX ? IY : IX;
      double    sum = 0;
      for (int a = 0; a < 2; ++a)
      {
        int ff[3] = { base[0], base[1], 0 };
        ff[t] += a;
        sum += face_value(B, l.oct, c, ff[0], ff[1], 0);
      }
      return sum / 2.0;
    }
    double sum = 0;
    for (int a2 = 0; a2 < 2; ++a2)
      for (int a1 = 0; a1 < 2; ++a1)
      {
        int ff[3] = { base[0], base[1], base[2] };
        ff[t1] += a1;
        ff[t2] += a2;
        sum += face_value(B, l.oct, c, ff[0], ff[1], ff[2]);
      }
    return sum / 4.0;
  }

  void
  fill_Bface_ghosted(const double * B)
  {
    const int b = lay.b, dim = m.dim;
    const int n_oct = m.nowned;
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int o = 0; o < n_oct; ++o)
    {
      for (int c = 0; c < 3; ++c)
      {
        const int nx = (c == IX) ? b + 5 : b + 4;
        const int ny = (c == IY) ? b + 5 : b + 4;
        const int nz = dim == 3 ? ((c == IZ) ? b + 5 : b + 4) : 1;
        for (int kk = 0; kk < nz; ++kk)
          for (int jj = 0; jj < ny; ++jj)
            for (int ii = 0; ii < nx; ++ii)
            {
              int f[3] = { ii - 2, jj - 2, dim == 3 ? kk - 2 : 0 };
              if (dim == 2 && c == IZ)
              {
                if (ii == b + 4 || jj == b + 4)
                  continue;
              }
              bool inner = true;
              for (int d = 0; d < dim; ++d)
              {
                const int hi = (d == c && !(dim == 2 && c == IZ)) ? b : b - 1;
                if (f[d] < 0 || f[d] > hi)
                  inner = false;
              }
              double v;
              if (inner)
                v = face_value(B, o, c, f[0], f[1], f[2]);
              else
                v = ghost_face(B, o, c, f);
              bg(c, f[0], f[1], f[2], o) = v;
            }
      }
    }
  }

  // -------------------------------------------------------------------
  // ConvertToPrimitivesVariablesFunctor::fill_ghosts
  // (scheme/ConvertToPrimitivesVariablesFunctor.cpp:639-719), cell_centered_prolongation_type
  // SIMPLE_COPY or EXTRAPOLATE_LINEAR_MINMOD.
  // -------------------------------------------------------------------

  // The five conserved variables at coarse-level cell c of block oct, as the minmod slopes of
  // the linear ghost prolongation read them: the cell itself inside the block; beyond it, the
  // neighbour's cell (same level), the mean of its 2^dim cells (finer, x fastest, then y, then
  // z) or its containing cell (coarser).  Restates kalypsso::core
  // StencilHelper::getNeighLoc + compute_minmod_slopes (called at
  // ConvertToPrimitivesVariablesFunctor.cpp:485-508, 571-597; not in the reference tree): false
  // when the cell lies beyond an outside block, whose neighbours are not part of the mesh.
  bool
  stencil_u(const double * U, int oct, const int c[3], double * u) const
  {
    const int  b = lay.b, dim = m.dim;
    const long ncell = lay.cells();
    bool       inside = true;
    for (int d = 0; d < dim; ++d)
      inside = inside && c[d] >= 0 && c[d] < b;
    if (inside)
    {
      for (int v = 0; v < 5; ++v)
        u[v] = U[size_t(oct * 5 + v) * ncell + lay.cidx(c[0], c[1], dim == 3 ? c[2] : 0)];
      return true;
    }
    if (oct >= m.nowned)
      return false;
    const Loc l = m.locate(oct, c[0], c[1], dim == 3 ? c[2] : 0);
    if (l.oct < 0)
      return false;
    if (l.rel == 1)
    {
      for (int v = 0; v < 5; ++v)
      {
        double sum = 0;
        for (int dk = 0; dk < (dim == 3 ? 2 : 1); ++dk)
          for (int dj = 0; dj < 2; ++dj)
            for (int di = 0; di < 2; ++di)
              sum += U[size_t(l.oct * 5 + v) * ncell + lay.cidx(l.ijk[0] + di, l.ijk[1] + dj, l.ijk[2] + dk)];
        u[v] = sum / double(1 << dim);
      }
    }
    else
    {
      for (int v = 0; v < 5; ++v)
        u[v] = U[size_t(l.oct * 5 + v) * ncell + lay.cidx(l.ijk[0], l.ijk[1], l.ijk[2])];
    }
    return true;
  }

  // linear_extrapolate_hydro_vars_using_limited_slopes (ConvertToPrimitivesVariablesFunctor.cpp:
  // 471-635): the conserved variables of ghost cell (i, j, k) from the coarse cell l that
  // contains it, u = u_c + 1/4 s_x du/dx + 1/4 s_y du/dy (+ 1/4 s_z du/dz), du/dx the minmod
  // slope (slope_type 1) over the coarse cell's neighbours along x (zero when one of them lies
  // beyond an outside block), s = -1 / +1 for the low / high half of the coarse cell.
  void
  extrapolate_linear_minmod(const double * U, const Loc & l, int i, int j, int k, double * u) const
  {
    const int dim = m.dim;
    const int cell[3] = { i, j, k };
    double    u0[5];
    const int c0[3] = { l.ijk[0], l.ijk[1], l.ijk[2] };
    stencil_u(U, l.oct, c0, u0);
    double slope[3][5] = {};
    for (int d = 0; d < dim; ++d)
    {
      int cp[3] = { c0[0], c0[1], c0[2] }, cm[3] = { c0[0], c0[1], c0[2] };
      cp[d] += 1;
      cm[d] -= 1;
      double up[5], um[5];
      const bool okp = stencil_u(U, l.oct, cp, up);
      const bool okm = stencil_u(U, l.oct, cm, um);
      for (int v = 0; v < 5; ++v)
        slope[d][v] = (okp && okm) ? slope_unsplit_scalar(u0[v], up[v], um[v], 1.0) : 0.0;
    }
    for (int v = 0; v < 5; ++v)
    {
      double x = u0[v];
      for (int d = 0; d < dim; ++d)
      {
        const int sgn = 2 * (((cell[d] % 2) + 2) % 2) - 1;
        x = x + 0.25 * double(sgn) * slope[d][v];
      }
      u[v] = x;
    }
  }

  bool
  convert_to_primitives(const double * U, const double * B)
  {
    const int  b = lay.b, dim = m.dim;
    const long nc = lay.gcells(2);
    const long total = long(m.nowned) * nc;
    const long ncell = lay.cells();
    bool       ok = true;
#pragma omp parallel for schedule(static) num_threads(nthreads) reduction(&& : ok)
    for (long gi = 0; gi < total; ++gi)
    {
      const int  o = int(gi / nc);
      const long ci = gi - long(o) * nc;
      const int  n = b + 4;
      int        i = int(ci % n) - 2;
      int        j = int((ci / n) % n) - 2;
      int        k = dim == 3 ? int(ci / (long(n) * n)) - 2 : 0;
      const bool inside = (i >= 0 && i < b && j >= 0 && j < b && (dim == 2 || (k >= 0 && k < b)));
      double     u[8];
      double     fm[6];
      if (inside)
      {
        const long cidx = lay.cidx(i, j, k);
        for (int v = 0; v < 5; ++v)
          u[v] = U[size_t(o * 5 + v) * ncell + cidx];
        if (dim == 2)
        {
          u[IA] = 0.5 * (bg(IX, i, j, 0, o) + bg(IX, i + 1, j, 0, o));
          u[IB] = 0.5 * (bg(IY, i, j, 0, o) + bg(IY, i, j + 1, 0, o));
          u[IC] = 0.5 * (bg(IZ, i, j, 0, o) + bg(IZ, i, j, 0, o));
        }
        else
        {
          u[IA] = 0.5 * (bg(IX, i, j, k, o) + bg(IX, i + 1, j, k, o));
          u[IB] = 0.5 * (bg(IY, i, j, k, o) + bg(IY, i, j + 1, k, o));
          u[IC] = 0.5 * (bg(IZ, i, j, k, o) + bg(IZ, i, j, k + 1, o));
        }
      }
      else
      {
        Loc l = m.locate(o, i, j, k);
        if (l.oct < 0)
        {
          ok = false;
          continue;
        }
        if (l.rel == -1 && s.prolongation == 1)
        {
          // coarser neighbour, EXTRAPOLATE_LINEAR_MINMOD: the extrapolated conserved variables
          // and the cell-centred field of this block's ghosted faces (:536-541, 623-628)
          extrapolate_linear_minmod(U, l, i, j, k, u);
          if (dim == 2)
          {
            u[IA] = 0.5 * (bg(IX, i, j, 0, o) + bg(IX, i + 1, j, 0, o));
            u[IB] = 0.5 * (bg(IY, i, j, 0, o) + bg(IY, i, j + 1, 0, o));
            u[IC] = bg(IZ, i, j, 0, o);
          }
          else
          {
            u[IA] = 0.5 * (bg(IX, i, j, k, o) + bg(IX, i + 1, j, k, o));
            u[IB] = 0.5 * (bg(IY, i, j, k, o) + bg(IY, i, j + 1, k, o));
            u[IC] = 0.5 * (bg(IZ, i, j, k, o) + bg(IZ, i, j, k + 1, o));
          }
        }
        else if (l.rel == 1)
        {
          // restriction: siblings average (x fastest, then y, then z)
          for (int v = 0; v < 5; ++v)
          {
            double sum = 0;
            for (int dk = 0; dk < (dim == 3 ? 2 : 1); ++dk)
              for (int dj = 0; dj < 2; ++dj)
                for (int di = 0; di < 2; ++di)
                  sum += U[size_t(l.oct * 5 + v) * ncell +
                           lay.cidx(l.ijk[0] + di, l.ijk[1] + dj, l.ijk[2] + dk)];
            u[v] = sum / double(1 << dim);
          }
          if (dim == 2)
          {
            u[IA] = 0.5 * (bg(IX, i, j, 0, o) + bg(IX, i + 1, j, 0, o));
            u[IB] = 0.5 * (bg(IY, i, j, 0, o) + bg(IY, i, j + 1, 0, o));
            u[IC] = bg(IZ, i, j, 0, o);
          }
          else
          {
            u[IA] = 0.5 * (bg(IX, i, j, k, o) + bg(IX, i + 1, j, k, o));
            u[IB] = 0.5 * (bg(IY, i, j, k, o) + bg(IY, i, j + 1, k, o));
            u[IC] = 0.5 * (bg(IZ, i, j, k, o) + bg(IZ, i, j, k + 1, o));
          }
        }
        else
        {
          // same level copy, or coarser with SIMPLE_COPY: read the neighbour's own cell
          const long cidx = lay.cidx(l.ijk[0], l.ijk[1], l.ijk[2]);
          for (int v = 0; v < 5; ++v)
            u[v] = U[size_t(l.oct * 5 + v) * ncell + cidx];
          const int ii = l.ijk[0], jj = l.ijk[1], kk = l.ijk[2];
          if (dim == 2)
          {
            u[IA] = 0.5 * (face_value(B, l.oct, IX, ii, jj, 0) + face_value(B, l.oct, IX, ii + 1, jj, 0));
            u[IB] = 0.5 * (face_value(B, l.oct, IY, ii, jj, 0) + face_value(B, l.oct, IY, ii, jj + 1, 0));
            u[IC] = 0.5 * (face_value(B, l.oct, IZ, ii, jj, 0) + face_value(B, l.oct, IZ, ii, jj, 0));
          }
          else
          {
            u[IA] = 0.5 * (face_value(B, l.oct, IX, ii, jj, kk) + face_value(B, l.oct, IX, ii + 1, jj, kk));
            u[IB] = 0.5 * (face_value(B, l.oct, IY, ii, jj, kk) + face_value(B, l.oct, IY, ii, jj + 1, kk));
            u[IC] = 0.5 * (face_value(B, l.oct, IZ, ii, jj, kk) + face_value(B, l.oct, IZ, ii, jj, kk + 1));
          }
        }
      }
      // face-centred field always from this block's ghosted face array
      if (dim == 2)
      {
        fm[AL] = bg(IX, i, j, 0, o);
        fm[AR] = bg(IX, i + 1, j, 0, o);
        fm[BL] = bg(IY, i, j, 0, o);
        fm[BR] = bg(IY, i, j + 1, 0, o);
        fm[CL] = bg(IZ, i, j, 0, o);
        fm[CR] = fm[CL];
      }
      else
      {
        fm[AL] = bg(IX, i, j, k, o);
        fm[AR] = bg(IX, i + 1, j, k, o);
        fm[BL] = bg(IY, i, j, k, o);
        fm[BR] = bg(IY, i, j + 1, k, o);
        fm[CL] = bg(IZ, i, j, k, o);
        fm[CR] = bg(IZ, i, j, k + 1, o);
      }
      double qq[8];
      if (!compute_primitives(u, qq, s))
        ok = false;
      q(i, j, k, ID, o) = qq[ID];
      q(i, j, k, IP, o) = qq[IP];
      q(i, j, k, IU, o) = qq[IU];
      q(i, j, k, IV, o) = qq[IV];
      q(i, j, k, IW, o) = qq[IW];
      q(i, j, k, IAL, o) = fm[AL];
      q(i, j, k, IAR, o) = fm[AR];
      q(i, j, k, IBL, o) = fm[BL];
      q(i, j, k, IBR, o) = fm[BR];
      q(i, j, k, ICL, o) = fm[CL];
      q(i, j, k, ICR, o) = fm[CR];
    }
    return ok;
  }

  // ComputeLimitedSlopesFunctor::slope_unsplit_scalar_mag (.cpp:148-221)
  double
  slope_mag(int i, int j, int k, int dir, int v1, int v2, int o)
  {
    int uv[3] = { 0, 0, 0 };
    uv[dir] = 1;
    if ((v1 == IAL && v2 == IAR && dir == IX) || (v1 == IBL && v2 == IBR && dir == IY) ||
        (m.dim == 3 && v1 == ICL && v2 == ICR && dir == IZ))
      return q(i, j, k, v2, o) - q(i, j, k, v1, o);
    const double datac = (q(i, j, k, v1, o) + q(i, j, k, v2, o)) * 0.5;
    const double datap =
      (q(i + uv[0], j + uv[1], k + uv[2], v1, o) + q(i + uv[0], j + uv[1], k + uv[2], v2, o)) * 0.5;
    const double datam =
      (q(i - uv[0], j - uv[1], k - uv[2], v1, o) + q(i - uv[0], j - uv[1], k - uv[2], v2, o)) * 0.5;
    return slope_unsplit_scalar(datac, datap, datam, s.slope_type);
  }

  // ComputeLimitedSlopesFunctor::compute_limited_slopes (.cpp:225-314)
  void
  compute_limited_slopes()
  {
    const int  b = lay.b, dim = m.dim;
    const long nc = lay.gcells(1);
    const long total = long(m.nowned) * nc;
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (long gi = 0; gi < total; ++gi)
    {
      const int  o = int(gi / nc);
      const long ci = gi - long(o) * nc;
      const int  n = b + 2;
      int        i = int(ci % n) - 1;
      int        j = int((ci / n) % n) - 1;
      int        k = dim == 3 ? int(ci / (long(n) * n)) - 1 : 0;
      const int  nh = dim == 3 ? 5 : 4; // nbvar_hydro_only<dim>() (models/MHDState.h:35-42)
      for (int v = 0; v < nh; ++v)
      {
        sl(IX, i, j, k, v, o) =
          slope_unsplit_scalar(q(i, j, k, v, o), q(i + 1, j, k, v, o), q(i - 1, j, k, v, o), s.slope_type);
        sl(IY, i, j, k, v, o) =
          slope_unsplit_scalar(q(i, j, k, v, o), q(i, j + 1, k, v, o), q(i, j - 1, k, v, o), s.slope_type);
        if (dim == 3)
          sl(IZ, i, j, k, v, o) = slope_unsplit_scalar(
            q(i, j, k, v, o), q(i, j, k + 1, v, o), q(i, j, k - 1, v, o), s.slope_type);
      }
      for (int d = 0; d < dim; ++d)
      {
        sl(d, i, j, k, IA, o) = slope_mag(i, j, k, d, IAL, IAR, o);
        sl(d, i, j, k, IB, o) = slope_mag(i, j, k, d, IBL, IBR, o);
        sl(d, i, j, k, IC, o) = slope_mag(i, j, k, d, ICL, ICR, o);
      }
    }
  }

  // ComputeUpdatedPrimitiveVariablesFunctor::update_primitive_variables (.cpp:284-546).
  // HANCOCK: the half-step predictor, with the gravity kick and the viscous term after the
  // source terms; RK2_SSP: q2 = the cell-centred state (.cpp:396-407, 531-542).
  void
  compute_primitives_predictor(double dt)
  {
    const int  b = lay.b, dim = m.dim;
    const long nc = lay.gcells(1);
    const long total = long(m.nowned) * nc;
    const double gamma = s.gamma0;
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (long gi = 0; gi < total; ++gi)
    {
      const int  o = int(gi / nc);
      const long ci = gi - long(o) * nc;
      const int  n = b + 2;
      int        i = int(ci % n) - 1;
      int        j = int((ci / n) % n) - 1;
      int        k = dim == 3 ? int(ci / (long(n) * n)) - 1 : 0;
      const double dx = dx_of(o);
      const double dtdx = dt / dx, dtdy = dt / dx, dtdz = dt / dx;
      const double r = q(i, j, k, ID, o);
      const double p = q(i, j, k, IP, o);
      double       u = q(i, j, k, IU, o);
      double       v = q(i, j, k, IV, o);
      double       w = q(i, j, k, IW, o);
      const double A = (q(i, j, k, IAL, o) + q(i, j, k, IAR, o)) * 0.5;
      const double B = (q(i, j, k, IBL, o) + q(i, j, k, IBR, o)) * 0.5;
      const double C = (q(i, j, k, ICL, o) + q(i, j, k, ICR, o)) * 0.5;
      if (s.time_integrator == 1)
      {
        q2(i, j, k, ID, o) = r;
        q2(i, j, k, IU, o) = u;
        q2(i, j, k, IV, o) = v;
        q2(i, j, k, IW, o) = w;
        q2(i, j, k, IP, o) = p;
        q2(i, j, k, IA, o) = A;
        q2(i, j, k, IB, o) = B;
        q2(i, j, k, IC, o) = C;
        continue;
      }
      const double drx = sl(IX, i, j, k, ID, o), dpx = sl(IX, i, j, k, IP, o);
      const double dux = sl(IX, i, j, k, IU, o), dvx = sl(IX, i, j, k, IV, o);
      const double dwx = sl(IX, i, j, k, IW, o), dAx = sl(IX, i, j, k, IA, o);
      const double dBx = sl(IX, i, j, k, IB, o), dCx = sl(IX, i, j, k, IC, o);
      const double dry = sl(IY, i, j, k, ID, o), dpy = sl(IY, i, j, k, IP, o);
      const double duy = sl(IY, i, j, k, IU, o), dvy = sl(IY, i, j, k, IV, o);
      const double dwy = sl(IY, i, j, k, IW, o), dAy = sl(IY, i, j, k, IA, o);
      const double dBy = sl(IY, i, j, k, IB, o), dCy = sl(IY, i, j, k, IC, o);
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
        const double drz = sl(IZ, i, j, k, ID, o), dpz = sl(IZ, i, j, k, IP, o);
        const double duz = sl(IZ, i, j, k, IU, o), dvz = sl(IZ, i, j, k, IV, o);
        const double dwz = sl(IZ, i, j, k, IW, o), dAz = sl(IZ, i, j, k, IA, o);
        const double dBz = sl(IZ, i, j, k, IB, o), dCz = sl(IZ, i, j, k, IC, o);
        sr0 = (-u * drx - dux * r) * dtdx + (-v * dry - dvy * r) * dtdy + (-w * drz - dwz * r) * dtdz;
        su0 = (-u * dux - (dpx + B * dBx + C * dCx) / r) * dtdx + (-v * duy + B * dAy / r) * dtdy +
              (-w * duz + C * dAz / r) * dtdz;
        sv0 = (-u * dvx + A * dBx / r) * dtdx + (-v * dvy - (dpy + A * dAy + C * dCy) / r) * dtdy +
              (-w * dvz + C * dBz / r) * dtdz;
        sw0 = (-u * dwx + A * dCx / r) * dtdx + (-v * dwy + B * dCy / r) * dtdy +
              (-w * dwz - (dpz + A * dAz + B * dBz) / r) * dtdz;
        sp0 = (-u * dpx - dux * gamma * p) * dtdx + (-v * dpy - dvy * gamma * p) * dtdy +
              (-w * dpz - dwz * gamma * p) * dtdz;
        sA0 = (u * dBy + B * duy - v * dAy - A * dvy) * dtdy + (u * dCz + C * duz - w * dAz - A * dwz) * dtdz;
        sB0 = (v * dAx + A * dvx - u * dBx - B * dux) * dtdx + (v * dCz + C * dvz - w * dBz - B * dwz) * dtdz;
        sC0 = (w * dAx + A * dwx - u * dCx - C * dux) * dtdx + (w * dBy + B * dwy - v * dCy - C * dvy) * dtdy;
      }
      // gravity predictor, half time step (.cpp:372-377 2D: u, v; :505-511 3D: u, v, w)
      if (s.gravity_enabled)
      {
        u += s.gravity[IX] * 0.5 * dt;
        v += s.gravity[IY] * 0.5 * dt;
        if (dim == 3)
          w += s.gravity[IZ] * 0.5 * dt;
      }
      // viscous predictor, add_viscous_predictor_{2d,3d} (.cpp:178-272, called at :380-383,
      // 514-517): per velocity component (2D: u, v), per axis, half of nu times the
      // second difference over dx, times dt/dx
      if (s.viscosity_enabled && s.viscosity_predictor)
      {
        const double nu = s.mu / q(i, j, k, ID, o);
        const double dtd[3] = { dtdx, dtdy, dtdz };
        double *     vel[3] = { &u, &v, &w };
        for (int c = 0; c < dim; ++c)
          for (int a = 0; a < dim; ++a)
          {
            int e[3] = { 0, 0, 0 };
            e[a] = 1;
            const double d2 = q(i + e[0], j + e[1], k + e[2], IU + c, o) + q(i - e[0], j - e[1], k - e[2], IU + c, o) -
                              2 * q(i, j, k, IU + c, o);
            *vel[c] += 0.5 * nu * d2 / dx * dtd[a];
          }
      }
      q2(i, j, k, ID, o) = r + 0.5 * sr0;
      q2(i, j, k, IU, o) = u + 0.5 * su0;
      q2(i, j, k, IV, o) = v + 0.5 * sv0;
      q2(i, j, k, IW, o) = w + 0.5 * sw0;
      q2(i, j, k, IP, o) = p + 0.5 * sp0;
      q2(i, j, k, IA, o) = A + 0.5 * sA0;
      q2(i, j, k, IB, o) = B + 0.5 * sB0;
      q2(i, j, k, IC, o) = C + 0.5 * sC0;
    }
  }

  // ComputeElectricFieldFunctor::compute_elec_field (.cpp:97-235)
  void
  compute_elec_field()
  {
    const int  b = lay.b, dim = m.dim;
    const long nc = lay.gcells(2);
    const long total = long(m.nowned) * nc;
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (long gi = 0; gi < total; ++gi)
    {
      const int  o = int(gi / nc);
      const long ci = gi - long(o) * nc;
      const int  n = b + 4;
      int        i = int(ci % n) - 2;
      int        j = int((ci / n) % n) - 2;
      int        k = dim == 3 ? int(ci / (long(n) * n)) - 2 : 0;
      if (dim == 2)
      {
        if (i > -2 && j > -2)
        {
          const double u = 0.25 * (q(i - 1, j - 1, 0, IU, o) + q(i - 1, j, 0, IU, o) +
                                   q(i, j - 1, 0, IU, o) + q(i, j, 0, IU, o));
          const double v = 0.25 * (q(i - 1, j - 1, 0, IV, o) + q(i - 1, j, 0, IV, o) +
                                   q(i, j - 1, 0, IV, o) + q(i, j, 0, IV, o));
          const double A = 0.5 * (q(i, j - 1, 0, IAL, o) + q(i, j, 0, IAL, o));
          const double B = 0.5 * (q(i - 1, j, 0, IBL, o) + q(i, j, 0, IBL, o));
          ef(i, j, 0, 0, o) = u * B - v * A;
        }
      }
      else if (i > -2 && j > -2 && k > -2)
      {
        {
          const double v = 0.25 * (q(i, j - 1, k - 1, IV, o) + q(i, j - 1, k, IV, o) +
                                   q(i, j, k - 1, IV, o) + q(i, j, k, IV, o));
          const double w = 0.25 * (q(i, j - 1, k - 1, IW, o) + q(i, j - 1, k, IW, o) +
                                   q(i, j, k - 1, IW, o) + q(i, j, k, IW, o));
          const double B = 0.5 * (q(i, j, k - 1, IBL, o) + q(i, j, k, IBL, o));
          const double C = 0.5 * (q(i, j - 1, k, ICL, o) + q(i, j, k, ICL, o));
          ef(i, j, k, IX, o) = v * C - w * B;
        }
        {
          const double u = 0.25 * (q(i - 1, j, k - 1, IU, o) + q(i - 1, j, k, IU, o) +
                                   q(i, j, k - 1, IU, o) + q(i, j, k, IU, o));
          const double w = 0.25 * (q(i - 1, j, k - 1, IW, o) + q(i - 1, j, k, IW, o) +
                                   q(i, j, k - 1, IW, o) + q(i, j, k, IW, o));
          const double A = 0.5 * (q(i, j, k - 1, IAL, o) + q(i, j, k, IAL, o));
          const double C = 0.5 * (q(i - 1, j, k, ICL, o) + q(i, j, k, ICL, o));
          ef(i, j, k, IY, o) = w * A - u * C;
        }
        {
          const double u = 0.25 * (q(i - 1, j - 1, k, IU, o) + q(i - 1, j, k, IU, o) +
                                   q(i, j - 1, k, IU, o) + q(i, j, k, IU, o));
          const double v = 0.25 * (q(i - 1, j - 1, k, IV, o) + q(i - 1, j, k, IV, o) +
                                   q(i, j - 1, k, IV, o) + q(i, j, k, IV, o));
          const double A = 0.5 * (q(i, j - 1, k, IAL, o) + q(i, j, k, IAL, o));
          const double B = 0.5 * (q(i - 1, j, k, IBL, o) + q(i, j, k, IBL, o));
          ef(i, j, k, IZ, o) = u * B - v * A;
        }
      }
    }
  }

  // ComputeSourceFaceMagFunctor::compute_sFaceMag (.cpp:114-203)
  void
  compute_sFaceMag(double dt)
  {
    const int  b = lay.b, dim = m.dim;
    const long nc = lay.gcells(1);
    const long total = long(m.nowned) * nc;
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (long gi = 0; gi < total; ++gi)
    {
      const int  o = int(gi / nc);
      const long ci = gi - long(o) * nc;
      const int  n = b + 2;
      int        i = int(ci % n) - 1;
      int        j = int((ci / n) % n) - 1;
      int        k = dim == 3 ? int(ci / (long(n) * n)) - 1 : 0;
      const double dx = dx_of(o);
      const double dtdx = dt / dx, dtdy = dt / dx, dtdz = dt / dx;
      if (dim == 2)
      {
        sfm(i, j, 0, IX, o) = (ef(i, j + 1, 0, 0, o) - ef(i, j, 0, 0, o)) * 0.5 * dtdy;
        sfm(i, j, 0, IY, o) = -(ef(i + 1, j, 0, 0, o) - ef(i, j, 0, 0, o)) * 0.5 * dtdx;
      }
      else
      {
        sfm(i, j, k, IX, o) = (ef(i, j + 1, k, IZ, o) - ef(i, j, k, IZ, o)) * 0.5 * dtdy -
                              (ef(i, j, k + 1, IY, o) - ef(i, j, k, IY, o)) * 0.5 * dtdz;
        sfm(i, j, k, IY, o) = -(ef(i + 1, j, k, IZ, o) - ef(i, j, k, IZ, o)) * 0.5 * dtdx +
                              (ef(i, j, k + 1, IX, o) - ef(i, j, k, IX, o)) * 0.5 * dtdz;
        sfm(i, j, k, IZ, o) = (ef(i + 1, j, k, IY, o) - ef(i, j, k, IY, o)) * 0.5 * dtdx -
                              (ef(i, j + 1, k, IX, o) - ef(i, j, k, IX, o)) * 0.5 * dtdy;
      }
    }
  }

  void
  get_state(int which, int i, int j, int k, int o, double * qs)
  {
    // which: 0 = q2, 1+d = slopes along d
    for (int v = 0; v < 8; ++v)
      qs[v] = which == 0 ? q2(i, j, k, v, o) : sl(which - 1, i, j, k, v, o);
  }

  // ComputeHydroFluxesAndStoreFunctor::reconstruct_state_{2d,3d} (.cpp:119-214)
  void
  reconstruct_state(const double * qc, int is, int js, int ks, int o, int dir, bool left, double * qr)
  {
    const double offset = left ? -0.5 : 0.5;
    double       dq[8];
    get_state(1 + dir, is, js, ks, o, dq);
    for (int v = 0; v < 8; ++v)
      qr[v] = qc[v] + offset * dq[v];
    qr[ID] = std::fmax(s.smallr, qr[ID]);
    qr[IP] = std::fmax(s.smallp * qr[ID], qr[IP]);
  }

  // ComputeHydroFluxesAndStoreFunctor::compute_fluxes_and_store_{2d,3d} (.cpp:218-503)
  void
  compute_fluxes_and_store(double dt, int dir)
  {
    const int b = lay.b, dim = m.dim;
    flux_shape[0] = flux_shape[1] = flux_shape[2] = b;
    if (dim == 2)
      flux_shape[2] = 1;
    flux_shape[dir] = b + 1;
    const long nc = fl_cells();
    const long total = long(m.nowned) * nc;
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (long gi = 0; gi < total; ++gi)
    {
      const int  o = int(gi / nc);
      const long ci = gi - long(o) * nc;
      int        i = int(ci % flux_shape[0]);
      int        j = int((ci / flux_shape[0]) % flux_shape[1]);
      int        k = dim == 3 ? int(ci / (long(flux_shape[0]) * flux_shape[1])) : 0;
      const double dx = dx_of(o);
      const double dtdS = dt / dx;
      double       qprim[8], qprim_n[8], qL[8], qR[8], flux[8];
      get_state(0, i, j, k, o, qprim);
      int sh[3] = { 0, 0, 0 };
      sh[dir] = 1;
      get_state(0, i - sh[0], j - sh[1], k - sh[2], o, qprim_n);
      reconstruct_state(qprim_n, i - sh[0], j - sh[1], k - sh[2], o, dir, false, qL);
      reconstruct_state(qprim, i, j, k, o, dir, true, qR);
      const int    vnorm = dir == IX ? IAL : (dir == IY ? IBL : ICL);
      const int    bn = dir == IX ? IA : (dir == IY ? IB : IC);
      const double BN = q(i, j, k, vnorm, o) + sfm(i, j, k, dir, o);
      qR[bn] = BN;
      qL[bn] = BN;
      if (dir == IY)
      {
        std::swap(qL[IU], qL[IV]);
        std::swap(qR[IU], qR[IV]);
        std::swap(qL[IA], qL[IB]);
        std::swap(qR[IA], qR[IB]);
      }
      else if (dir == IZ)
      {
        std::swap(qL[IU], qL[IW]);
        std::swap(qR[IU], qR[IW]);
        std::swap(qL[IA], qL[IC]);
        std::swap(qR[IA], qR[IC]);
      }
      riemann_hlld(qL, qR, flux, s);
      if (dir == IY)
      {
        std::swap(flux[IU], flux[IV]);
        std::swap(flux[IA], flux[IB]);
      }
      else if (dir == IZ)
      {
        std::swap(flux[IU], flux[IW]);
        std::swap(flux[IA], flux[IC]);
      }
      for (int v = 0; v < 8; ++v)
        flux[v] = flux[v] * dtdS;
      // set_flux: 2D stores ID,IP,IU,IV,IW,IC; 3D all 8 (.h:297-339)
      for (int v = 0; v < 8; ++v)
      {
        if (dim == 2 && (v == IA || v == IB))
          continue;
        fl(i, j, k, v, o) = flux[v];
      }
    }
  }

  // ComputeViscousFluxesAndStoreFunctor (scheme/ComputeViscousFluxesAndStoreFunctor.cpp:96-352):
  // the viscous flux at the low face of flux cell (i,j,k) along dir, from the primitive
  // velocities of Q (not the predicted Q2).  Every gradient is the backward difference at cell
  // (i,j,k) along each axis (.cpp:96-163); the stress is mu (grad v + grad v^T) minus its trace
  // part (1/2 in 2D, 1/3 in 3D); the energy flux uses the velocity of cell (i,j,k).  Stored
  // times dt/dx in the flux array; density (and 2D Bz) flux 0 (set_flux, .h:208-259).
  void
  compute_viscous_fluxes_and_store(double dt, int dir)
  {
    const int b = lay.b, dim = m.dim;
    flux_shape[0] = flux_shape[1] = flux_shape[2] = b;
    if (dim == 2)
      flux_shape[2] = 1;
    flux_shape[dir] = b + 1;
    const long   nc = fl_cells();
    const long   total = long(m.nowned) * nc;
    const double mu = s.mu;
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (long gi = 0; gi < total; ++gi)
    {
      const int  o = int(gi / nc);
      const long ci = gi - long(o) * nc;
      const int  i = int(ci % flux_shape[0]);
      const int  j = int((ci / flux_shape[0]) % flux_shape[1]);
      const int  k = dim == 3 ? int(ci / (long(flux_shape[0]) * flux_shape[1])) : 0;
      const double dx = dx_of(o);
      const double dtdS = dt / dx;
      // g[c][a] = d v_c / d x_a
      double g[3][3] = { { 0, 0, 0 }, { 0, 0, 0 }, { 0, 0, 0 } };
      double vel[3] = { q(i, j, k, IU, o), q(i, j, k, IV, o), dim == 3 ? q(i, j, k, IW, o) : 0.0 };
      for (int a = 0; a < dim; ++a)
      {
        int e[3] = { 0, 0, 0 };
        e[a] = 1;
        for (int c = 0; c < dim; ++c)
          g[c][a] = (q(i, j, k, IU + c, o) - q(i - e[0], j - e[1], k - e[2], IU + c, o)) / dx;
      }
      double fU = 0, fV = 0, fW = 0, fE = 0;
      if (dim == 2)
      {
        const double dv = g[0][0] + g[1][1];
        if (dir == IX)
        {
          const double txx = 2 * mu * (g[0][0] - 0.5 * dv);
          const double txy = mu * (g[1][0] + g[0][1]);
          fU = -txx;
          fV = -txy;
          fE = -vel[0] * txx - vel[1] * txy;
        }
        else
        {
          const double tyy = 2 * mu * (g[1][1] - 0.5 * dv);
          const double txy = mu * (g[1][0] + g[0][1]);
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
        if (dir == IX)
        {
          t0 = 2 * mu * (g[0][0] - third * dv);
          t1 = mu * (g[1][0] + g[0][1]);
          t2 = mu * (g[2][0] + g[0][2]);
        }
        else if (dir == IY)
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
      double flux[8] = { 0.0, fE, fU, fV, fW, 0.0, 0.0, 0.0 };
      for (int v = 0; v < 8; ++v)
      {
        if (dim == 2 && (v == IA || v == IB))
          continue;
        fl(i, j, k, v, o) = flux[v] * dtdS;
      }
    }
  }

  // ReadFluxesAndConservativeUpdateFunctor::get_flux_from_fine_neighbor (.h:232-263):
  // sum over the 2^(dim-1) fine faces covering the coarse face (first transverse index
  // fastest), read on the fine block's right (use_right) or left face.
  void
  flux_from_fine(int o, int dir, int i, int j, int k, int side, double * fx)
  {
    const int b = lay.b, dim = m.dim;
    int       c[3] = { i, j, k };
    c[dir] = side == 0 ? -1 : b;
    Loc l = m.locate(o, c[0], c[1], c[2]);
    if (l.oct < 0 || l.rel != 1)
      throw std::runtime_error("flux_from_fine: no finer neighbour");
    int base[3] = { l.ijk[0], l.ijk[1], l.ijk[2] };
    // right flux of the fine block = face index b (side 0, neighbour on the left);
    // left flux = face index 0 (side 1, neighbour on the right)
    base[dir] = side == 0 ? b : 0;
    int t[2], nt = 0;
    for (int d = 0; d < dim; ++d)
      if (d != dir)
        t[nt++] = d;
    const int nv = 8;
    for (int v = 0; v < nv; ++v)
      fx[v] = 0;
    const int vars2d[6] = { ID, IP, IU, IV, IW, IC };
    const int nvars = dim == 2 ? 6 : 5;
    for (int vi = 0; vi < nvars; ++vi)
    {
      const int v = dim == 2 ? vars2d[vi] : vi;
      double    sum = 0;
      if (dim == 2)
      {
        for (int a = 0; a < 2; ++a)
        {
          int f[3] = { base[0], base[1], 0 };
          f[t[0]] += a;
          sum += fine_flux_read(l.oct, dir, f, v);
        }
      }
      else
      {
        for (int a1 = 0; a1 < 2; ++a1)
          for (int a0 = 0; a0 < 2; ++a0)
          {
            int f[3] = { base[0], base[1], base[2] };
            f[t[0]] += a0;
            f[t[1]] += a1;
            sum += fine_flux_read(l.oct, dir, f, v);
          }
      }
      fx[v] = sum;
    }
  }
  double
  fine_flux_read(int o, int dir, int f[3], int v)
  {
    (void)dir;
    return fl(f[0], f[1], f[2], v, o);
  }

  // ReadFluxesAndConservativeUpdateFunctor::read_fluxes_and_update_{2d,3d} (.cpp:105-436)
  void
  read_fluxes_and_update(double * U2, double * B2, int dir)
  {
    const int  b = lay.b, dim = m.dim;
    const long ncell = lay.cells();
    const long total = long(m.nowned) * ncell;
    const double volume_ratio = double(1 << dim);
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (long gi = 0; gi < total; ++gi)
    {
      const int  o = int(gi / ncell);
      const long ci = gi - long(o) * ncell;
      int        i = int(ci % b);
      int        j = int((ci / b) % b);
      int        k = dim == 3 ? int(ci / (long(b) * b)) : 0;
      int        c[3] = { i, j, k };
      const bool lo_finer = conformal[size_t(o) * 2 * dim + 2 * dir + 0] != 0;
      const bool hi_finer = conformal[size_t(o) * 2 * dim + 2 * dir + 1] != 0;
      double     acc[8] = { 0, 0, 0, 0, 0, 0, 0, 0 };
      const int  vars2d[6] = { ID, IP, IU, IV, IW, IC };
      const int  nvars = dim == 2 ? 6 : 5;
      if (c[dir] == 0 && lo_finer)
      {
        double fx[8];
        flux_from_fine(o, dir, i, j, k, 0, fx);
        for (int vi = 0; vi < nvars; ++vi)
        {
          const int v = dim == 2 ? vars2d[vi] : vi;
          acc[v] = acc[v] + fx[v] / volume_ratio;
        }
      }
      else
      {
        for (int vi = 0; vi < nvars; ++vi)
        {
          const int v = dim == 2 ? vars2d[vi] : vi;
          acc[v] = acc[v] + fl(i, j, k, v, o);
        }
      }
      if (c[dir] == b - 1 && hi_finer)
      {
        double fx[8];
        flux_from_fine(o, dir, i, j, k, 1, fx);
        for (int vi = 0; vi < nvars; ++vi)
        {
          const int v = dim == 2 ? vars2d[vi] : vi;
          acc[v] = acc[v] - fx[v] / volume_ratio;
        }
      }
      else
      {
        int ip[3] = { i, j, k };
        ip[dir] += 1;
        for (int vi = 0; vi < nvars; ++vi)
        {
          const int v = dim == 2 ? vars2d[vi] : vi;
          acc[v] = acc[v] - fl(ip[0], ip[1], ip[2], v, o);
        }
      }
      const long cidx = lay.cidx(i, j, k);
      for (int v = 0; v < 5; ++v)
        U2[size_t(o * 5 + v) * ncell + cidx] += acc[v];
      if (dim == 2)
        B2[size_t(o) * lay.fper() + lay.fidx(IZ, i, j, 0)] += acc[IC];
    }
  }

  // ComputeEmfAndStoreFunctor::reconstruct_state_2d_at_edge (.cpp:105-219)
  enum
  {
    LB = 0,
    RT = 1,
    RB = 2,
    LT = 3
  };
  double
  limited_slope_q(int i, int j, int k, int var, int o, int dir)
  {
    if (dir == IX)
      return slope_unsplit_scalar(q(i, j, k, var, o), q(i + 1, j, k, var, o), q(i - 1, j, k, var, o), s.slope_type);
    if (dir == IY)
      return slope_unsplit_scalar(q(i, j, k, var, o), q(i, j + 1, k, var, o), q(i, j - 1, k, var, o), s.slope_type);
    return slope_unsplit_scalar(q(i, j, k, var, o), q(i, j, k + 1, var, o), q(i, j, k - 1, var, o), s.slope_type);
  }

  void
  reconstruct_edge_2d(int is, int js, int o, int loc, double * qo)
  {
    get_state(0, is, js, 0, o, qo);
    int sign_x = 1, sign_y = 1, ia = is, ja = js, ib = is, jb = js, sign_a = 1, sign_b = 1;
    if (loc == LB)
    {
      sign_x = -1; sign_y = -1; sign_a = -1; sign_b = -1;
    }
    else if (loc == RT)
    {
      ia = is + 1; jb = js + 1;
    }
    else if (loc == RB)
    {
      ia = is + 1; sign_y = -1; sign_a = -1;
    }
    else if (loc == LT)
    {
      jb = js + 1; sign_x = -1; sign_b = -1;
    }
    double A = q(ia, ja, 0, IAL, o);
    double B = q(ib, jb, 0, IBL, o);
    if (s.time_integrator == 0) // HANCOCK only (ComputeEmfAndStoreFunctor.cpp:189-193)
    {
      A += sfm(ia, ja, 0, IX, o);
      B += sfm(ib, jb, 0, IY, o);
    }
    const double dAy = limited_slope_q(ia, ja, 0, IA, o, IY);
    const double dBx = limited_slope_q(ib, jb, 0, IB, o, IX);
    double       dqX[8], dqY[8];
    get_state(1 + IX, is, js, 0, o, dqX);
    get_state(1 + IY, is, js, 0, o, dqY);
    const double sx = double(sign_x), sy = double(sign_y);
    qo[ID] += 0.5 * (sx * dqX[ID] + sy * dqY[ID]);
    qo[IP] += 0.5 * (sx * dqX[IP] + sy * dqY[IP]);
    qo[IU] += 0.5 * (sx * dqX[IU] + sy * dqY[IU]);
    qo[IV] += 0.5 * (sx * dqX[IV] + sy * dqY[IV]);
    qo[IW] += 0.5 * (sx * dqX[IW] + sy * dqY[IW]);
    qo[IA] = A + 0.5 * (double(sign_a) * dAy);
    qo[IB] = B + 0.5 * (double(sign_b) * dBx);
    qo[IC] += 0.5 * (sx * dqX[IC] + sy * dqY[IC]);
    qo[ID] = std::fmax(s.smallr, qo[ID]);
    qo[IP] = std::fmax(s.smallp * qo[ID], qo[IP]);
  }

  // ComputeEmfAndStoreFunctor::reconstruct_state_3d_at_edge (.cpp:223-367), including its
  // cumulative IB / IA updates when the edge is along Y or X (.cpp:324-362).
  void
  reconstruct_edge_3d(int is, int js, int ks, int o, int loc, int dir0, int dir1, double * qo)
  {
    get_state(0, is, js, ks, o, qo);
    int sdq0 = 1, sdq1 = 1, sb0 = 1, sb1 = 1;
    int ijk0[3] = { is, js, ks }, ijk1[3] = { is, js, ks };
    if (loc == LB)
    {
      sdq0 = -1; sdq1 = -1; sb0 = -1; sb1 = -1;
    }
    else if (loc == RT)
    {
      ijk0[dir0] += 1; ijk1[dir1] += 1;
    }
    else if (loc == RB)
    {
      ijk0[dir0] += 1; sdq1 = -1; sb0 = -1;
    }
    else if (loc == LT)
    {
      ijk1[dir1] += 1; sdq0 = -1; sb1 = -1;
    }
    double B0 = q(ijk0[0], ijk0[1], ijk0[2], IA + dir0, o);
    double B1 = q(ijk1[0], ijk1[1], ijk1[2], IA + dir1, o);
    if (s.time_integrator == 0) // HANCOCK only (ComputeEmfAndStoreFunctor.cpp:297-301)
    {
      B0 += sfm(ijk0[0], ijk0[1], ijk0[2], dir0, o);
      B1 += sfm(ijk1[0], ijk1[1], ijk1[2], dir1, o);
    }
    const double dB0d1 = limited_slope_q(ijk0[0], ijk0[1], ijk0[2], IA + dir0, o, dir1);
    const double dB1d0 = limited_slope_q(ijk1[0], ijk1[1], ijk1[2], IA + dir1, o, dir0);
    double       dq0[8], dq1[8];
    get_state(1 + dir0, is, js, ks, o, dq0);
    get_state(1 + dir1, is, js, ks, o, dq1);
    const double s0 = double(sdq0), s1 = double(sdq1);
    qo[ID] += 0.5 * (s0 * dq0[ID] + s1 * dq1[ID]);
    qo[IP] += 0.5 * (s0 * dq0[IP] + s1 * dq1[IP]);
    qo[IU] += 0.5 * (s0 * dq0[IU] + s1 * dq1[IU]);
    qo[IV] += 0.5 * (s0 * dq0[IV] + s1 * dq1[IV]);
    qo[IW] += 0.5 * (s0 * dq0[IW] + s1 * dq1[IW]);
    if (dir0 == IX)
      qo[IA] = B0 + 0.5 * (double(sb0) * dB0d1);
    else
      qo[IA] += 0.5 * (s0 * dq0[IA] + s1 * dq1[IA]);
    if (dir0 == IY)
      qo[IB] = B0 + 0.5 * (double(sb0) * dB0d1);
    else
      qo[IB] += 0.5 * (s0 * dq0[IB] + s1 * dq1[IB]);
    if (dir1 == IY)
      qo[IB] = B1 + 0.5 * (double(sb1) * dB1d0);
    else
      qo[IB] += 0.5 * (s0 * dq0[IB] + s1 * dq1[IB]);
    if (dir1 == IZ)
      qo[IC] = B1 + 0.5 * (double(sb1) * dB1d0);
    else
      qo[IC] += 0.5 * (s0 * dq0[IC] + s1 * dq1[IC]);
    qo[ID] = std::fmax(s.smallr, qo[ID]);
    qo[IP] = std::fmax(s.smallp * qo[ID], qo[IP]);
  }

  // ComputeEmfAndStoreFunctor::compute_emf_and_store_{2d,3d} (.cpp:371-590)
  void
  compute_emf_and_store(double dt)
  {
    const int  b = lay.b, dim = m.dim;
    const long nc = emf_cells();
    const long total = long(m.nowned) * nc;
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (long gi = 0; gi < total; ++gi)
    {
      const int  o = int(gi / nc);
      const long ci = gi - long(o) * nc;
      const int  n = b + 1;
      int        i = int(ci % n);
      int        j = int((ci / n) % n);
      int        k = dim == 3 ? int(ci / (long(n) * n)) : 0;
      const double dx = dx_of(o);
      const double dtdx = dt / dx;
      double       qE[4][8];
      if (dim == 2)
      {
        reconstruct_edge_2d(i, j, o, LB, qE[ILB]);
        reconstruct_edge_2d(i - 1, j - 1, o, RT, qE[IRT]);
        reconstruct_edge_2d(i - 1, j, o, RB, qE[IRB]);
        reconstruct_edge_2d(i, j - 1, o, LT, qE[ILT]);
        const double e = compute_emf(qE, EMFZ, s);
        em(emf, i, j, 0, 0, o) = e * dtdx;
        continue;
      }
      if (k < b)
      {
        reconstruct_edge_3d(i, j, k, o, LB, IX, IY, qE[ILB]);
        reconstruct_edge_3d(i - 1, j - 1, k, o, RT, IX, IY, qE[IRT]);
        reconstruct_edge_3d(i - 1, j, k, o, RB, IX, IY, qE[IRB]);
        reconstruct_edge_3d(i, j - 1, k, o, LT, IX, IY, qE[ILT]);
        const double e = compute_emf(qE, EMFZ, s);
        em(emf, i, j, k, IZ, o) = e * dtdx;
      }
      if (j < b)
      {
        reconstruct_edge_3d(i, j, k, o, LB, IX, IZ, qE[ILB]);
        reconstruct_edge_3d(i - 1, j, k - 1, o, RT, IX, IZ, qE[IRT]);
        reconstruct_edge_3d(i - 1, j, k, o, RB, IX, IZ, qE[IRB]);
        reconstruct_edge_3d(i, j, k - 1, o, LT, IX, IZ, qE[ILT]);
        for (int v = 0; v < 8; ++v)
          std::swap(qE[IRB][v], qE[ILT][v]);
        const double e = compute_emf(qE, EMFY, s);
        em(emf, i, j, k, IY, o) = e * dtdx;
      }
      if (i < b)
      {
        reconstruct_edge_3d(i, j, k, o, LB, IY, IZ, qE[ILB]);
        reconstruct_edge_3d(i, j - 1, k - 1, o, RT, IY, IZ, qE[IRT]);
        reconstruct_edge_3d(i, j - 1, k, o, RB, IY, IZ, qE[IRB]);
        reconstruct_edge_3d(i, j, k - 1, o, LT, IY, IZ, qE[ILT]);
        const double e = compute_emf(qE, EMFX, s);
        em(emf, i, j, k, IX, o) = e * dtdx;
      }
    }
  }

  // -------------------------------------------------------------------
  // CorrectEmfAtBlockBorderFunctor::correct_emf (scheme/CorrectEmfAtBlockBorderFunctor.cpp:94-360).
  // Edge siblings (kalypsso::core StencilHelper::getEdgeSiblingLoc, absent) restated:
  // the four cells around a border edge are indexed by quadrant (s1,s2) in {0,1}^2 over the two
  // transverse directions (cell index = edge index - 1 + s).  The block's own quadrant rounds
  // the edge index down (clamped to b-1).  The three slots are the other quadrants; a slot whose
  // cell lies in this block is invalid, otherwise it names the leaf holding that cell and the
  // edge's coordinates in it.  Sums run in canonical quadrant order (0,0),(1,0),(0,1),(1,1) so
  // that every block sharing an edge forms the same value.
  // -------------------------------------------------------------------
  struct Slot
  {
    bool valid = false;
    bool outside = false;
    int  oct = -1;
    int  rel = 0;
    int  e[3] = { 0, 0, 0 }; // edge coordinates in the sibling (first fine edge if finer)
  };

  Slot
  edge_slot(int o, int a, int t1, int t2, int e[3], int s1, int s2)
  {
    const int b = lay.b, dim = m.dim;
    Slot      sl_;
    int       c[3] = { e[0], e[1], e[2] };
    c[t1] = e[t1] - 1 + s1;
    c[t2] = e[t2] - 1 + s2;
    bool inside = true;
    for (int d = 0; d < dim; ++d)
      if (c[d] < 0 || c[d] >= b)
        inside = false;
    if (inside)
      return sl_;
    Loc l = m.locate(o, c[0], c[1], c[2]);
    if (l.oct < 0)
      return sl_;
    sl_.valid = true;
    sl_.oct = l.oct;
    sl_.rel = l.rel;
    sl_.outside = l.oct >= m.nowned;
    const int L = m.level[o];
    for (int d = 0; d < dim; ++d)
    {
      // global edge/cell coordinate at the current level
      long       G = long(m.coord[3 * o + d]) * b + e[d];
      const int  Ls = L + l.rel;
      const long scale_num = l.rel == 1 ? 2 : 1;
      long       Gs;
      if (l.rel == -1)
        Gs = G / 2; // unused for coarser siblings
      else
        Gs = G * scale_num;
      long       loc = Gs - long(m.coord[3 * l.oct + d]) * b;
      const long n = m.ncells(Ls, d);
      if (d == a)
      {
        // along the edge: cell index (first of the two fine segments when finer)
        if (loc < 0)
          loc += n;
        if (loc >= n)
          loc -= n;
        if (loc >= b && m.periodic[d])
          loc = ((loc % n) + n) % n;
      }
      else
      {
        if (loc < 0)
          loc += n;
        if (loc > b)
          loc -= n;
      }
      sl_.e[d] = int(loc);
    }
    if (dim == 2)
      sl_.e[2] = 0;
    return sl_;
  }

  void
  correct_emf()
  {
    const int  b = lay.b, dim = m.dim;
    const int  ndir = dim == 3 ? 3 : 1;
    const long nc = emf_cells();
    const double surface_ratio = double(1 << (dim - 1));
    const long total = long(m.nowned) * nc * ndir;
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (long gi = 0; gi < total; ++gi)
    {
      const int  o = int(gi / (nc * ndir));
      const long r = gi - long(o) * nc * ndir;
      const int  cdir = int(r / nc);
      const long ci = r - long(cdir) * nc;
      const int  n = b + 1;
      int        e[3] = { int(ci % n), int((ci / n) % n), dim == 3 ? int(ci / (long(n) * n)) : 0 };
      const int  a = dim == 3 ? cdir : IZ; // edge direction
      int        t1, t2;
      if (a == IX)
      {
        t1 = IY; t2 = IZ;
      }
      else if (a == IY)
      {
        t1 = IX; t2 = IZ;
      }
      else
      {
        t1 = IX; t2 = IY;
      }
      if (dim == 3 && e[a] >= b)
        continue; // not an edge of this direction
      const bool border1 = e[t1] == 0 || e[t1] == b;
      const bool border2 = e[t2] == 0 || e[t2] == b;
      if (!border1 && !border2)
        continue; // norm_v == 0
      const int comp = dim == 3 ? a : 0;
      // own quadrant
      const int own1 = e[t1] < b ? 1 : 0;
      const int own2 = e[t2] < b ? 1 : 0;
      Slot      slots[4];
      for (int s2 = 0; s2 < 2; ++s2)
        for (int s1 = 0; s1 < 2; ++s1)
        {
          if (s1 == own1 && s2 == own2)
            continue;
          slots[s1 + 2 * s2] = edge_slot(o, a, t1, t2, e, s1, s2);
        }
      bool coarser = false, finer = false, all_same = true;
      for (int qd = 0; qd < 4; ++qd)
      {
        if (qd == own1 + 2 * own2)
          continue;
        const Slot & S = slots[qd];
        if (S.valid && S.rel < 0)
          coarser = true;
        if (S.valid && S.rel > 0)
          finer = true;
        if (!(S.valid && S.rel == 0))
          all_same = false;
      }
      const double own_val = em(emf, e[0], e[1], e[2], comp, o);
      double       out = own_val;
      bool         write = false;
      if (coarser)
      {
        double sum = 0;
        int    count = 0;
        for (int qd = 0; qd < 4; ++qd)
        {
          if (qd == own1 + 2 * own2)
          {
            sum += own_val;
            count++;
            continue;
          }
          const Slot & S = slots[qd];
          if (S.valid && S.rel == 0 && !S.outside)
          {
            sum += em(emf, S.e[0], S.e[1], S.e[2], comp, S.oct);
            count++;
          }
        }
        out = sum / double(count);
        write = true;
      }
      else if (finer)
      {
        double sum = 0;
        int    count = 0;
        for (int qd = 0; qd < 4; ++qd)
        {
          if (qd == own1 + 2 * own2)
            continue;
          const Slot & S = slots[qd];
          if (S.valid && S.rel > 0 && !S.outside)
          {
            sum += em(emf, S.e[0], S.e[1], S.e[2], comp, S.oct);
            if (dim == 3)
            {
              int e2[3] = { S.e[0], S.e[1], S.e[2] };
              e2[a] += 1;
              sum += em(emf, e2[0], e2[1], e2[2], comp, S.oct);
            }
            count++;
          }
        }
        if (count > 0)
        {
          out = sum / double(count) / surface_ratio;
          write = true;
        }
      }
      if (all_same)
      {
        double sum = 0;
        int    count = 0;
        for (int qd = 0; qd < 4; ++qd)
        {
          if (qd == own1 + 2 * own2)
          {
            sum += own_val;
            count++;
            continue;
          }
          const Slot & S = slots[qd];
          if (S.valid && !S.outside)
          {
            sum += em(emf, S.e[0], S.e[1], S.e[2], comp, S.oct);
            count++;
          }
        }
        out = sum / double(count);
        write = true;
      }
      if (write)
        em(emf2, e[0], e[1], e[2], comp, o) = out;
    }
  }

  // ReadEmfAndUpdateFunctor::read_emf_and_update_{2d,3d} (.cpp:58-163)
  void
  read_emf_and_update(double * B2)
  {
    const int  b = lay.b, dim = m.dim;
    const long nf = lay.fper();
    const long total = long(m.nowned) * nf;
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (long gi = 0; gi < total; ++gi)
    {
      const int  o = int(gi / nf);
      const long fi = gi - long(o) * nf;
      int        c = 0;
      long       r = fi;
      while (c < 2 && r >= lay.fsize(c))
      {
        r -= lay.fsize(c);
        ++c;
      }
      const int nx = (c == IX && !(dim == 2 && c == IZ)) ? b + 1 : b;
      const int ny = (c == IY) ? b + 1 : b;
      int       i = int(r % nx);
      int       j = int((r / nx) % ny);
      int       k = dim == 3 ? int(r / (long(nx) * ny)) : 0;
      double &  Bv = B2[size_t(o) * nf + fi];
      if (dim == 2)
      {
        if (c == IX)
        {
          const double e0 = em(emf2, i, j, 0, 0, o);
          const double e1 = em(emf2, i, j + 1, 0, 0, o);
          Bv += (e1 - e0);
        }
        else if (c == IY)
        {
          const double e1 = em(emf2, i + 1, j, 0, 0, o);
          const double e0 = em(emf2, i, j, 0, 0, o);
          Bv -= (e1 - e0);
        }
        continue;
      }
      double dB = 0;
      if (c == IX)
      {
        dB += (em(emf2, i, j + 1, k, IZ, o) - em(emf2, i, j, k, IZ, o));
        dB -= (em(emf2, i, j, k + 1, IY, o) - em(emf2, i, j, k, IY, o));
      }
      else if (c == IY)
      {
        dB += (em(emf2, i, j, k + 1, IX, o) - em(emf2, i, j, k, IX, o));
        dB -= (em(emf2, i + 1, j, k, IZ, o) - em(emf2, i, j, k, IZ, o));
      }
      else
      {
        dB += (em(emf2, i + 1, j, k, IY, o) - em(emf2, i, j, k, IY, o));
        dB -= (em(emf2, i, j + 1, k, IX, o) - em(emf2, i, j, k, IX, o));
      }
      Bv += dB;
    }
  }

  // AddGravitySourceTerm::operator() (scheme/AddGravitySourceTerm.cpp:42-100), owned cells
  void
  add_gravity_source_term(const double * U, double * U2, double dt)
  {
    const int  dim = m.dim;
    const long ncell = lay.cells();
    const long total = long(m.nowned) * ncell;
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (long gi = 0; gi < total; ++gi)
    {
      const long   o = gi / ncell, ci = gi - o * ncell;
      auto         at = [&](int var) -> double & { return U2[size_t(o * 5 + var) * ncell + ci]; };
      const double rho_old = U[size_t(o * 5 + ID) * ncell + ci];
      const double rho_new = at(ID);
      double       rhou = at(IU), rhov = at(IV), rhow = dim == 3 ? at(IW) : 0.0;
      double       ekin_old = 0.5 * (rhou * rhou + rhov * rhov) / rho_new;
      if (dim == 3)
        ekin_old += 0.5 * (rhow * rhow) / rho_new;
      rhou += dt * s.gravity[IX] * (rho_old + rho_new) / 2;
      rhov += dt * s.gravity[IY] * (rho_old + rho_new) / 2;
      if (dim == 3)
        rhow += dt * s.gravity[IZ] * (rho_old + rho_new) / 2;
      at(IU) = rhou;
      at(IV) = rhov;
      if (dim == 3)
        at(IW) = rhow;
      double ekin_new = 0.5 * (rhou * rhou + rhov * rhov) / rho_new;
      if (dim == 3)
        ekin_new += 0.5 * (rhow * rhow) / rho_new;
      at(IP) += (ekin_new - ekin_old);
    }
  }

  // GodunovImplemV0::do_time_step (scheme/GodunovImplemV0.cpp:101-207).  Contract: U2 == U
  // and B2 == B on entry (the solver copies them); on exit U2, B2 hold t+dt.
  int
  do_time_step(const double * U, double * U2, const double * B, double * B2, double dt)
  {
    try
    {
      fill_Bface_ghosted(B);
      if (!convert_to_primitives(U, B))
      {
        err = "negative internal energy (abort_when_negative_eint) or missing neighbour";
        return -2;
      }
      compute_limited_slopes();
      compute_primitives_predictor(dt);
      compute_elec_field();
      compute_sFaceMag(dt);
      for (int dir = 0; dir < m.dim; ++dir)
      {
        compute_fluxes_and_store(dt, dir);
        read_fluxes_and_update(U2, B2, dir);
      }
      if (s.viscosity_enabled) // GodunovImplemV0.cpp:179-192
        for (int dir = 0; dir < m.dim; ++dir)
        {
          compute_viscous_fluxes_and_store(dt, dir);
          read_fluxes_and_update(U2, B2, dir);
        }
      compute_emf_and_store(dt);
      emf2 = emf; // Kokkos::deep_copy(m_emf2, m_emf) (GodunovImplemV0.cpp:606)
      correct_emf();
      read_emf_and_update(B2);
      if (s.gravity_enabled) // GodunovImplemV0.cpp:200-205
        add_gravity_source_term(U, U2, dt);
    }
    catch (std::exception & ex)
    {
      err = ex.what();
      return -1;
    }
    return 0;
  }

  // ComputeDtMHDFunctor::compute_cfl (scheme/ComputeDtMHDFunctor.cpp:86-151) reduced with
  // Kokkos::Max, then SolverGodunovMHD::compute_dt_local (SolverGodunovMHD.cpp:377-403).
  double
  compute_inv_dt(const double * U, const double * B)
  {
    const int  b = lay.b, dim = m.dim;
    const long ncell = lay.cells();
    const long total = long(m.nowned) * ncell;
    double     invDt = 0.0;
#pragma omp parallel for schedule(static) num_threads(nthreads) reduction(max : invDt)
    for (long gi = 0; gi < total; ++gi)
    {
      const int  o = int(gi / ncell);
      const long ci = gi - long(o) * ncell;
      int        i = int(ci % b);
      int        j = int((ci / b) % b);
      int        k = dim == 3 ? int(ci / (long(b) * b)) : 0;
      const double dx = dx_of(o);
      double       u[8], qq[8], v[3];
      for (int vv = 0; vv < 5; ++vv)
        u[vv] = U[size_t(o * 5 + vv) * ncell + ci];
      if (dim == 2)
      {
        u[IA] = 0.5 * (face_value(B, o, IX, i, j, 0) + face_value(B, o, IX, i + 1, j, 0));
        u[IB] = 0.5 * (face_value(B, o, IY, i, j, 0) + face_value(B, o, IY, i, j + 1, 0));
        u[IC] = 0.5 * (face_value(B, o, IZ, i, j, 0) + face_value(B, o, IZ, i, j, 0));
      }
      else
      {
        u[IA] = 0.5 * (face_value(B, o, IX, i, j, k) + face_value(B, o, IX, i + 1, j, k));
        u[IB] = 0.5 * (face_value(B, o, IY, i, j, k) + face_value(B, o, IY, i, j + 1, k));
        u[IC] = 0.5 * (face_value(B, o, IZ, i, j, k) + face_value(B, o, IZ, i, j, k + 1));
      }
      compute_primitives(u, qq, s);
      find_speed_info(qq, dim, s, v);
      double cand;
      if (!s.gravity_enabled)
      {
        if (dim == 2)
          cand = v[IX] / dx + v[IY] / dx;
        else
          cand = v[IX] / dx + v[IY] / dx + v[IZ] / dx;
      }
      else
      {
        // ComputeDtMHDFunctor::compute_cfl_with_gravity (.cpp:108-249): g dt^2/(2dx) + |v| dt/dx
        // <= cfl, i.e. each v/dx scaled by k / (sqrt(1 + 2k) - 1), k = sum|g| dx / |v|^2 >= 1e-4
        double v2 = v[IX] * v[IX] + v[IY] * v[IY];
        if (dim == 3)
          v2 += v[IZ] * v[IZ];
        double kk = std::fabs(s.gravity[IX]) + std::fabs(s.gravity[IY]);
        if (dim == 3)
          kk += std::fabs(s.gravity[IZ]);
        kk *= dx / v2;
        kk = std::fmax(kk, 1e-4);
        const double factor = kk / (std::sqrt(1.0 + 2.0 * kk) - 1.0);
        cand = 0.0;
        cand += (v[IX] * factor) / dx;
        cand += (v[IY] * factor) / dx;
        if (dim == 3)
          cand += (v[IZ] * factor) / dx;
      }
      invDt = std::fmax(invDt, cand);
    }
    return invDt;
  }

  // FillOutsideCellFunctor (border_conditions/FillOutside.cpp:97-395), for the outside blocks
  // appended after the owned ones.  Each outside block is the same-level image of the leaf
  // across the domain border (its `source`).
  std::vector<int> outside_src; // per outside block: source (inside) block
  std::vector<int> outside_dirs; // per outside block: 3 ints, outward components (-1,0,1)

  int
  fill_outside(double * U, double * B)
  {
    const int  b = lay.b, dim = m.dim;
    const long ncell = lay.cells();
    for (int oo = 0; oo < m.noutside; ++oo)
    {
      const int o = m.nowned + oo;
      const int src = outside_src[oo];
      const int * od = &outside_dirs[3 * oo];
      // hydro: sweep the faces in order XMIN,XMAX,YMIN,YMAX(,ZMIN,ZMAX); the last matching wins
      for (long ci = 0; ci < ncell; ++ci)
      {
        int cc[3] = { int(ci % b), int((ci / b) % b), dim == 3 ? int(ci / (long(b) * b)) : 0 };
        for (int face = 0; face < 2 * dim; ++face)
        {
          const int d = face / 2;
          const int want = (face % 2 == 0) ? -1 : +1;
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
          const long ii = lay.cidx(cin[0], cin[1], cin[2]);
          for (int v = 0; v < 5; ++v)
            U[size_t(o * 5 + v) * ncell + ci] = U[size_t(src * 5 + v) * ncell + ii];
          if (bct == 3)
            U[size_t(o * 5 + IU + d) * ncell + ci] *= -1;
        }
      }
      // face B: clamp the face index to the first/last layer along each outward direction
      const long nf = lay.fper();
      for (long fi = 0; fi < nf; ++fi)
      {
        int  c = 0;
        long r = fi;
        while (c < 2 && r >= lay.fsize(c))
        {
          r -= lay.fsize(c);
          ++c;
        }
        const int nx = (c == IX && !(dim == 2 && c == IZ)) ? b + 1 : b;
        const int ny = (c == IY) ? b + 1 : b;
        int       f[3] = { int(r % nx), int((r / nx) % ny), dim == 3 ? int(r / (long(nx) * ny)) : 0 };
        int       fin[3] = { f[0], f[1], f[2] };
        for (int d = 0; d < dim; ++d)
        {
          if (od[d] == 0)
            continue;
          const int face = 2 * d + (od[d] < 0 ? 0 : 1);
          if (m.bc[face] != 2)
            continue; // WALL for B is a TODO in the reference
          int ext = (d == c && !(dim == 2 && c == IZ)) ? b + 1 : b;
          fin[d] = od[d] < 0 ? 0 : ext - 1;
        }
        B[size_t(o) * nf + fi] = B[size_t(src) * nf + lay.fidx(c, fin[0], fin[1], fin[2])];
      }
    }
    return 0;
  }
};

} // namespace oracle

// ---------------------------------------------------------------------------
// C ABI used by the tests (ctypes)
// ---------------------------------------------------------------------------
extern "C"
{

  // settings: gamma0, cfl, slope_type, smallr, smallc, scaling, abort_when_negative_eint,
  // time_integrator (0 HANCOCK, 1 RK2_SSP), gravity_enabled, gx, gy, gz, viscosity_enabled, mu,
  // viscous predictor
  void *
  oracle_create(int dim, int b, const double * settings, int nowned, int noutside,
                const int * levels, const int * coords, const int * nbrick, const int * periodic,
                const int * bc, const int * outside_src, const int * outside_dirs, int nthreads)
  {
    auto * S = new oracle::Solver();
    S->s.gamma0 = settings[0];
    S->s.cfl = settings[1];
    S->s.slope_type = settings[2];
    S->s.smallr = settings[3];
    S->s.smallc = settings[4];
    S->s.scaling = settings[5];
    S->s.abort_when_negative_eint = settings[6] != 0.0;
    S->s.time_integrator = int(settings[7]);
    S->s.gravity_enabled = settings[8] != 0.0;
    for (int d = 0; d < 3; ++d)
      S->s.gravity[d] = d < dim ? settings[9 + d] : 0.0;
    S->s.viscosity_enabled = settings[12] != 0.0;
    S->s.mu = settings[13];
    S->s.viscosity_predictor = settings[14] != 0.0;
    S->s.prolongation = int(settings[15]);
    S->m.dim = dim;
    S->m.b = b;
    S->m.nowned = nowned;
    S->m.noutside = noutside;
    for (int d = 0; d < 3; ++d)
    {
      S->m.nbrick[d] = nbrick[d];
      S->m.periodic[d] = periodic[d];
    }
    for (int f = 0; f < 6; ++f)
      S->m.bc[f] = bc ? bc[f] : 0;
    const int nt = nowned + noutside;
    S->m.level.assign(levels, levels + nt);
    S->m.coord.assign(coords, coords + 3 * nt);
    for (int o = 0; o < nt; ++o)
      S->m.hash[oracle::Mesh::key(levels[o], coords[3 * o], coords[3 * o + 1], coords[3 * o + 2])] = o;
    if (noutside > 0)
    {
      S->outside_src.assign(outside_src, outside_src + noutside);
      S->outside_dirs.assign(outside_dirs, outside_dirs + 3 * noutside);
    }
    S->nthreads = nthreads > 0 ? nthreads : 1;
    S->setup();
    return S;
  }

  void
  oracle_destroy(void * h)
  {
    delete static_cast<oracle::Solver *>(h);
  }

  void
  oracle_set_threads(void * h, int n)
  {
    static_cast<oracle::Solver *>(h)->nthreads = n > 0 ? n : 1;
  }

  int
  oracle_do_time_step(void * h, const double * U, double * U2, const double * B, double * B2, double dt)
  {
    return static_cast<oracle::Solver *>(h)->do_time_step(U, U2, B, B2, dt);
  }

  double
  oracle_compute_inv_dt(void * h, const double * U, const double * B)
  {
    return static_cast<oracle::Solver *>(h)->compute_inv_dt(U, B);
  }

  int
  oracle_fill_outside(void * h, double * U, double * B)
  {
    return static_cast<oracle::Solver *>(h)->fill_outside(U, B);
  }

  // LinearCombination<dim>::apply of the RK2-SSP driver (SolverGodunovMHD.cpp:583-588,
  // 617-622).  The functor lives in kalypsso::core, which is not in the reference tree (a CMake
  // target, kalypsso::core, with no pinned version: CMakeLists.txt:76).  Restated from its call
  // sites and the two-stage SSP Runge-Kutta scheme it implements (Shu & Osher 1988): the result
  // is written to the second array, y = c0 * x + c1 * y, over every block (num_octants =
  // get_nb_total()) of the cell array and of the face array.
  void
  oracle_linear_combination(void * h, const double * U, double * U_RK, const double * B, double * B_RK, double c0,
                            double c1)
  {
    auto *     S = static_cast<oracle::Solver *>(h);
    const long nt = long(S->m.nowned) + S->m.noutside;
    const long nu = nt * 5 * S->lay.cells(), nb = nt * S->lay.fper();
    for (long i = 0; i < nu; ++i)
      U_RK[i] = c0 * U[i] + c1 * U_RK[i];
    for (long i = 0; i < nb; ++i)
      B_RK[i] = c0 * B[i] + c1 * B_RK[i];
  }

  const char *
  oracle_last_error(void * h)
  {
    return static_cast<oracle::Solver *>(h)->err.c_str();
  }

  // Direct access to single physics functions, for known-answer tests.
  void
  oracle_riemann_hlld(const double * settings, const double * ql, const double * qr, double * flux)
  {
    oracle::Settings s;
    s.gamma0 = settings[0];
    s.smallr = settings[3];
    s.smallc = settings[4];
    s.smallp = s.smallc * s.smallc / s.gamma0;
    double a[8], b[8];
    std::memcpy(a, ql, sizeof(a));
    std::memcpy(b, qr, sizeof(b));
    oracle::riemann_hlld(a, b, flux, s);
  }

  double
  oracle_compute_emf(const double * settings, const double * qedge /* 4x8: RT,RB,LT,LB */, int dir)
  {
    oracle::Settings s;
    s.gamma0 = settings[0];
    s.smallr = settings[3];
    s.smallc = settings[4];
    s.smallp = s.smallc * s.smallc / s.gamma0;
    double qE[4][8];
    std::memcpy(qE, qedge, sizeof(qE));
    return oracle::compute_emf(qE, dir, s);
  }

  void
  oracle_compute_primitives(const double * settings, const double * u, double * q)
  {
    oracle::Settings s;
    s.gamma0 = settings[0];
    s.smallr = settings[3];
    s.smallc = settings[4];
    s.smallp = s.smallc * s.smallc / s.gamma0;
    oracle::compute_primitives(u, q, s);
  }

  double
  oracle_slope(double q, double qp, double qm, double slope_type)
  {
    return oracle::slope_unsplit_scalar(q, qp, qm, slope_type);
  }

  void
  oracle_speed_info(const double * settings, const double * q, int dim, double * out)
  {
    oracle::Settings s;
    s.gamma0 = settings[0];
    oracle::find_speed_info(q, dim, s, out);
  }
}

extern "C"
{
  // Debug/testing hook: run the stages up to the fluxes of direction `dir` and copy the
  // flux array (owned blocks x 8 x flux cells) to `out`.
  int
  oracle_stage_fluxes(void * h, const double * U, const double * B, double dt, int dir, double * out)
  {
    auto * S = static_cast<oracle::Solver *>(h);
    try
    {
      S->fill_Bface_ghosted(B);
      S->convert_to_primitives(U, B);
      S->compute_limited_slopes();
      S->compute_primitives_predictor(dt);
      S->compute_elec_field();
      S->compute_sFaceMag(dt);
      S->compute_fluxes_and_store(dt, dir);
    }
    catch (std::exception & ex)
    {
      S->err = ex.what();
      return -1;
    }
    std::memcpy(out, S->Fl.data(), S->Fl.size() * sizeof(double));
    return 0;
  }
}

extern "C"
{
  // Debug/testing hook: after oracle_stage_fluxes, copy one scratch array
  // (0 Bg, 1 Q, 2 Q2, 3 Sx, 4 Sy, 5 Sz, 6 E, 7 sF) into out; returns its length.
  long
  oracle_scratch(void * h, int which, double * out)
  {
    auto *                      S = static_cast<oracle::Solver *>(h);
    const std::vector<double> * v[8] = { &S->Bg, &S->Q, &S->Q2, &S->Sx, &S->Sy, &S->Sz, &S->E, &S->sF };
    if (which < 0 || which > 7)
      return -1;
    if (out)
      std::memcpy(out, v[which]->data(), v[which]->size() * sizeof(double));
    return long(v[which]->size());
  }
}

// ---------------------------------------------------------------------------
// The AMR cycle's device work (SolverGodunovMHD.cpp:263-319): the refinement indicator and
// flags (compute_refine_flags, :1094-1185) and the remap of U and face B to the adapted leaves
// (map_userdata_after_adapt, :1191-1269).  The reference delegates both to kalypsso::core
// (ComputeRefineFlags, UserDataRemapper; CMake target kalypsso::core, CMakeLists.txt:76, no
// version pinned), which is not in its tree.  Restated from published methods:
//  * Loehner's second-derivative error estimate (CMAME 61, 1987), unsplit (with the mixed
//    derivatives), noise filter eps = [amr] epsilon_lohner, block indicator = max over cells;
//  * cell data: SIMPLE_COPY prolongation, restriction by the mean of the 2^d fine cells summed
//    in pairs (restricting a prolongation is then exact);
//  * face B: injection on the parent's face planes and the mean of the two bracketing parent
//    faces inside the parent cell (prolongation), the mean of the 2^(d-1) fine faces summed in
//    pairs (restriction).
// ---------------------------------------------------------------------------
namespace oracle
{

// Refinement variables: 0-4 a conserved variable of U (rho, E, rho u, rho v, rho w); 10-17 a
// derived quantity of ComputeDerivedQuantities (utils/ComputeDerivedQuantities.h:42-51, 120-228):
// thermal pressure, speed of sound, specific kinetic energy, magnetic pressure, local Mach
// number, vx, vy, vz, from the cell's primitives (compute_primitives of U and the cell-centred
// face field).  The derived array is computed for the owned blocks only and is zero elsewhere
// (a zero-initialised view, :104, 109-111), so outside blocks read as zero.
enum
{
  kDerivedFirst = 10,
  kThermalPressure = 10,
  kSpeedOfSound = 11,
  kSpecificEkin = 12,
  kMagneticPressure = 13,
  kLocalMach = 14,
  kVx = 15,
  kVy = 16,
  kVz = 17
};

// variable v of block o at its cell (i, j, k) inside [0, b)
static double
field_cell(const Solver & S, const double * U, const double * B, int v, int o, int i, int j, int k)
{
  const int  dim = S.m.dim;
  const long ncell = S.lay.cells();
  const long ci = S.lay.cidx(i, j, k);
  if (v < kDerivedFirst)
    return U[(long(o) * 5 + v) * ncell + ci];
  if (o >= S.m.nowned)
    return 0.0;
  double u[8], q[8];
  for (int n = 0; n < 5; ++n)
    u[n] = U[(long(o) * 5 + n) * ncell + ci];
  if (dim == 2)
  {
    u[IA] = 0.5 * (S.face_value(B, o, IX, i, j, 0) + S.face_value(B, o, IX, i + 1, j, 0));
    u[IB] = 0.5 * (S.face_value(B, o, IY, i, j, 0) + S.face_value(B, o, IY, i, j + 1, 0));
    u[IC] = 0.5 * (S.face_value(B, o, IZ, i, j, 0) + S.face_value(B, o, IZ, i, j, 0));
  }
  else
  {
    u[IA] = 0.5 * (S.face_value(B, o, IX, i, j, k) + S.face_value(B, o, IX, i + 1, j, k));
    u[IB] = 0.5 * (S.face_value(B, o, IY, i, j, k) + S.face_value(B, o, IY, i, j + 1, k));
    u[IC] = 0.5 * (S.face_value(B, o, IZ, i, j, k) + S.face_value(B, o, IZ, i, j, k + 1));
  }
  compute_primitives(u, q, S.s);
  switch (v)
  {
  case kThermalPressure: return q[IP];
  case kSpeedOfSound: return std::sqrt(S.s.gamma0 * q[IP] / q[ID]);
  case kSpecificEkin:
    return dim == 2 ? 0.5 * (q[IU] * q[IU] + q[IV] * q[IV]) : 0.5 * (q[IU] * q[IU] + q[IV] * q[IV] + q[IW] * q[IW]);
  case kMagneticPressure: return 0.5 * (q[IA] * q[IA] + q[IB] * q[IB] + q[IC] * q[IC]);
  case kLocalMach:
  {
    double un = q[IU] * q[IU] + q[IV] * q[IV];
    if (dim == 3)
      un += q[IW] * q[IW];
    un = std::sqrt(un);
    const double cs = std::sqrt(S.s.gamma0 * q[IP] / q[ID]);
    return un / cs;
  }
  case kVx: return q[IU];
  case kVy: return q[IV];
  case kVz: return dim == 3 ? q[IW] : 0.0;
  default: throw std::runtime_error("refine indicator: unknown variable");
  }
}

// variable v of owned block o at cell (i, j, k), one ghost layer at most: a neighbour's cell
// (same level), the mean of its 2^dim cells (finer; x fastest, then y, then z) or its
// containing cell (coarser)
static double
u_cell(const Solver & S, const double * U, const double * B, int v, int o, int i, int j, int k)
{
  const int b = S.lay.b, dim = S.m.dim;
  if (i >= 0 && i < b && j >= 0 && j < b && (dim == 2 || (k >= 0 && k < b)))
    return field_cell(S, U, B, v, o, i, j, k);
  const Loc l = S.m.locate(o, i, j, k);
  if (l.oct < 0)
    throw std::runtime_error("refine indicator: missing neighbour");
  if (l.rel == 1)
  {
    double sum = 0;
    for (int dk = 0; dk < (dim == 3 ? 2 : 1); ++dk)
      for (int dj = 0; dj < 2; ++dj)
        for (int di = 0; di < 2; ++di)
          sum += field_cell(S, U, B, v, l.oct, l.ijk[0] + di, l.ijk[1] + dj, l.ijk[2] + dk);
    return sum / double(1 << dim);
  }
  return field_cell(S, U, B, v, l.oct, l.ijk[0], l.ijk[1], l.ijk[2]);
}

// SIMPLE_GRADIENT (the [amr] refine_criterion of kalypsso::core's ComputeRefineFlags, absent from
// the reference tree; restated): the largest normalised jump between the cell and its 2 dim face
// neighbours, |q_n - q_c| / max(|q_n|, |q_c|), zero where both vanish; directions in order x, y,
// z, the low neighbour before the high one
static double
gradient_at(const Solver & S, const double * U, const double * B, int v, int o, int i, int j, int k)
{
  const double u0 = u_cell(S, U, B, v, o, i, j, k);
  double       e = 0.0;
  for (int a = 0; a < S.m.dim; ++a)
    for (int sg = -1; sg <= 1; sg += 2)
    {
      int d[3] = { 0, 0, 0 };
      d[a] = sg;
      const double un = u_cell(S, U, B, v, o, i + d[0], j + d[1], k + d[2]);
      const double m = std::fmax(std::fabs(un), std::fabs(u0));
      const double x = m > 0.0 ? std::fabs(un - u0) / m : 0.0;
      e = std::fmax(e, x);
    }
  return e;
}

static double
lohner_at(const Solver & S, const double * U, const double * B, int v, int o, int i, int j, int k, double eps)
{
  const int dim = S.m.dim;
  auto u = [&](const int d[3]) { return u_cell(S, U, B, v, o, i + d[0], j + d[1], k + d[2]); };
  double num = 0, den = 0;
  for (int a = 0; a < dim; ++a)
    for (int c = 0; c < dim; ++c)
    {
      if (a == c)
      {
        int p[3] = { 0, 0, 0 }, z[3] = { 0, 0, 0 }, q[3] = { 0, 0, 0 };
        p[a] = 1;
        q[a] = -1;
        const double up = u(p), u0 = u(z), um = u(q);
        const double n = (up - u0) - (u0 - um);
        const double d = std::fabs(up - u0) + std::fabs(u0 - um) +
                         eps * (std::fabs(up) + 2.0 * std::fabs(u0) + std::fabs(um));
        num += n * n;
        den += d * d;
      }
      else
      {
        int pp[3] = { 0, 0, 0 }, pm[3] = { 0, 0, 0 }, mp[3] = { 0, 0, 0 }, mm[3] = { 0, 0, 0 };
        pp[a] = 1, pp[c] = 1;
        pm[a] = 1, pm[c] = -1;
        mp[a] = -1, mp[c] = 1;
        mm[a] = -1, mm[c] = -1;
        const double upp = u(pp), upm = u(pm), ump = u(mp), umm = u(mm);
        const double n = ((upp - upm) - (ump - umm)) * 0.25;
        const double d = (std::fabs(upp - upm) + std::fabs(ump - umm)) * 0.25 +
                         eps * (std::fabs(upp) + std::fabs(upm) + std::fabs(ump) + std::fabs(umm)) * 0.25;
        num += n * n;
        den += d * d;
      }
    }
  return den > 0.0 ? std::sqrt(num / den) : 0.0;
}

// one leaf's remap source: 0 copy, 1 prolongation (octant), 2 restriction (children)
struct RemapSrc
{
  int kind = 0, oct = 0, src[8] = { -1, -1, -1, -1, -1, -1, -1, -1 };
};

static int
parity(int v)
{
  return __builtin_popcount(unsigned(v)) & 1;
}

// Face B of a prolonged leaf.  For each parent cell inside the leaf, its 2^d fine cells take:
// on the parent cell's face planes the parent's face (injection), or, on the leaf's boundary
// plane across which an old leaf of this level lies (nbr), that leaf's faces; inside the parent
// cell the mean of the two opposite outer fine faces of each row plus the minimum-norm
// correction that makes every fine cell divergence free: with the fine cells as the nodes of
// the cube graph and the interior faces as its edges, the face between cells n0 < n1 gains
// phi[n0] - phi[n1] with L phi = -div, L the graph Laplacian, solved on its Walsh basis
// (w_s(n) = (-1)^|s & n|, eigenvalue 2 |s|).  2D B_z (cell centred) is copied.
static void
prolong_faces(const Layout & lay, const double * Bo, const RemapSrc & r, const int nbr[6], const int o3[3], double * bn)
{
  const int  dim = lay.dim, b = lay.b, h = b / 2, nc = 1 << dim, nm = 1 << (dim - 1);
  const long fper = lay.fper();
  const double * P = Bo + long(r.src[0]) * fper;
  for (int pk = 0; pk < (dim == 3 ? h : 1); ++pk)
    for (int pj = 0; pj < h; ++pj)
      for (int pi = 0; pi < h; ++pi)
      {
        const int pl[3] = { pi, pj, pk };
        int       pc[3] = { 0, 0, 0 };
        for (int d = 0; d < dim; ++d)
          pc[d] = o3[d] * h + pl[d];
        double outer[3][2][4], inner[3][4];
        for (int c = 0; c < dim; ++c)
        {
          const int t0 = c == 0 ? 1 : 0, t1 = c == 2 ? 1 : 2;
          for (int side = 0; side < 2; ++side)
          {
            const int o = (pl[c] == (side ? h - 1 : 0)) ? nbr[2 * c + side] : -1;
            for (int m = 0; m < nm; ++m)
            {
              if (o >= 0)
              {
                int g[3] = { 0, 0, 0 };
                g[c] = side ? 0 : b;
                g[t0] = 2 * pl[t0] + (m & 1);
                if (dim == 3)
                  g[t1] = 2 * pl[t1] + ((m >> 1) & 1);
                outer[c][side][m] = Bo[long(o) * fper + lay.fidx(c, g[0], g[1], g[2])];
              }
              else
              {
                int q[3] = { pc[0], pc[1], pc[2] };
                q[c] += side;
                outer[c][side][m] = P[lay.fidx(c, q[0], q[1], q[2])];
              }
            }
          }
          for (int m = 0; m < nm; ++m)
            inner[c][m] = 0.5 * (outer[c][0][m] + outer[c][1][m]);
        }
        // fine cell n = (x bit 0, y bit 1, z bit 2); its row m along c from its transverse bits
        auto rowm = [&](int c, int n) {
          const int t0 = c == 0 ? 1 : 0, t1 = c == 2 ? 1 : 2;
          return ((n >> t0) & 1) | (dim == 3 ? ((n >> t1) & 1) << 1 : 0);
        };
        double div[8];
        for (int n = 0; n < nc; ++n)
        {
          double sdiv = 0;
          for (int c = 0; c < dim; ++c)
          {
            const int    m = rowm(c, n);
            const bool   high = (n >> c) & 1;
            const double lo = high ? inner[c][m] : outer[c][0][m];
            const double hi = high ? outer[c][1][m] : inner[c][m];
            sdiv = sdiv + (hi - lo);
          }
          div[n] = sdiv;
        }
        double phi[8] = { 0, 0, 0, 0, 0, 0, 0, 0 };
        for (int sw = 1; sw < nc; ++sw)
        {
          double w = 0;
          for (int n = 0; n < nc; ++n)
            w = parity(sw & n) ? w - div[n] : w + div[n];
          const double coef = w / double(nc * 2 * __builtin_popcount(unsigned(sw)));
          for (int n = 0; n < nc; ++n)
            phi[n] = parity(sw & n) ? phi[n] + coef : phi[n] - coef;
        }
        for (int c = 0; c < dim; ++c)
        {
          const int t0 = c == 0 ? 1 : 0, t1 = c == 2 ? 1 : 2;
          for (int m = 0; m < nm; ++m)
          {
            const int n0 = ((m & 1) << t0) | (dim == 3 ? ((m >> 1) & 1) << t1 : 0);
            inner[c][m] = inner[c][m] + (phi[n0] - phi[n0 | (1 << c)]);
            int g[3] = { 2 * pl[0], 2 * pl[1], 2 * pl[2] };
            g[t0] += m & 1;
            if (dim == 3)
              g[t1] += (m >> 1) & 1;
            bn[lay.fidx(c, g[0], g[1], g[2])] = outer[c][0][m];
            g[c] += 1;
            bn[lay.fidx(c, g[0], g[1], g[2])] = inner[c][m];
            g[c] += 1;
            bn[lay.fidx(c, g[0], g[1], g[2])] = outer[c][1][m];
          }
        }
        if (dim == 2)
          for (int dj = 0; dj < 2; ++dj)
            for (int di = 0; di < 2; ++di)
              bn[lay.fidx(2, 2 * pl[0] + di, 2 * pl[1] + dj, 0)] = P[lay.fidx(2, pc[0], pc[1], 0)];
      }
}

} // namespace oracle

extern "C"
{
  // per owned block: the maximum over its cells of Loehner's estimate of U variable v (and,
  // when ecell is not null, every cell's estimate, block by block)
  // Block indicators of compute_refine_flags (SolverGodunovMHD.cpp:1094-1185) over nvar refinement
  // variables (codes above): criterion 0 LOHNER_UNSPLIT (noise filter eps), 1 SIMPLE_GRADIENT; the
  // maximum over the block's cells and the variables (the flags of several variables then follow
  // from that maximum: refine when any variable asks it, coarsen when all do).  ecell (optional):
  // the per-cell maximum over the variables.
  int
  oracle_refine_indicator_vars(void * h, const double * U, const double * B, int nvar, const int * vars,
                               int criterion, double eps, double * emax, double * ecell)
  {
    auto *     S = static_cast<oracle::Solver *>(h);
    const int  b = S->lay.b, dim = S->m.dim;
    const long ncell = S->lay.cells();
    try
    {
      for (int o = 0; o < S->m.nowned; ++o)
      {
        double e = 0.0;
        for (long ci = 0; ci < ncell; ++ci)
        {
          const int i = int(ci % b), j = int((ci / b) % b), k = dim == 3 ? int(ci / (long(b) * b)) : 0;
          double    xc = 0.0;
          for (int n = 0; n < nvar; ++n)
          {
            const double x = criterion == 1 ? oracle::gradient_at(*S, U, B, vars[n], o, i, j, k)
                                            : oracle::lohner_at(*S, U, B, vars[n], o, i, j, k, eps);
            xc = std::fmax(xc, (x >= 0.0) ? x : HUGE_VAL);
          }
          e = std::fmax(e, xc);
          if (ecell)
            ecell[long(o) * ncell + ci] = xc;
        }
        emax[o] = e;
      }
    }
    catch (std::exception & ex)
    {
      S->err = ex.what();
      return -1;
    }
    return 0;
  }

  int
  oracle_refine_indicator(void * h, const double * U, int v, double eps, double * emax, double * ecell)
  {
    return oracle_refine_indicator_vars(h, U, nullptr, 1, &v, 0, eps, emax, ecell);
  }

  // flags from the indicator: +1 refine, -1 coarsen, 0 keep
  void
  oracle_refine_flags(int n, const int * levels, const double * emax, int level_min, int level_max, double eps_refine,
                      double eps_coarsen, int8_t * flags)
  {
    for (int o = 0; o < n; ++o)
    {
      const int l = levels[o];
      const double e = emax[o];
      flags[o] = l < level_min ? 1
                 : l > level_max ? -1
                 : (e > eps_refine && l < level_max) ? 1
                 : (e < eps_coarsen && l > level_min) ? -1
                                                       : 0;
    }
  }

  // U_new, B_new of the nnew leaves (levels, coords) from U_old, B_old of the nold leaves
  int
  oracle_remap(int dim, int b, const int * nbrick, const int * periodic, int nold, const int * old_levels,
               const int * old_coords, int nnew, const int * new_levels, const int * new_coords, const double * Uo,
               const double * Bo, double * Un, double * Bn)
  {
    oracle::Layout lay;
    lay.dim = dim;
    lay.b = b;
    std::unordered_map<uint64_t, int> idx;
    for (int o = 0; o < nold; ++o)
      idx[oracle::Mesh::key(old_levels[o], old_coords[3 * o], old_coords[3 * o + 1], old_coords[3 * o + 2])] = o;
    auto find = [&](int l, int x, int y, int z) {
      if (l < 0)
        return -1;
      auto it = idx.find(oracle::Mesh::key(l, x, y, z));
      return it == idx.end() ? -1 : it->second;
    };
    const long ncell = lay.cells(), fper = lay.fper();
    const int  nch = 1 << dim;
    for (int n = 0; n < nnew; ++n)
    {
      const int L = new_levels[n], x = new_coords[3 * n], y = new_coords[3 * n + 1], z = new_coords[3 * n + 2];
      oracle::RemapSrc r;
      if ((r.src[0] = find(L, x, y, z)) >= 0)
        r.kind = 0;
      else if ((r.src[0] = find(L - 1, x >> 1, y >> 1, dim == 3 ? z >> 1 : 0)) >= 0)
      {
        r.kind = 1;
        r.oct = (x & 1) | ((y & 1) << 1) | (dim == 3 ? (z & 1) << 2 : 0);
      }
      else
      {
        r.kind = 2;
        for (int c = 0; c < nch; ++c)
          if ((r.src[c] = find(L + 1, 2 * x + (c & 1), 2 * y + ((c >> 1) & 1), dim == 3 ? 2 * z + ((c >> 2) & 1) : 0)) <
              0)
            return -1;
      }
      // the old leaf of this level across each face plane (periodic wrap), or -1
      int nbr[6] = { -1, -1, -1, -1, -1, -1 };
      for (int c = 0; c < dim; ++c)
        for (int side = 0; side < 2; ++side)
        {
          long      q[3] = { x, y, z };
          const long nb = long(nbrick[c]) << L;
          q[c] += side ? 1 : -1;
          if (q[c] < 0 || q[c] >= nb)
          {
            if (!periodic[c])
              continue;
            q[c] = (q[c] + nb) % nb;
          }
          nbr[2 * c + side] = find(L, int(q[0]), int(q[1]), int(q[2]));
        }
      double *       un = Un + long(n) * 5 * ncell;
      double *       bn = Bn + long(n) * fper;
      if (r.kind == 0)
      {
        std::memcpy(un, Uo + long(r.src[0]) * 5 * ncell, 5 * ncell * sizeof(double));
        std::memcpy(bn, Bo + long(r.src[0]) * fper, fper * sizeof(double));
        continue;
      }
      const int o3[3] = { r.oct & 1, (r.oct >> 1) & 1, (r.oct >> 2) & 1 };
      // cells
      for (int v = 0; v < 5; ++v)
        for (int k = 0; k < (dim == 3 ? b : 1); ++k)
          for (int j = 0; j < b; ++j)
            for (int i = 0; i < b; ++i)
            {
              double val;
              if (r.kind == 1)
              {
                const double * P = Uo + (long(r.src[0]) * 5 + v) * ncell;
                val = P[lay.cidx((o3[0] * b + i) / 2, (o3[1] * b + j) / 2, dim == 3 ? (o3[2] * b + k) / 2 : 0)];
              }
              else
              {
                // coarse cell (i, j, k) covers fine cells 2i..2i+1 of the child holding them
                const int cx = 2 * i / b, cy = 2 * j / b, cz = dim == 3 ? 2 * k / b : 0;
                const double * F = Uo + (long(r.src[cx | (cy << 1) | (cz << 2)]) * 5 + v) * ncell;
                // summed in pairs (x pairs, then y, then z): the mean of equal values is exact
                double p[4];
                for (int dk = 0; dk < (dim == 3 ? 2 : 1); ++dk)
                  for (int dj = 0; dj < 2; ++dj)
                  {
                    const int fi = 2 * i - cx * b, fj = 2 * j - cy * b + dj, fk = dim == 3 ? 2 * k - cz * b + dk : 0;
                    p[2 * dk + dj] = F[lay.cidx(fi, fj, fk)] + F[lay.cidx(fi + 1, fj, fk)];
                  }
                const double sum = dim == 3 ? (p[0] + p[1]) + (p[2] + p[3]) : p[0] + p[1];
                val = sum / double(nch);
              }
              un[v * ncell + lay.cidx(i, j, k)] = val;
            }
      // faces
      if (r.kind == 1)
      {
        prolong_faces(lay, Bo, r, nbr, o3, bn);
        continue;
      }
      for (int c = 0; c < 3; ++c)
      {
        const bool cellc = dim == 2 && c == 2;
        const int  ex[3] = { (!cellc && c == 0) ? b + 1 : b, (!cellc && c == 1) ? b + 1 : b,
                             dim == 3 ? (c == 2 ? b + 1 : b) : 1 };
        for (int k = 0; k < ex[2]; ++k)
          for (int j = 0; j < ex[1]; ++j)
            for (int i = 0; i < ex[0]; ++i)
            {
              const int f[3] = { i, j, k };
              double    val;
              const int side = (!cellc && f[c] == b) ? 1 : 0;
              const int nbo = (!cellc && (f[c] == 0 || f[c] == b)) ? nbr[2 * c + side] : -1;
              if (nbo >= 0)
              {
                // a boundary plane shared with an old leaf of this level keeps that leaf's faces
                int g[3] = { i, j, k };
                g[c] = side ? 0 : b;
                val = Bo[long(nbo) * fper + lay.fidx(c, g[0], g[1], g[2])];
              }
              else
              {
                int ch[3] = { 0, 0, 0 }, fl[3] = { 0, 0, 0 };
                for (int d = 0; d < dim; ++d)
                {
                  ch[d] = 2 * f[d] >= b ? 1 : 0;
                  fl[d] = 2 * f[d] - ch[d] * b;
                }
                const double * F = Bo + long(r.src[ch[0] | (ch[1] << 1) | (ch[2] << 2)]) * fper;
                if (cellc)
                {
                  const double r0 = F[lay.fidx(2, fl[0], fl[1], 0)] + F[lay.fidx(2, fl[0] + 1, fl[1], 0)];
                  const double r1 = F[lay.fidx(2, fl[0], fl[1] + 1, 0)] + F[lay.fidx(2, fl[0] + 1, fl[1] + 1, 0)];
                  val = (r0 + r1) / 4.0;
                }
                else
                {
                  // pairs along the first transverse direction, then the second
                  const int t0 = c == 0 ? 1 : 0, t1 = c == 2 ? 1 : 2;
                  double    p[2] = { 0, 0 };
                  for (int d1 = 0; d1 < (dim == 3 ? 2 : 1); ++d1)
                  {
                    int q[3] = { fl[0], fl[1], fl[2] };
                    if (dim == 3)
                      q[t1] += d1;
                    int q1[3] = { q[0], q[1], q[2] };
                    q1[t0] += 1;
                    p[d1] = F[lay.fidx(c, q[0], q[1], q[2])] + F[lay.fidx(c, q1[0], q1[1], q1[2])];
                  }
                  val = (dim == 3 ? p[0] + p[1] : p[0]) / double(1 << (dim - 1));
                }
              }
              bn[lay.fidx(c, i, j, k)] = val;
            }
      }
    }
    return 0;
  }
}
