// Device-side arithmetic of the line-driven disk-wind step (2-D spherical
// r-theta, HD + IDEAL, one passive tracer), gfx950 FP64.
//
// Expressions keep the reference's operation order (compiled with
// -ffp-contract=off).  Arithmetic that involves only +, -, *, / and sqrt
// is bit-identical to the reference CPU build; pow / exp / log10 / cosh
// come from the device math library and may differ from the host libm in
// the last bit.
#pragma once
#include <hip/hip_runtime.h>
#include "pmx_hd_math.h"

// The cooling arithmetic compiles for the host too, so host-side tools can
// exercise exactly the device code.
#define PMX_HD __host__ __device__ __forceinline__
#define PMX_HD_NOINLINE __host__ __device__ inline

namespace pmx {
namespace wind {

enum { TRC = 5, NVW = 6, NANG_MAX = 64 };

// Iteration caps of the cooling loops.  The reference's loops have no
// bound for non-finite states (a NaN energy makes every Cash-Karp step a
// rejection); these caps end such loops and the cell is flagged.  Finite
// states never reach them: the Cash-Karp driver stops after 1000 accepted
// sub-steps (the reference aborts there) and each rejection shrinks the
// step by at least 0.9.
constexpr int BRACKET_MAX = 20000;
constexpr int CK45_ITER_MAX = 20000;
constexpr int SUBSTEP_MAX = 20000;

// Physical constants of Src/pluto.h:296-321
constexpr double C_amu = 1.66053886e-24;
constexpr double C_c = 2.99792458e10;
constexpr double C_kB = 1.3806505e-16;
constexpr double C_mp = 1.67262171e-24;
constexpr double C_sigmaT = 6.6524e-25;

// Problem constants (host-evaluated with the reference's expressions).
struct Const {
  double gamma, gmm1;
  double UD, UL, UV, UT, UP, UA, KELVIN;   // unit system, Src/pluto.h:700-716
  double mu;                               // g_inputParam[MU]
  double mu_tab;                           // MeanMolecularWeight() of TABULATED cooling
  double gm_code;                          // G M / (UL UV^2), cv_idl/init.c:367-382
  double sigma_e;                          // sigmaT / amu / 1.18, line_connect.c:836
  double krad, alpharad;                   // force-multiplier parameters
  double lx;                               // L_star * f_x
  double t_x;                              // T_x
  double E_cost;                           // UL / UD / UV^3, cv_idl/radiat.c:47
  double dfloor, pfloor;                   // internal-boundary floors (code units); 0: off
  int char_lim;                            // CHAR_LIMITING YES
  int ent;                                 // ENTROPY_SWITCH ALWAYS
  int hll;                                 // [Solver] hll instead of hllc
  int roe;                                 // [Solver] roe
  int fast_div;                            // line-force quotients through div_rc
  int fm_fast;                             // factored force-multiplier power (see force_mult_pair)
  double first_dt;                         // runtime->first_dt (dt-too-small stop)
  int fm_points;                           // > 0: tabulated force multiplier (krad = alpharad = 999)
  int midplane;                            // the cv_idl side-0 pass resets the last polar row
};

// ---------------------------------------------------------------------
// The production problem's internal boundary, UserDefBoundary(side 0)
// (Test_Problems/LineDrivenWind/cv_idl/init.c:196-273), on one cell's
// primitive state: a density below DFLOOR is raised to it with the
// velocities scaled by rho/DFLOOR (momentum kept) and the pressure set to
// keep the sound speed, then held above 500 K; the tracer is cleared; the
// pressure is held above pfloor.  Returns 1 if the cell changed (its Uc
// is then PrimToCons of the new state).  cs^2 is formed as cs*cs (the
// correctly rounded square) where the reference calls pow(cs, 2); glibc
// 2.35's pow(x, 2) differs from x*x by one ulp on ~0.09% of random inputs.
// ---------------------------------------------------------------------
PMX_HD int floor_cell(const Const& k, double* v) {
  const double tfloor = 5.e2;
  int changed = 0;
  if (v[RHO] < k.dfloor) {
    if (v[RHO] < 0.0) v[RHO] = k.dfloor;
    const double cs = sqrt(k.gamma * v[PRS] / v[RHO]);
    const double dfact = v[RHO] / k.dfloor;
    v[RHO] = k.dfloor;
    v[VX1] = dfact * v[VX1];
    v[VX2] = dfact * v[VX2];
    v[VX3] = dfact * v[VX3];
    v[PRS] = (cs * cs) * v[RHO] / k.gamma;
    const double temp = v[PRS] / v[RHO] * k.KELVIN * k.mu;
    if (temp < tfloor) v[PRS] = v[RHO] * tfloor / (k.KELVIN * k.mu);
    v[TRC] = 0.0;
    changed = 1;
  }
  if (v[PRS] < k.pfloor) {
    v[PRS] = k.pfloor;
    changed = 1;
  }
  return changed;
}

// ---------------------------------------------------------------------
// prim <-> cons with the tracer (Src/HD/mappers.c:65-67, :247-249)
// ---------------------------------------------------------------------
PMX_HD void p2c6(const double* v, double* u, double gmm1) {
  prim_to_cons(v, u, gmm1);
  u[TRC] = v[RHO] * v[TRC];
}

PMX_HD int c2p6(double* u, double* v, double gmm1) {
  int f = cons_to_prim(u, v, gmm1);
  v[TRC] = u[TRC] * (1.0 / u[RHO]);
  return f;
}

// ConsToPrim with FLAG_ENTROPY, which ENTROPY_SWITCH ALWAYS sets on every
// zone (flag_shock.c:140-143, HD/mappers.c:129-204, :247-249): the pressure
// from the conserved entropy ue = rho s, s = p / rho^gamma, and the energy
// redefined from it.
PMX_HD int c2p6_entropy(double* u, double ue, double* v, double gmm1) {
  const double small_rho = 1.e-12, small_p = 1.e-12;  // Src/globals.h:115-116
  int fail = 0;
  double m2 = u[VX1] * u[VX1] + u[VX2] * u[VX2] + u[VX3] * u[VX3];
  if (u[RHO] < 0.0) { u[RHO] = small_rho; fail = 1; }
  v[RHO] = u[RHO];
  double tau = 1.0 / u[RHO];
  v[VX1] = u[VX1] * tau;
  v[VX2] = u[VX2] * tau;
  v[VX3] = u[VX3] * tau;
  double kin = 0.5 * m2 / u[RHO];
  if (u[ENG] < 0.0) { u[ENG] = small_p / gmm1 + kin; fail |= 2; }
  double rhog1 = pow(v[RHO], gmm1);
  v[PRS] = ue * rhog1;
  if (v[PRS] < 0.0) { v[PRS] = small_p; fail |= 4; }
  u[ENG] = v[PRS] / gmm1 + kin;
  v[TRC] = u[TRC] * tau;
  return fail;
}

// The entropy scalar Vc[ENTR] = p / rho^gamma that ComputeEntropy sets on
// every zone once the boundary is in place (entropy_switch.c:14-37,
// EOS/Ideal/eos.c:77-96).
PMX_HD double entropy_of(const double* v, double gamma) {
  return v[PRS] / pow(v[RHO], gamma);
}

// PLM of one scalar (the entropy) as the tracer: van Leer, or minmod for a
// FLAG_MINMOD zone (plm_states.c:177-258; characteristic limiting leaves
// scalars on the van Leer limiter, :651-660).
PMX_HD void plm_scalar(double sm1, double s0, double sp1, const double* c, bool mm, double& sp,
                       double& sm) {
  double dvp = (sp1 - s0) * c[2], dvm = (s0 - sm1) * c[3], dl;
  if (mm) dl = (dvp * dvm > 0.0 ? (fabs(dvp) < fabs(dvm) ? dvp : dvm) : 0.0);
  else dl = (dvp * dvm > 0.0 ? (dvp) * (dvm) * (c[0] * (dvm) + c[1] * (dvp)) /
                                  ((dvp) * (dvp) + (dvm) * (dvm) + (c[0] + c[1] - 2.0) * (dvp) * (dvm))
                            : 0.0);
  sp = s0 + dl * c[4];
  sm = s0 - dl * c[5];
}

// General-grid van Leer PLM of one cell (Src/States/plm_states.c:141-258
// with plm_coeffs.h:140-142): c = {cp, cm, wp, wm, dp, dm}.
PMX_HD void plm_vl_general(const double* vm1, const double* v0,
                                               const double* vp1, const double* c,
                                               double* vp, double* vm) {
#pragma unroll
  for (int nv = 0; nv < NVW; nv++) {
    double dvp = (vp1[nv] - v0[nv]) * c[2];
    double dvm = (v0[nv] - vm1[nv]) * c[3];
    double dl = (dvp * dvm > 0.0
                     ? (dvp) * (dvm) * (c[0] * (dvm) + c[1] * (dvp)) /
                           ((dvp) * (dvp) + (dvm) * (dvm) + (c[0] + c[1] - 2.0) * (dvp) * (dvm))
                     : 0.0);
    vp[nv] = v0[nv] + dl * c[4];
    vm[nv] = v0[nv] - dl * c[5];
  }
}

// The same reconstruction with the minmod slope that SHOCK_FLATTENING
// MULTID imposes on FLAG_MINMOD zones (plm_states.c:181-186, SET_MM_LIMITER
// and ABS_MIN of plm_coeffs.h:76-77 and macros.h:153).
PMX_HD void plm_mm_general(const double* vm1, const double* v0, const double* vp1,
                           const double* c, double* vp, double* vm) {
#pragma unroll
  for (int nv = 0; nv < NVW; nv++) {
    double dvp = (vp1[nv] - v0[nv]) * c[2];
    double dvm = (v0[nv] - vm1[nv]) * c[3];
    double dl = (dvp * dvm > 0.0 ? (fabs(dvp) < fabs(dvm) ? dvp : dvm) : 0.0);
    vp[nv] = v0[nv] + dl * c[4];
    vm[nv] = v0[nv] - dl * c[5];
  }
}

// PLM on characteristic variables (CHAR_LIMITING YES, plm_states.c:479-676)
// for one zone of a DIR-sweep, with the adiabatic HD primitive eigenvectors
// (eigenv.c:92-189) and PrimToChar (eigenv.c:575-605).  The reference
// clears the eigenvector arrays before each sweep (ResetState), so the
// projection R.dw runs over a matrix whose other entries are zero; it is
// written here as that full 5 x 5 product so that it also propagates
// non-finite differences as the reference does.  mm: FLAG_MINMOD zone.
template <int DIR>
PMX_HD void plm_char_general(const double* vm1, const double* v0, const double* vp1,
                             const double* c, double gamma, bool mm, double* vp, double* vm) {
  constexpr int VXn = VX1 + DIR, VXt = DIR ? VX3 : VX2, VXb = DIR ? VX1 : VX3;
  const double cp = c[0], cm = c[1];
  double dvp[NVW], dvm[NVW];
#pragma unroll
  for (int nv = 0; nv < NVW; nv++) {
    dvp[nv] = (vp1[nv] - v0[nv]) * c[2];
    dvm[nv] = (v0[nv] - vm1[nv]) * c[3];
  }
  const double cs2 = gamma * v0[PRS] / v0[RHO];            // SoundSpeed2, eos.c:33
  const double cs = sqrt(cs2), rhocs = v0[RHO] * cs, rho_cs = v0[RHO] / cs;
  double R[5][5];
#pragma unroll
  for (int a = 0; a < 5; a++)
#pragma unroll
    for (int b = 0; b < 5; b++) R[a][b] = 0.0;
  R[RHO][0] = 0.5 * rho_cs; R[VXn][0] = -0.5; R[PRS][0] = 0.5 * rhocs;
  R[RHO][1] = 0.5 * rho_cs; R[VXn][1] = 0.5;  R[PRS][1] = 0.5 * rhocs;
  R[RHO][2] = 1.0; R[VXt][3] = 1.0; R[VXb][4] = 1.0;
  const double L0p = 1.0 / rhocs, L2p = -1.0 / cs2;
  double dwp[5], dwm[5], dw[5];
  dwm[0] = -1.0 * dvm[VXn] + L0p * dvm[PRS];
  dwm[1] = 1.0 * dvm[VXn] + L0p * dvm[PRS];
  dwm[2] = dvm[RHO] + L2p * dvm[PRS];
  dwm[3] = dvm[VXt];
  dwm[4] = dvm[VXb];
  dwp[0] = -1.0 * dvp[VXn] + L0p * dvp[PRS];
  dwp[1] = 1.0 * dvp[VXn] + L0p * dvp[PRS];
  dwp[2] = dvp[RHO] + L2p * dvp[PRS];
  dwp[3] = dvp[VXt];
  dwp[4] = dvp[VXb];
#pragma unroll
  for (int k = 0; k < 5; k++) {
    const double p = dwp[k], m = dwm[k];
    if (mm) dw[k] = (p * m > 0.0 ? (fabs(p) < fabs(m) ? p : m) : 0.0);
    else dw[k] = (p * m > 0.0 ? (p) * (m) * (cp * (m) + cm * (p)) /
                                    ((p) * (p) + (m) * (m) + (cp + cm - 2.0) * (p) * (m))
                              : 0.0);
  }
  double dl[NVW];
#pragma unroll
  for (int nv = 0; nv < 5; nv++) {
    double dc = 0.0;
#pragma unroll
    for (int k = 0; k < 5; k++) dc += dw[k] * R[nv][k];
    if (dvp[nv] * dvm[nv] > 0.0) {
      const double a = cp * dvp[nv], b = cm * dvm[nv];
      const double d2v = fabs(a) < fabs(b) ? a : b;                        // ABS_MIN
      dl[nv] = (d2v * dc > 0.0 ? (fabs(d2v) < fabs(dc) ? d2v : dc) : 0.0);  // MINMOD_LIMITER
    } else {
      dl[nv] = 0.0;
    }
  }
  {
    const double p = dvp[TRC], m = dvm[TRC];
    dl[TRC] = (p * m > 0.0 ? (p) * (m) * (cp * (m) + cm * (p)) /
                                 ((p) * (p) + (m) * (m) + (cp + cm - 2.0) * (p) * (m))
                           : 0.0);
  }
#pragma unroll
  for (int nv = 0; nv < NVW; nv++) {
    vp[nv] = v0[nv] + dl[nv] * c[4];
    vm[nv] = v0[nv] - dl[nv] * c[5];
  }
}

// ---------------------------------------------------------------------
// Line force: the force multiplier of one angle bin of LineForce
// (line_connect.c:873-902), 1 + M with M = krad t^alpha capped at 4400,
// t = sigma_e rho v_th / dvds, and M = 0 where dvds <= 0 (or the -999 flag).
//
// Both LineForce calls of a cell (x1 state and cell centre) see the same
// dvds per bin, so t^alpha is evaluated as A * dvds^-alpha with the
// per-call factor A = (sigma_e rho v_th)^alpha computed once per cell:
// one power per bin serves both calls and the division by dvds goes away.
// The per-bin power is exp2(-alpha log2(dvds)): the rounding of the
// product -alpha log2(dvds) (|.| < 40 for any double dvds the line force
// meets) moves the result by at most |alpha log2(dvds)| 2^-53 relative,
// a few 1e-15, where the library pow would carry extra precision at
// several times the cost.  The product differs from pow(t, alpha) by a few
// 1e-15 relative in all; with krad = 0 both forms give M = 0 exactly.
// ---------------------------------------------------------------------
// x / c for c > 0, given rc = 1 / c correctly rounded, bit-identical to
// the IEEE quotient when x = 0 or |x| and |x / c| lie in [2^-900, 2^900]:
// q = x rc and two corrections q += (x - c q) rc, the residual exact in an
// FMA.  By Markstein's theorem (rc correctly rounded, q within an ulp of
// x / c, no underflow or overflow) the last correction gives the correctly
// rounded quotient; the first brings q within an ulp in every case.  Five
// FP64 instructions and no branch, where the division takes about eleven.
PMX_HD double div_rc(double x, double c, double rc) {
  double q = x * rc;
  q = fma(fma(-q, c, x), rc, q);
  q = fma(fma(-q, c, x), rc, q);
  return copysign(q, x);   // +-0 keeps the sign of x, as x / c does for c > 0
}

// The line force's quotients: div_rc where the run's inputs keep every
// operand inside its range (FAST, decided once per run on the host), the
// division otherwise.
template <bool FAST>
PMX_HD double qdiv(double x, double c, double rc) {
  if constexpr (FAST) return div_rc(x, c, rc);
  else return x / c;
}

PMX_HD double force_factor(const Const& k, double rho_cgs, double v_th) {
  return pow(k.sigma_e * rho_cgs * v_th, k.alpharad);
}

// linterp() of line_connect.c:781-811 over one cell's force-multiplier fit:
// x = log10 t, xa = log10 t of the fit points, ya = log10 M of the cell at
// the fit points (stride ys).  The first point not below x is found by a
// linear scan (a NaN x stops at the first point, as there); 10^ of the
// end value outside the fit range, of the linear interpolation inside;
// a NaN result is 0.
PMX_HD double linterp_fm(double x, const double* xa, const double* ya, long long ys, int n) {
  int idx = 0;
  while (idx < n && xa[idx] < x) idx++;
  double result;
  if (idx == 0) {
    result = pow(10.0, ya[0]);
  } else if (idx >= n) {
    result = pow(10.0, ya[(long long)(n - 1) * ys]);
  } else {
    const double x_low = xa[idx - 1], x_high = xa[idx];
    const double y_low = ya[(long long)(idx - 1) * ys], y_high = ya[(long long)idx * ys];
    const double slope = (y_high - y_low) / (x_high - x_low);
    result = pow(10.0, y_low + slope * (x - x_low));
  }
  if (isnan(result)) result = 0.0;
  return result;
}

// Tabulated force multiplier of one LineForce call and bin
// (line_connect.c:876-896): t = sigma_e rho v_th / dvds in the reference's
// operation order, M = linterp(log10 t) capped at 4400, 0 where dvds <= 0.
PMX_HD double force_mult_tab(const Const& k, double rho_cgs, double v_th, double dvds,
                             const double* lt, const double* lm, long long ys) {
  double M = 0.0;
  if (dvds > 0.0) {
    const double t_UV = k.sigma_e * rho_cgs * v_th / dvds;
    M = linterp_fm(log10(t_UV), lt, lm, ys, k.fm_points);
    if (M > 4400.) M = 4400.;
  }
  return 1.0 + M;
}

// ---------------------------------------------------------------------
// The reference's power M = krad pow(t, alpha), t = num / dvds rounded as
// the reference rounds it, with the logarithms shared between the calls:
//   log2 t = log2 num - log2 dvds + log2(1 + eps),
// where eps = (t dvds - num) / num is the quotient's relative rounding
// error (the residual is exact in an FMA) and log2(1 + eps) = eps log2 e
// to 2^-106.  A cell takes log2 num once per call and a bin log2 dvds once
// for both calls, each in double-double (~2^-60), then 2^(alpha log2 t)
// from a double-double argument, rounded once: within 0.6 ulp of the exact
// power (0.575 the largest error found over the wind's ranges) and equal
// to glibc's, which is correctly rounded in practice, for 98.6% of the
// arguments (within an ulp of it for the rest).  About half the two
// library powers' work per bin.
// ---------------------------------------------------------------------
struct DD {
  double hi, lo;
};

PMX_HD DD dd_two_sum(double a, double b) {   // a + b = hi + lo exactly
  const double s = a + b;
  const double bb = s - a;
  return {s, (a - (s - bb)) + (b - bb)};
}

constexpr double LOG2E_HI = 1.4426950408889634, LOG2E_LO = 2.0355273740931033e-17;
constexpr double LN2_HI = 0.6931471805599453, LN2_LO = 2.3190468138462996e-17;

// log2 x for finite x > 0, double-double.  x = m 2^e with m in
// [sqrt(1/2), sqrt(2)); ln m = 2 atanh(s), s = (m - 1) / (m + 1), |s| <=
// 0.1716: 2 s + 2 s^3 (1/3 + s^2/5 + ... + s^20/23), truncation < 2^-66.
PMX_HD DD log2_dd(double x) {
  int e;
  double m = frexp(x, &e);                 // m in [0.5, 1)
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    e -= 1;
  }
  const double a = m - 1.0;                // exact
  const DD b = dd_two_sum(1.0, m);
  const double sh = a / b.hi;
  const double r = fma(-sh, b.hi, a) - sh * b.lo;   // a - sh (b.hi + b.lo)
  const double sl = r / b.hi;
  const double z = sh * sh;
  double p = 1.0 / 23.0;
  p = fma(p, z, 1.0 / 21.0);
  p = fma(p, z, 1.0 / 19.0);
  p = fma(p, z, 1.0 / 17.0);
  p = fma(p, z, 1.0 / 15.0);
  p = fma(p, z, 1.0 / 13.0);
  p = fma(p, z, 1.0 / 11.0);
  p = fma(p, z, 1.0 / 9.0);
  p = fma(p, z, 1.0 / 7.0);
  p = fma(p, z, 1.0 / 5.0);
  p = fma(p, z, 1.0 / 3.0);
  const DD lnm = dd_two_sum(2.0 * sh, 2.0 * sl + 2.0 * sh * z * p);
  const double ph = lnm.hi * LOG2E_HI;
  const double pl = fma(lnm.hi, LOG2E_HI, -ph) + (lnm.hi * LOG2E_LO + lnm.lo * LOG2E_HI);
  const DD q = dd_two_sum((double)e, ph);
  return dd_two_sum(q.hi, q.lo + pl);
}

// 2^y for a double-double y, one rounding at the end (|y| below 1030; the
// far ranges overflow or underflow).  y = n + f, |f| <= 1/2; e^g with g =
// f ln 2 = gh + gl: 1 + gh + gh^2 (1/2! + gh/3! + ... + gh^12/14!) (Taylor,
// truncation < 2^-62) times (1 + gl).
PMX_HD double exp2_dd(double yh, double yl) {
  if (!(yh < 1030.0)) return yh + yh;      // +inf, or NaN as it came
  if (yh < -1100.0) return 0.0;
  const double n = rint(yh);
  const double f = yh - n;                 // exact
  const double gh = f * LN2_HI;
  const double gl = fma(f, LN2_HI, -gh) + (f * LN2_LO + yl * LN2_HI);
  double q = 1.0 / 87178291200.0;          // 1/14!
  q = fma(q, gh, 1.0 / 6227020800.0);
  q = fma(q, gh, 1.0 / 479001600.0);
  q = fma(q, gh, 1.0 / 39916800.0);
  q = fma(q, gh, 1.0 / 3628800.0);
  q = fma(q, gh, 1.0 / 362880.0);
  q = fma(q, gh, 1.0 / 40320.0);
  q = fma(q, gh, 1.0 / 5040.0);
  q = fma(q, gh, 1.0 / 720.0);
  q = fma(q, gh, 1.0 / 120.0);
  q = fma(q, gh, 1.0 / 24.0);
  q = fma(q, gh, 1.0 / 6.0);
  q = fma(q, gh, 0.5);
  const double u = gh * gh * q;            // e^gh - 1 - gh
  const DD one = dd_two_sum(1.0, gh);
  const double r = one.hi + (one.lo + (u + gl * (1.0 + gh + u)));
  return ldexp(r, (int)n);
}

// log2 of one LineForce call's sigma_e rho v_th, and its reciprocal for the
// quotient's rounding error (any few correct bits do).
struct PowNum {
  double num, rn;
  DD L;
};

PMX_HD PowNum pow_num(double num) {
  PowNum p;
  p.num = num;
  p.rn = 1.0 / num;
  p.L = (num >= 2.2250738585072014e-308 && num < HUGE_VAL) ? log2_dd(num) : DD{0.0, 0.0};
  return p;
}

// krad pow(t, alpha) for t = num / dvds, dvds > 0, given D = log2 dvds.
// A t or num that is not a normal positive number (0, subnormal, inf,
// NaN: where the quotient's residual need not be exact) takes the library
// power.
PMX_HD double force_power_shared(const Const& k, const PowNum& a, double dvds, DD D) {
  const double t = a.num / dvds;           // the reference's t (line_connect.c:880)
  const double tiny = 2.2250738585072014e-308;   // smallest normal double
  if (!(t >= tiny && t < HUGE_VAL && a.num >= tiny && a.num < HUGE_VAL))
    return k.krad * pow(t, k.alpharad);
  const double eps = fma(t, dvds, -a.num) * a.rn;
  const DD d = dd_two_sum(a.L.hi, -D.hi);
  const DD dl = dd_two_sum(d.hi, d.lo + ((a.L.lo - D.lo) + eps * LOG2E_HI));
  const double yh = k.alpharad * dl.hi;
  const double yl = fma(k.alpharad, dl.hi, -yh) + k.alpharad * dl.lo;
  return k.krad * exp2_dd(yh, yl);
}

// The reference's form of both calls' multipliers for one bin
// (line_connect.c:879-893): t = (sigma_e rho v_th) / dvds and M = krad
// pow(t, alpha) per call, capped at 4400, the powers through
// force_power_shared (log2 dvds once for both calls).
PMX_HD void force_mult_shared(const Const& k, const PowNum& a1, const PowNum& a2, double dvds,
                              double& f1, double& f2) {
  const double M_max = 4400.;
  double M1 = 0.0, M2 = 0.0;
  if (dvds > 0.0) {
    const DD D = log2_dd(dvds);
    M1 = force_power_shared(k, a1, dvds, D);
    M2 = force_power_shared(k, a2, dvds, D);
    if (M1 > M_max) M1 = M_max;
    if (M2 > M_max) M2 = M_max;
  }
  f1 = 1.0 + M1;
  f2 = 1.0 + M2;
}

PMX_HD void force_mult_pair(const Const& k, double A1, double A2, double dvds,
                            double& f1, double& f2) {
  const double M_max = 4400.;
  double M1 = 0.0, M2 = 0.0;
  if (dvds > 0.0) {
    double P = exp2(-k.alpharad * log2(dvds));
    M1 = k.krad * (A1 * P);
    M2 = k.krad * (A2 * P);
    if (M1 > M_max) M1 = M_max;
    if (M2 > M_max) M2 = M_max;
  }
  f1 = 1.0 + M1;
  f2 = 1.0 + M2;
}

// ---------------------------------------------------------------------
// BLONDIN cooling (Src/Cooling/BLONDIN/cooling.c)
// ---------------------------------------------------------------------
struct Blondin {
  double xi, sqxi, sqsqxi, ne, nH, n, tx, E, hc_init, dt_share;
  double pre[5];   // comp_h, comp_c, xray_h, brem_c, line_c prefactors (cooling.c:74-79)
};

// The five prefactors of a cell from a [5][sv] array, or 1.0 for all
// without one (what read_sirocco_heatcool sets at start-up).
PMX_HD void load_pre(const double* pre, long long sv, long long cell, double* out) {
#pragma unroll
  for (int q = 0; q < 5; q++) out[q] = pre ? pre[q * sv + cell] : 1.0;
}

PMX_HD double ne_rat(double T) {   // cooling.c:210-222
  double ne_ratio;
  if (T < 1.5e4) {
    ne_ratio = 1e-2 + pow(10.0, (-51.59417133 + 12.27740153 * log10(T)));
  } else if (T >= 1.5e4 && T < 3.3e4) {
    ne_ratio = pow(10.0, (-3.80749689 + 0.86092628 * log10(T)));
  } else {
    ne_ratio = 1.21;
  }
  return ne_ratio;
}

// heatcool(), cooling.c:173-202, with the cell's prefactors
PMX_HD double heatcool(Blondin& b, double T) {
  double sqT = sqrt(T);
  b.ne = b.nH * ne_rat(T);
  double comp_heat = b.pre[0] * (8.9e-36 * b.xi * b.tx);
  double comp_cool = b.pre[1] * (8.9e-36 * b.xi * (4.0 * T));
  double xray_heat = b.pre[2] * (1.5e-21 * (b.sqsqxi / sqT));
  double line_cool = b.pre[4] * ((1e-16 * exp(-1.3e5 / T) / b.sqxi / T) +
                                 fmin(fmin(1e-24, 5e-27 * sqT), 1.5e-17 / T));
  double brem_cool = b.pre[3] * (3.3e-27 * sqT);
  return b.nH * (b.ne * comp_heat + b.nH * xray_heat - b.ne * comp_cool - b.ne * line_cool -
                 b.ne * brem_cool);
}

PMX_HD double zfunc(Blondin& b, double temp) {   // cooling.c:233-237
  return (temp * b.n * C_kB / (2.0 / 3.0)) - b.E - b.dt_share * (b.hc_init + heatcool(b, temp)) / 2.0;
}

template <int WHICH>   // 0: zfunc, 1: heatcool
PMX_HD double bfunc(Blondin& b, double x) {
  return WHICH ? heatcool(b, x) : zfunc(b, x);
}

// Brent root finder, cooling.c:250-319 (ITMAX 100, EPS 3e-8)
template <int WHICH>
PMX_HD_NOINLINE double zbrent(Blondin& bl, double x1, double x2, double tol) {
  const int ITMAX = 100;
  const double EPS = 3.0e-8;
  double a = x1, b = x2, c = x2, d = 0.0, e = 0.0;
  double fa = bfunc<WHICH>(bl, a), fb = bfunc<WHICH>(bl, b), fc = fb;
  if (fb * fa > 0.0) return b;
  for (int iter = 1; iter <= ITMAX; iter++) {
    if (fb * fc > 0.0) { c = a; fc = fa; e = d = b - a; }
    if (fabs(fc) < fabs(fb)) { a = b; b = c; c = a; fa = fb; fb = fc; fc = fa; }
    double tol1 = 2.0 * EPS * fabs(b) + 0.5 * tol;
    double xm = 0.5 * (c - b);
    if (fabs(xm) <= tol1 || fb == 0.0) return b;
    if (fabs(e) >= tol1 && fabs(fa) > fabs(fb)) {
      double p, q, r, s = fb / fa;
      if (a == c) {
        p = 2.0 * xm * s;
        q = 1.0 - s;
      } else {
        q = fa / fc;
        r = fb / fc;
        p = s * (2.0 * xm * q * (q - r) - (b - a) * (r - 1.0));
        q = (q - 1.0) * (r - 1.0) * (s - 1.0);
      }
      if (p > 0.0) q = -q;
      p = fabs(p);
      double min1 = 3.0 * xm * q - fabs(tol1 * q), min2 = fabs(e * q);
      if (2.0 * p < (min1 < min2 ? min1 : min2)) { e = d; d = p / q; }
      else { d = xm; e = d; }
    } else {
      d = xm;
      e = d;
    }
    a = b;
    fa = fb;
    if (fabs(d) > tol1) b += d;
    else b += (xm > 0.0 ? fabs(tol1) : -fabs(tol1));
    fb = bfunc<WHICH>(bl, b);
  }
  return b;
}

// One cell of BlondinCooling (cooling.c:50-165).  Returns the new
// pressure (code units), or p unchanged below 1e4 K.
PMX_HD double blondin_cell(const Const& k, double rho_code, double p,
                                               double r_code, double xi_s, double tr_s,
                                               double t, double dt, const double* pre) {
  const double minT = 1.e4;
  Blondin b;
#pragma unroll
  for (int q = 0; q < 5; q++) b.pre[q] = pre[q];
  b.dt_share = dt * k.UT;
  double r = r_code * k.UL;
  double rho = rho_code * k.UD;
  b.E = (p * k.UP) / (k.gamma - 1);
  b.nH = rho / (1.43 * C_mp);
  if (t <= 3.0) {
    b.xi = k.lx / b.nH / r / r;
    b.tx = k.t_x;
  } else {
    b.xi = xi_s;
    b.tx = tr_s;
  }
  b.n = rho / (k.mu * C_mp);
  double T = b.E * (2.0 / 3.0) / (b.n * C_kB);
  if (T < minT) return p;
  b.sqxi = sqrt(b.xi);
  b.sqsqxi = pow(b.xi, 0.25);
  b.hc_init = heatcool(b, T);
  double t_l = T * 0.9, t_u = T * 1.1;
  double test = zfunc(b, t_l) * zfunc(b, t_u);
  // The widening ends by itself once t_u overflows (about 7400 passes);
  // the cap only bounds the loop for non-finite states.
  for (int widen = 0; test > 0 && test == test && widen < BRACKET_MAX; widen++) {
    t_l *= 0.9;
    t_u *= 1.1;
    test = zfunc(b, t_l) * zfunc(b, t_u);
  }
  double T_f;
  if (test != test) {
    T_f = T;
  } else {
    T_f = zbrent<0>(b, t_l, t_u, 1.0);
    double hc_final = heatcool(b, T_f);
    if (hc_final * b.hc_init < 0.0) T_f = zbrent<1>(b, fmin(T_f, T), fmax(T_f, T), 1.0);
  }
  T_f = dmax(T_f, minT);
  double E_f = T_f / (2.0 / 3.0) * (b.n * C_kB);
  return E_f * (k.gamma - 1) / k.UP;
}

// ---------------------------------------------------------------------
// Sub-cycled plain integrator: CoolingSource (cv_idl/cooling_source.c)
// with Radiat (cv_idl/radiat.c) and SolveODE_RKF12 / SolveODE_CK45
// (Src/Cooling/cooling_ode_solver.c).  Only the RHOE slot evolves; rho
// and the tracer are fixed during the sub-step.
// ---------------------------------------------------------------------
struct RadCell {
  double rho;      // code units
  double trc;
  double x1;       // code units
  double xi_s, tr_s;
  double pre[5];   // comp_h, comp_c, xray_h, brem_c, line_c (cooling_source.c:101-105)
  double t;        // g_time
  mutable int nev; // Radiat() evaluations so far (cost accounting)
};

// Radiat(): returns the RHOE rate; may floor *E like the reference.
PMX_HD double radiat(const Const& k, const RadCell& R, double& E) {
  const double minT = 1.e4;
  R.nev++;
  double prs = E * (k.gamma - 1.0);
  if (prs < 0.0) { prs = 1.e-12; E = prs / (k.gamma - 1.0); }
  double T = prs / R.rho * k.KELVIN * k.mu;
  if (T < minT) return 0.0;
  double rho = R.rho * k.UD, nH = rho / (1.43 * C_mp);
  double r = R.x1 * k.UL, xi, tx;
  if (R.t < 3.0) { xi = k.lx / nH / r / r; tx = k.t_x; }
  else { xi = R.xi_s; tx = R.tr_s; }
  double sqT = sqrt(T), sqxi = sqrt(xi), sqsqxi = pow(xi, 0.25);
  double ne = 1.21 * nH;
  // cv_idl/radiat.c:118-123
  double comp_heat = R.pre[0] * (8.9e-36 * xi * tx);
  double comp_cool = R.pre[1] * (8.9e-36 * xi * (4.0 * T));
  double xray_heat = R.pre[2] * (1.5e-21 * (sqsqxi / sqT));
  double line_cool = R.pre[4] * ((1e-16 * exp(-1.3e5 / T) / sqxi / T) +
                                 fmin(fmin(1e-24, 5e-27 * sqT), 1.5e-17 / T));
  double brem_cool = R.pre[3] * (3.3e-27 * sqT);
  double lambda = nH * (ne * comp_heat + nH * xray_heat - ne * comp_cool - ne * line_cool -
                        ne * brem_cool);
  double rhs = (R.trc > 0.01) ? lambda * k.E_cost : 0.0;
  // 1 - 1/cosh(y), y = (T/minT)^12, rounds to 1.0 exactly once 1/cosh(y) <
  // 2^-54 (y > 38): from T/minT >= 1.4 on (y >= 56.7, 1/cosh(y) < 5e-25,
  // far beyond any last-bit error of pow or cosh) the factor is 1.0 and
  // the two calls are skipped, bit for bit the same result
  const double tr = T / minT;
  if (!(tr >= 1.4)) rhs *= 1.0 - 1.0 / cosh(pow(tr, 12));
  return rhs;
}

// SolveODE_RKF12: returns err, writes the 2nd-order result.
PMX_HD double rkf12(const Const& k, const RadCell& R, double E0, double k1,
                                        double& E2nd, double dt) {
  double E1 = E0 + 0.5 * dt * k1;
  double k2 = radiat(k, R, E1);
  E2nd = E0 + dt * k2;
  double E1st = E0 + dt * k1;
  double vscal = fabs(E0) + dt * fabs(k1);
  double err = 0.0;
  double scrh = fabs(E2nd - E1st) / fabs(vscal);
  return dmax(err, scrh);
}

// SolveODE_CK45 over [0, dt]: writes the 5th-order result, returns the
// proposed next step.  E0 and k1 belong to the caller and are not
// changed.  `nsub` counts accepted sub-steps (the reference aborts the
// run past 1000; here the caller sees the count and flags the cell).
PMX_HD_NOINLINE double ck45(const Const& k, const RadCell& R, double E0, double k1, double& E5th,
                       double dt, double tol, int& nsub) {
  const double c1 = 37.0 / 378.0, c2 = 0.0, c3 = 250.0 / 621.0, c4 = 125.0 / 594.0, c5 = 0.0,
               c6 = 512.0 / 1771.0;
  const double cs1 = 2825.0 / 27648.0, cs2 = 0.0, cs3 = 18575.0 / 48384.0,
               cs4 = 13525.0 / 55296.0, cs5 = 277.0 / 14336.0, cs6 = 0.25;
  const double b21 = 0.2, b31 = 3.0 / 40.0, b32 = 9.0 / 40.0, b41 = 0.3, b42 = -0.9, b43 = 1.2;
  const double b51 = -11.0 / 54.0, b52 = 2.5, b53 = -70.0 / 27.0, b54 = 35.0 / 27.0;
  const double b61 = 1631.0 / 55296.0, b62 = 175.0 / 512.0, b63 = 575.0 / 13824.0,
               b64 = 44275.0 / 110592.0, b65 = 253.0 / 4096.0;
  double t = 0.0, tend = dt;
  int ksub = 0;
  for (int iter = 0;; iter++) {
    if (iter >= CK45_ITER_MAX) { ksub = 1001; break; }
    double E1, k2, k3, k4, k5, k6;
    E1 = E0 + dt * b21 * k1;
    k2 = radiat(k, R, E1);
    E1 = E0 + dt * (b31 * k1 + b32 * k2);
    k3 = radiat(k, R, E1);
    E1 = E0 + dt * (b41 * k1 + b42 * k2 + b43 * k3);
    k4 = radiat(k, R, E1);
    E1 = E0 + dt * (b51 * k1 + b52 * k2 + b53 * k3 + b54 * k4);
    k5 = radiat(k, R, E1);
    E1 = E0 + dt * (b61 * k1 + b62 * k2 + b63 * k3 + b64 * k4 + b65 * k5);
    k6 = radiat(k, R, E1);
    E5th = E0 + dt * (c1 * k1 + c2 * k2 + c3 * k3 + c4 * k4 + c5 * k5 + c6 * k6);
    double E4th = E0 + dt * (cs1 * k1 + cs2 * k2 + cs3 * k3 + cs4 * k4 + cs5 * k5 + cs6 * k6);
    double vscal = fabs(E0) + dt * fabs(k1);
    double err = 0.0;
    double scrh = fabs(E5th - E4th) / fabs(vscal);
    err = dmax(err, scrh);
    err /= tol;
    if (err < 1.0) {
      ksub++;
      err = dmax(0.0, 1.e-18);
      t += dt;
      double dt_grow = 0.9 * dt * pow(err, -0.2);
      dt_grow = dmin(dt_grow, 5.0 * dt);
      dt = dt_grow;
      if (fabs(t / tend - 1.0) < 1.e-9) break;
      if (dt > (tend - t)) dt = tend - t;
      E0 = E5th;
      k1 = radiat(k, R, E0);
      if (ksub > 1000) break;
    } else {
      double dt_shrink = 0.9 * dt * pow(err, -0.25);
      dt = dmax(dt_shrink, 0.05 * dt);
    }
  }
  nsub = ksub;
  return dt;
}

// End of CoolingSource for one cell (cv_idl/cooling_source.c): new
// pressure from the integrated energy E1, the 1e4 K floor and the cell's
// dt * maxCoolingRate / |dp/p| candidate for dt_cool.
PMX_HD double finish_cell(const Const& k, const RadCell& R, double p, double E1,
                                              double dt, double& dt_cool_cell) {
  const double minT = 1.e4, maxCoolingRate = 0.1;
  double T0 = p / R.rho * k.KELVIN * k.mu_tab;
  double prs = E1 * (k.gamma - 1.0);
  double T1 = prs / R.rho * k.KELVIN * k.mu_tab;
  if (prs < 0.0) prs = 1.e-12;
  if (T1 < minT && T0 > minT) prs = minT * R.rho / (k.KELVIN * k.mu_tab);
  double e = fabs(prs / p - 1.0);
  dt_cool_cell = dt * maxCoolingRate / e;
  return prs;
}

// One cell of CoolingSource: returns the new pressure and the cell's
// dt * maxCoolingRate / |dp/p| candidate for dt_cool.
// plain_cell_from: the same from E0 and k1 = Radiat(E0) already formed
// (E0 as Radiat left it), for a caller that has evaluated the first rate.
PMX_HD double plain_cell_from(const Const& k, const RadCell& R, double p, double dt, double E0,
                              double k1, double& dt_cool_cell, int& bad) {
  const double min_tol = 2.e-5;
  double E1 = E0;
  int stiff = (fabs(k1) > 0.5 / dt ? 1 : 0);
  double err = 0.0;
  if (!stiff) {
    err = rkf12(k, R, E0, k1, E1, dt);
    err /= min_tol;
  }
  if (stiff || err > 1.0) {
    double dt1 = 0.5 / (fabs(k1) + 1.0 / dt), t1 = 0.0;
    int done = 0, outer = 0;
    do {
      if (++outer > SUBSTEP_MAX) { bad = 1; break; }
      if ((t1 + dt1) >= dt) { dt1 = dt - t1; done = 1; }
      int nsub = 0;
      double dts = ck45(k, R, E0, k1, E1, dt1, min_tol, nsub);
      if (nsub > 1000) bad = 1;
      t1 += dt1;
      dt1 = dts;
      if (done) break;
      E0 = E1;
      k1 = radiat(k, R, E0);
    } while (!done);
  }
  return finish_cell(k, R, p, E1, dt, dt_cool_cell);
}

PMX_HD double plain_cell(const Const& k, const RadCell& R, double p, double dt,
                         double& dt_cool_cell, int& bad) {
  double E0 = p / (k.gamma - 1.0);
  const double k1 = radiat(k, R, E0);
  return plain_cell_from(k, R, p, dt, E0, k1, dt_cool_cell, bad);
}

// ---------------------------------------------------------------------
// Refinement of a surrogate guess for a stiff cell (no counterpart in the
// reference).  Over one cooling step rho and the tracer are fixed, so the
// energy obeys the scalar ODE dE/dt = f(E) = Radiat(E), whose solution
// moves monotonically from E0 towards the first zero of f (an equilibrium)
// or, when cooling, towards the 1e4 K cut below which f = 0.  The exact end
// state over dt is the E with
//      t(E) = integral_{E0}^{E} dE' / f(E') = dt,
// solved here by Newton's method from the network's guess.  The integral
// uses composite 4-point Gauss-Legendre quadrature, in u = ln E for a path
// that stays away from an equilibrium, and in w = ln|E_eq - E| towards an
// equilibrium E_eq (there the integrand tends to the constant 1/lambda of
// the linearised relaxation, so the quadrature stays accurate).
// ---------------------------------------------------------------------
constexpr double GLX[4] = {-0.8611363115940526, -0.3399810435848563, 0.3399810435848563,
                           0.8611363115940526};
constexpr double GLW[4] = {0.3478548451374638, 0.6521451548625461, 0.6521451548625461,
                           0.3478548451374638};
constexpr int REFINE_ITERS = 10;     // Newton iterations (damped steps included)
constexpr int REFINE_ROOT_ITERS = 60;
constexpr int QUAD_STACK = 40;       // adaptive quadrature: bisection depth
constexpr int QUAD_PANELS = 400;     // adaptive quadrature: panels examined per integral
constexpr double QUAD_TOL = 1e-6;    // absolute time tolerance, in units of dt
constexpr double REACH_TOL = 1e-2;   // loose tolerance of the equilibrium reach test
constexpr double EQ_PROBE = 1e-3;    // relative step beyond the guess probing for f = 0

PMX_HD double rate(const Const& k, const RadCell& R, double E) {
  double e = E;
  return radiat(k, R, e);
}

// Integrand of the time integral in the variable x: in u = ln E (mode 0)
// dt/du = E / f(E); in w = ln|Eeq - E| with E = Eeq - s0 e^w (mode 1)
// dt/dw = -s0 e^w / f(E).  False where f does not have the sign s0.
PMX_HD bool integrand(const Const& k, const RadCell& R, int mode, double Eeq, double s0, double x,
                      double& g) {
  double E, dEdx;
  if (mode == 0) {
    E = exp(x);
    dEdx = E;
  } else {
    const double ew = exp(x);
    E = Eeq - s0 * ew;
    dEdx = -s0 * ew;
  }
  const double f = rate(k, R, E);
  if (!(f * s0 > 0.0)) return false;
  g = dEdx / f;
  return true;
}

// 4-point Gauss-Legendre on one panel
PMX_HD bool gl4(const Const& k, const RadCell& R, int mode, double Eeq, double s0, double a,
                double b, double& val, double& xbad) {
  const double h = b - a, c = 0.5 * (a + b);
  val = 0.0;
  for (int q = 0; q < 4; q++) {
    double g;
    const double x = c + 0.5 * h * GLX[q];
    if (!integrand(k, R, mode, Eeq, s0, x, g)) {
      xbad = x;
      return false;
    }
    val += 0.5 * h * GLW[q] * g;
  }
  return true;
}

// Adaptive integral of the time integrand from a to b: panels are halved
// until the two halves agree with the whole panel to within the panel's
// share of the tolerance `tol` (Richardson-corrected sum); false if f
// changes sign at a node (xbad = that node, else NaN) or the panel budget
// runs out.
PMX_HD_NOINLINE bool t_adapt(const Const& k, const RadCell& R, int mode, double Eeq, double s0,
                             double a, double b, double tol, double& t, double& xbad) {
  t = 0.0;
  xbad = NAN;
  if (a == b) return true;
  double sa[QUAD_STACK], sb[QUAD_STACK], sw[QUAD_STACK];
  int sp = 0, panels = 0;
  const double span = fabs(b - a);
  if (!gl4(k, R, mode, Eeq, s0, a, b, sw[0], xbad)) return false;
  sa[0] = a; sb[0] = b; sp = 1;
  while (sp > 0) {
    sp--;
    const double pa = sa[sp], pb = sb[sp], whole = sw[sp], m = 0.5 * (pa + pb);
    double left, right;
    if (!gl4(k, R, mode, Eeq, s0, pa, m, left, xbad)) return false;
    if (!gl4(k, R, mode, Eeq, s0, m, pb, right, xbad)) return false;
    const double diff = left + right - whole;
    if (++panels > QUAD_PANELS) return false;
    if (fabs(diff) <= tol * fabs(pb - pa) / span) {
      t += left + right + diff / 255.0;
    } else {
      if (sp + 2 > QUAD_STACK) return false;
      sa[sp] = m; sb[sp] = pb; sw[sp] = right; sp++;
      sa[sp] = pa; sb[sp] = m; sw[sp] = left; sp++;
    }
  }
  return true;
}

// Surrogate-guided end state of one stiff cell.  E0 > 0, f0 = f(E0) != 0,
// Eg the guess.  Returns 1 with *Ef set, or a negative code (which check
// failed) when the cell must take the plain integrator: the path reaches
// the 1e4 K cut, a quadrature node or iterate crosses an equilibrium where
// it should not, a root is not bracketed, Newton does not converge, or the
// final accuracy check fails.
PMX_HD_NOINLINE int refine_cell(const Const& k, const RadCell& R, double E0, double f0, double dt,
                           double Eg, double& Ef) {
  const double s0 = f0 > 0.0 ? 1.0 : -1.0;
  const double tol = QUAD_TOL * dt;
  const double Ecut = 1.e4 / (k.KELVIN * k.mu) * R.rho / (k.gamma - 1.0);   // T = 1e4 K
  const double u0 = log(E0);
  if (!(Eg > 0.0) || !(Eg - Eg == 0.0)) return -1;   // not positive and finite
  if (s0 * (Eg - E0) <= 0.0) Eg = E0 * exp(s0 * 1e-3);   // guess behind the start
  if (s0 < 0.0 && Eg <= Ecut) return -2;
  double fg = rate(k, R, Eg);
  double Ea = E0, Eb = 0.0;          // bracket of an equilibrium (f(Ea) s0 > 0 >= f(Eb) s0)
  bool have_eq = false;
  if (!(fg * s0 > 0.0)) {
    Eb = Eg;
    have_eq = true;
  } else {
    // A guess that has (almost) relaxed to an equilibrium: the sign change
    // of f is right beyond it, and the cheap equilibrium branch applies.
    const double Et = Eg * exp(s0 * EQ_PROBE);
    if (!(s0 < 0.0 && Et <= Ecut) && !(rate(k, R, Et) * s0 > 0.0)) {
      Ea = Eg;
      Eb = Et;
      have_eq = true;
    }
  }
  if (!have_eq) {
    // Newton in u from the guess
    double u = log(Eg), t, xbad;
    bool ok0 = t_adapt(k, R, 0, 0.0, s0, u0, u, tol, t, xbad);
    if (!ok0 && xbad == xbad) {        // an equilibrium before the guess
      Eb = exp(xbad);
      have_eq = true;
    } else if (!ok0) {
      // the panel budget ran out: the guess sits next to an equilibrium,
      // where the integrand is nearly singular; look just beyond it
      const double dl[4] = {1e-4, 1e-3, 1e-2, 1e-1};
      for (int q = 0; q < 4 && !have_eq; q++) {
        const double Et = Eg * exp(s0 * dl[q]);
        if (s0 < 0.0 && Et <= Ecut) break;
        if (!(rate(k, R, Et) * s0 > 0.0)) {
          Ea = Eg;
          Eb = Et;
          have_eq = true;
        }
      }
      if (!have_eq) return -3;
    }
    double fu = fg;
    for (int it = 0; it < REFINE_ITERS && !have_eq; it++) {
      const double E = exp(u);
      double du = -(t - dt) * fu / E;
      double un = u + du;
      if (s0 * (un - u0) <= 0.0) un = 0.5 * (u + u0);   // stay ahead of the start
      // Near an equilibrium the linear step overshoots far past it; never
      // step across the cut, go half way towards it instead (the sign test
      // below then finds the equilibrium, if there is one, in between).
      if (s0 < 0.0 && un <= log(Ecut)) un = 0.5 * (u + log(Ecut));
      const double En = exp(un);
      if (s0 < 0.0 && En <= Ecut) return -4;
      const double fn = rate(k, R, En);
      if (!(fn * s0 > 0.0)) {          // stepped across an equilibrium
        Ea = E;
        Eb = En;
        have_eq = true;
        break;
      }
      double dtinc;
      if (!t_adapt(k, R, 0, 0.0, s0, u, un, tol, dtinc, xbad)) {
        if (!(xbad == xbad)) return -5;
        Ea = E;                        // an equilibrium between the iterates
        Eb = exp(xbad);
        have_eq = true;
        break;
      }
      t += dtinc;
      u = un;
      fu = fn;
      if (fabs(du) <= 1e-7) {
        Ef = exp(u);
        return 1;
      }
    }
    if (!have_eq) return -8;
  }
  // ---- an equilibrium lies between Ea and Eb: locate it (Illinois) -----
  if (s0 < 0.0 && Eb <= Ecut) return -9;
  double fa = rate(k, R, Ea), fb = rate(k, R, Eb);
  if (fb == 0.0) {
    double Tb = Eb * (k.gamma - 1.0) / R.rho * k.KELVIN * k.mu;
    if (Tb < 1.e4) return -10;          // the cut, not an equilibrium
  }
  double Eeq = Eb;
  {
    double a = Ea, b = Eb;
    int side = 0;
    bool found = false;
    for (int it = 0; it < REFINE_ROOT_ITERS && !found; it++) {
      double c = (a * fb - b * fa) / (fb - fa);
      if (!(c > fmin(a, b) && c < fmax(a, b))) c = 0.5 * (a + b);
      const double fc = rate(k, R, c);
      if (fc == 0.0) {
        Eeq = c;
        found = true;
      } else if (fc * s0 > 0.0) {
        a = c; fa = fc;
        if (side == -1) fb *= 0.5;
        side = -1;
      } else {
        b = c; fb = fc;
        if (side == 1) fa *= 0.5;
        side = 1;
      }
      if (!found && fabs(b - a) <= 1e-12 * fabs(c)) {
        Eeq = 0.5 * (a + b);
        found = true;
      }
    }
    if (!found) return -11;
  }
  if (s0 < 0.0 && Eeq <= Ecut * 1.0001) return -12;
  // ---- reached within dt?  time to come within 1e-6 of E_eq -----------
  const double w0 = log(fabs(Eeq - E0));
  const double wend = log(1e-6 * Eeq);
  if (!(wend < w0)) {                 // started next to the equilibrium
    Ef = Eeq;
    return 1;
  }
  // Only t_reach <= dt matters here, so a loose tolerance suffices when
  // the margin is large; otherwise the reach time is recomputed tightly.
  double treach;
  double xb;
  if (t_adapt(k, R, 1, Eeq, s0, w0, wend, REACH_TOL * dt, treach, xb) &&
      treach <= (1.0 - 2.0 * REACH_TOL) * dt) {
    Ef = Eeq;
    return 1;
  }
  if (!t_adapt(k, R, 1, Eeq, s0, w0, wend, tol, treach, xb)) return -13;
  if (treach <= dt) {
    Ef = Eeq;
    return 1;
  }
  // ---- not reached: Newton in w for t(w) = dt --------------------------
  double w = (s0 * (Eeq - Eg) > 0.0 && s0 * (Eg - E0) > 0.0) ? log(fabs(Eeq - Eg)) : 0.5 * (w0 + wend);
  double t;
  if (!t_adapt(k, R, 1, Eeq, s0, w0, w, tol, t, xb)) return -14;
  for (int it = 0; it < REFINE_ITERS; it++) {
    const double ew = exp(w);
    const double fw = rate(k, R, Eeq - s0 * ew);
    if (!(fw * s0 > 0.0)) return -15;
    const double dtdw = (-s0 * ew) / fw;     // < 0
    double dw = -(t - dt) / dtdw;
    double wn = w + dw;
    if (wn > w0) wn = 0.5 * (w + w0);
    if (wn < wend) wn = 0.5 * (w + wend);
    double dtinc;
    if (!t_adapt(k, R, 1, Eeq, s0, w, wn, tol, dtinc, xb)) return -16;
    t += dtinc;
    w = wn;
    if (fabs(dw) <= 1e-7) {
      Ef = Eeq - s0 * exp(w);
      return 1;
    }
  }
  return -20;
}

}  // namespace wind
}  // namespace pmx
