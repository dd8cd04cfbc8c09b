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
};

// ---------------------------------------------------------------------
// prim <-> cons with the tracer (Src/HD/mappers.c:65-67, :247-249)
// ---------------------------------------------------------------------
__device__ __forceinline__ void p2c6(const double* v, double* u, double gmm1) {
  prim_to_cons(v, u, gmm1);
  u[TRC] = v[RHO] * v[TRC];
}

__device__ __forceinline__ int c2p6(double* u, double* v, double gmm1) {
  int f = cons_to_prim(u, v, gmm1);
  v[TRC] = u[TRC] * (1.0 / u[RHO]);
  return f;
}

// General-grid van Leer PLM of one cell (Src/States/plm_states.c:141-258
// with plm_coeffs.h:140-142): c = {cp, cm, wp, wm, dp, dm}.
__device__ __forceinline__ void plm_vl_general(const double* vm1, const double* v0,
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

// ---------------------------------------------------------------------
// Line force: one angle bin of LineForce (line_connect.c:873-902) for the
// state (rho_cgs, v_th).  Returns the (1 + M) factor.
// ---------------------------------------------------------------------
__device__ __forceinline__ double force_mult(const Const& k, double rho_cgs, double v_th,
                                             double dvds) {
  const double M_max = 4400.;
  double M;
  if (dvds > 0.0) {
    double t_UV = k.sigma_e * rho_cgs * v_th / dvds;
    M = k.krad * pow(t_UV, k.alpharad);
    if (M > M_max) M = M_max;
  } else {
    M = 0.0;
  }
  return 1.0 + M;
}

// ---------------------------------------------------------------------
// BLONDIN cooling (Src/Cooling/BLONDIN/cooling.c)
// ---------------------------------------------------------------------
struct Blondin {
  double xi, sqxi, sqsqxi, ne, nH, n, tx, E, hc_init, dt_share;
};

__device__ __forceinline__ double ne_rat(double T) {   // cooling.c:210-222
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

// heatcool(), cooling.c:173-202 with every prefactor equal to 1.0 (the
// values read_sirocco_heatcool leaves at start-up)
__device__ __forceinline__ double heatcool(Blondin& b, double T) {
  double sqT = sqrt(T);
  b.ne = b.nH * ne_rat(T);
  double comp_heat = 1.0 * (8.9e-36 * b.xi * b.tx);
  double comp_cool = 1.0 * (8.9e-36 * b.xi * (4.0 * T));
  double xray_heat = 1.0 * (1.5e-21 * (b.sqsqxi / sqT));
  double line_cool = 1.0 * ((1e-16 * exp(-1.3e5 / T) / b.sqxi / T) +
                            fmin(fmin(1e-24, 5e-27 * sqT), 1.5e-17 / T));
  double brem_cool = 1.0 * (3.3e-27 * sqT);
  return b.nH * (b.ne * comp_heat + b.nH * xray_heat - b.ne * comp_cool - b.ne * line_cool -
                 b.ne * brem_cool);
}

__device__ __forceinline__ double zfunc(Blondin& b, double temp) {   // cooling.c:233-237
  return (temp * b.n * C_kB / (2.0 / 3.0)) - b.E - b.dt_share * (b.hc_init + heatcool(b, temp)) / 2.0;
}

template <int WHICH>   // 0: zfunc, 1: heatcool
__device__ __forceinline__ double bfunc(Blondin& b, double x) {
  return WHICH ? heatcool(b, x) : zfunc(b, x);
}

// Brent root finder, cooling.c:250-319 (ITMAX 100, EPS 3e-8)
template <int WHICH>
__device__ double zbrent(Blondin& bl, double x1, double x2, double tol) {
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
__device__ __forceinline__ double blondin_cell(const Const& k, double rho_code, double p,
                                               double r_code, double xi_s, double tr_s,
                                               double t, double dt) {
  const double minT = 1.e4;
  Blondin b;
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
  double t;        // g_time
};

// Radiat(): returns the RHOE rate; may floor *E like the reference.
__device__ __forceinline__ double radiat(const Const& k, const RadCell& R, double& E) {
  const double minT = 1.e4;
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
  double comp_heat = 1.0 * (8.9e-36 * xi * tx);
  double comp_cool = 1.0 * (8.9e-36 * xi * (4.0 * T));
  double xray_heat = 1.0 * (1.5e-21 * (sqsqxi / sqT));
  double line_cool = 1.0 * ((1e-16 * exp(-1.3e5 / T) / sqxi / T) +
                            fmin(fmin(1e-24, 5e-27 * sqT), 1.5e-17 / T));
  double brem_cool = 1.0 * (3.3e-27 * sqT);
  double lambda = nH * (ne * comp_heat + nH * xray_heat - ne * comp_cool - ne * line_cool -
                        ne * brem_cool);
  double rhs = (R.trc > 0.01) ? lambda * k.E_cost : 0.0;
  rhs *= 1.0 - 1.0 / cosh(pow(T / minT, 12));
  return rhs;
}

// SolveODE_RKF12: returns err, writes the 2nd-order result.
__device__ __forceinline__ double rkf12(const Const& k, const RadCell& R, double E0, double k1,
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
__device__ double ck45(const Const& k, const RadCell& R, double E0, double k1, double& E5th,
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

// One cell of CoolingSource: returns the new pressure and the cell's
// dt * maxCoolingRate / |dp/p| candidate for dt_cool.
__device__ __forceinline__ double plain_cell(const Const& k, const RadCell& R, double p,
                                             double dt, double& dt_cool_cell, int& bad) {
  const double min_tol = 2.e-5, minT = 1.e4, maxCoolingRate = 0.1;
  double T0 = p / R.rho * k.KELVIN * k.mu_tab;
  double E0 = p / (k.gamma - 1.0), E1 = E0;
  double k1 = radiat(k, R, E0);
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
  double prs = E1 * (k.gamma - 1.0);
  double T1 = prs / R.rho * k.KELVIN * k.mu_tab;
  if (prs < 0.0) prs = 1.e-12;
  if (T1 < minT && T0 > minT) prs = minT * R.rho / (k.KELVIN * k.mu_tab);
  double e = fabs(prs / p - 1.0);
  dt_cool_cell = dt * maxCoolingRate / e;
  return prs;
}

}  // namespace wind
}  // namespace pmx
