// Device-side HD arithmetic for the fused stage kernels (gfx950, FP64).
//
// Every expression keeps the operation order of the reference so that the
// GPU result is bit-identical to PLUTO's CPU build (the library is compiled
// with -ffp-contract=off, FP64 division and sqrt are IEEE correctly rounded
// on CDNA).  The functions work on 5-component primitive / conservative
// tuples (rho, v1, v2, v3, p) / (rho, m1, m2, m3, E) held in registers.
#pragma once
#include <hip/hip_runtime.h>

namespace pmx {

enum { RHO = 0, VX1 = 1, VX2 = 2, VX3 = 3, PRS = 4, ENG = 4, NV = 5 };
enum { REC_PLM_VL = 0, REC_PLM_MC = 1, REC_PLM_MM = 2, REC_PLM_DEFAULT = 3,
       REC_PLM_FLAT = 4, REC_PPM = 5 };
enum { SOL_HLLC = 0, SOL_ROE = 1 };

__host__ __device__ __forceinline__ double dmax(double a, double b) { return a >= b ? a : b; }  // macros.h:147
__host__ __device__ __forceinline__ double dmin(double a, double b) { return a <= b ? a : b; }  // macros.h:150
__device__ __forceinline__ double absmin(double a, double b) {                        // macros.h:153
  return fabs(a) < fabs(b) ? a : b;
}
__device__ __forceinline__ double minmod(double a, double b) {                        // macros.h:160
  return a * b > 0.0 ? (fabs(a) < fabs(b) ? a : b) : 0.0;
}

// Src/HD/mappers.c:44-57 (EOS IDEAL)
__device__ __forceinline__ void prim_to_cons(const double* v, double* u, double gmm1) {
  double rho = v[RHO];
  u[RHO] = rho;
  u[VX1] = rho * v[VX1];
  u[VX2] = rho * v[VX2];
  u[VX3] = rho * v[VX3];
  double e = v[VX1] * v[VX1] + v[VX2] * v[VX2] + v[VX3] * v[VX3];
  u[ENG] = 0.5 * rho * e + v[PRS] / gmm1;
}

// Src/HD/mappers.c:113-205 (EOS IDEAL, no entropy switch).  Floors the
// conserved state in place exactly like the reference; returns a bit mask
// (1 = rho<0, 2 = E<0, 4 = p<0).
__device__ __forceinline__ int cons_to_prim(double* u, double* v, double gmm1) {
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
  v[PRS] = gmm1 * (u[ENG] - kin);
  if (v[PRS] < 0.0) {
    v[PRS] = small_p;
    u[ENG] = v[PRS] / gmm1 + kin;
    fail |= 4;
  }
  return fail;
}

// ---------------------------------------------------------------------
// PLM slope limiters on a uniform Cartesian grid (plm_coeffs.h:71-126)
// ---------------------------------------------------------------------
template <int REC>
__device__ __forceinline__ double plm_slope(int nv, double dvp, double dvm) {
  int lim = REC;
  if (REC == REC_PLM_DEFAULT)  // plm_states.c:202-236: MC(rho) VL(v) MM(p)
    lim = (nv == RHO) ? REC_PLM_MC : (nv == PRS ? REC_PLM_MM : REC_PLM_VL);
  if (lim == REC_PLM_VL) return dvp * dvm > 0.0 ? 2.0 * dvp * dvm / (dvp + dvm) : 0.0;
  if (lim == REC_PLM_MM) return dvp * dvm > 0.0 ? absmin(dvp, dvm) : 0.0;
  if (lim == REC_PLM_MC) {
    if (dvp * dvm > 0.0) {
      double qc = 0.5 * (dvm + dvp), scrh = 2.0 * absmin(dvp, dvm);
      return absmin(qc, scrh);
    }
    return 0.0;
  }
  return 0.0;
}

// PLM states of one cell from v[c-1], v[c], v[c+1] (plm_states.c:141-258,
// UNIFORM_CARTESIAN_GRID: dp = dm = 1/2).  vp = right face, vm = left face.
template <int REC>
__device__ __forceinline__ void plm_states(const double* vl, const double* vc,
                                           const double* vr, double* vp, double* vm) {
#pragma unroll
  for (int nv = 0; nv < NV; nv++) {
    double dl = plm_slope<REC>(nv, vr[nv] - vc[nv], vc[nv] - vl[nv]);
    vp[nv] = vc[nv] + dl * 0.5;
    vm[nv] = vc[nv] - dl * 0.5;
  }
}

// PPM (PPM_ORDER 4) interface value at c+1/2 from v[c-1..c+2], bounded by
// minmod against v[c+1]-v[c] (ppm_states.c:150-161; uniform Cartesian
// weights ppm_coeffs.c:490-494).
__device__ __forceinline__ double ppm_iface(double a, double b, double c, double d) {
  const double w0 = -1.0 / 12.0, w1 = 7.0 / 12.0, w2 = 7.0 / 12.0, w3 = -1.0 / 12.0;
  double q = w0 * a + w1 * b + w2 * c + w3 * d;
  double dv = c - b;
  double dvp = q - b;
  return b + minmod(dvp, dv);
}

// PPM parabolic limiter for one cell (ppm_states.c:199-215, Cartesian:
// cm = cp = (3+1)/(3-1) = 2).  Ip = interface value at c+1/2, Im at c-1/2.
__device__ __forceinline__ void ppm_limit(double v, double Ip, double Im, double& vp, double& vm) {
  const double cm = (3.0 + 1.0) / (3.0 - 1.0), cp = (3.0 + 1.0) / (3.0 - 1.0);
  double dvp = Ip - v, dvm = Im - v;
  if (dvp * dvm >= 0.0) {
    dvp = dvm = 0.0;
  } else {
    if (fabs(dvp) >= cm * fabs(dvm)) dvp = -cm * dvm;
    else if (fabs(dvm) >= cp * fabs(dvp)) dvm = -cp * dvp;
  }
  vp = v + dvp;
  vm = v + dvm;
}

// ---------------------------------------------------------------------
// Riemann solvers at one interface in direction DIR.
// Output: flux[5] (normal momentum flux WITHOUT pressure), press, cmax.
// ---------------------------------------------------------------------
template <int DIR>
__device__ __forceinline__ void phys_flux(const double* v, const double* u, double* f) {
  // Src/HD/fluxes.c:36-44
  f[RHO] = u[VX1 + DIR];
  f[VX1] = u[VX1] * v[VX1 + DIR];
  f[VX2] = u[VX2] * v[VX1 + DIR];
  f[VX3] = u[VX3] * v[VX1 + DIR];
  f[ENG] = (u[ENG] + v[PRS]) * v[VX1 + DIR];
}

// Src/HD/hllc.c:57-179, HLL_Speed DAVIS estimate hll_speed.c:76-84,
// MaxSignalSpeed eigenv.c:45-65, SoundSpeed2 EOS/Ideal/eos.c:33.
// Only the star state on the side that is selected is formed: the other
// one never reaches the output.  BRANCHLESS forms it with the same
// operations on every path and selects at the end (bit-identical): the
// Cartesian stage takes that form (Sedov 256^3: 4247-4249 vs 4205-4217
// Mcell-updates/s), the wind stage the branching one (the other form is
// 1% slower there).
#ifndef PMX_HLLC_BRANCHLESS
#define PMX_HLLC_BRANCHLESS 0
#endif
template <int DIR, bool BRANCHLESS = (PMX_HLLC_BRANCHLESS != 0)>
__device__ __forceinline__ void hllc(const double* vL, const double* vR, double gamma,
                                     double gmm1, double* flux, double& press, double& cmax) {
  const int n = VX1 + DIR;
  double uL[NV], uR[NV];
  prim_to_cons(vL, uL, gmm1);
  prim_to_cons(vR, uR, gmm1);
  double a2L = gamma * vL[PRS] / vL[RHO];
  double a2R = gamma * vR[PRS] / vR[RHO];
  double aL = sqrt(a2L), aR = sqrt(a2R);
  double SL = dmin(vL[n] - aL, vR[n] - aR);
  double SR = dmax(vL[n] + aL, vR[n] + aR);
  cmax = dmax(fabs(SL), fabs(SR));
  if (BRANCHLESS) {
    // Same operations on every path, selected at the end: a supersonic
    // interface takes the upwind physical flux (hllc.c:75-84), the others
    // the star flux of the side the contact leaves behind.
    const bool supL = SL > 0.0, supR = !supL && SR < 0.0;
    double qL = vL[PRS] + uL[n] * (vL[n] - SL);
    double qR = vR[PRS] + uR[n] * (vR[n] - SR);
    double wL = vL[RHO] * (vL[n] - SL);
    double wR = vR[RHO] * (vR[n] - SR);
    double vs = (qR - qL) / (wR - wL);
    const bool left = supL || (!supR && vs >= 0.0);
    double v[NV], u[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) {
      v[k] = left ? vL[k] : vR[k];
      u[k] = left ? uL[k] : uR[k];
    }
    double S = left ? SL : SR;
    double vx = v[n];
    double f[NV];
    phys_flux<DIR>(v, u, f);
    double us[NV];
    us[RHO] = u[RHO] * (S - vx) / (S - vs);
#pragma unroll
    for (int k = VX1; k <= VX3; k++) us[k] = us[RHO] * (k == n ? vs : v[k]);
    us[ENG] = u[ENG] / v[RHO] + (vs - vx) * (vs + v[PRS] / (v[RHO] * (S - vx)));
    us[ENG] *= us[RHO];
    const bool sup = supL || supR;
#pragma unroll
    for (int k = 0; k < NV; k++) {
      double fs = f[k] + S * (us[k] - u[k]);
      flux[k] = sup ? f[k] : fs;
    }
    press = v[PRS];
    return;
  }
  if (SL > 0.0) {
    phys_flux<DIR>(vL, uL, flux);
    press = vL[PRS];
    return;
  }
  if (SR < 0.0) {
    phys_flux<DIR>(vR, uR, flux);
    press = vR[PRS];
    return;
  }
  double qL = vL[PRS] + uL[n] * (vL[n] - SL);
  double qR = vR[PRS] + uR[n] * (vR[n] - SR);
  double wL = vL[RHO] * (vL[n] - SL);
  double wR = vR[RHO] * (vR[n] - SR);
  double vs = (qR - qL) / (wR - wL);
  // select side, then form that side's star state and flux
  const bool left = vs >= 0.0;
  const double* v = left ? vL : vR;
  const double* u = left ? uL : uR;
  double S = left ? SL : SR;
  double vx = v[n];
  double f[NV];
  phys_flux<DIR>(v, u, f);
  double us[NV];
  us[RHO] = u[RHO] * (S - vx) / (S - vs);
#pragma unroll
  for (int k = VX1; k <= VX3; k++) us[k] = us[RHO] * (k == n ? vs : v[k]);
  us[ENG] = u[ENG] / v[RHO] + (vs - vx) * (vs + v[PRS] / (v[RHO] * (S - vx)));
  us[ENG] *= us[RHO];
#pragma unroll
  for (int k = 0; k < NV; k++) flux[k] = f[k] + S * (us[k] - u[k]);
  press = v[PRS];
}

// Src/HD/roe.c:86-345 (ROE_AVERAGE YES, EOS IDEAL).  The (n, t, b)
// component order is the cyclic one of Src/set_indexes.c.
template <int DIR, int NDIM>
__device__ __forceinline__ void roe(const double* ql, const double* qr, double gamma,
                                    double gmm1, double* flux, double& press, double& cmax) {
  const int vn = VX1 + DIR;
  const int t = VX1 + (DIR + 1) % 3;
  const int b = VX1 + (DIR + 2) % 3;
  const double delta = 1.e-7;
  double gmm1_inv = 1.0 / gmm1;
  double uL[NV], uR[NV], fL[NV], fR[NV];
  prim_to_cons(ql, uL, gmm1);
  prim_to_cons(qr, uR, gmm1);
  double a2L = gamma * ql[PRS] / ql[RHO];
  double a2R = gamma * qr[PRS] / qr[RHO];
  phys_flux<DIR>(ql, uL, fL);
  phys_flux<DIR>(qr, uR, fR);
  double dv[NV], um[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) dv[k] = qr[k] - ql[k];
  double s = sqrt(qr[RHO] / ql[RHO]);
  um[RHO] = ql[RHO] * s;
  s = 1.0 / (1.0 + s);
  double c = 1.0 - s;
  um[VX1] = s * ql[VX1] + c * qr[VX1];
  um[VX2] = s * ql[VX2] + c * qr[VX2];
  um[VX3] = s * ql[VX3] + c * qr[VX3];
  double vel2 = um[VX1] * um[VX1] + um[VX2] * um[VX2] + um[VX3] * um[VX3];
  double hl = 0.5 * (ql[VX1] * ql[VX1] + ql[VX2] * ql[VX2] + ql[VX3] * ql[VX3]);
  hl += a2L * gmm1_inv;
  double hr = 0.5 * (qr[VX1] * qr[VX1] + qr[VX2] * qr[VX2] + qr[VX3] * qr[VX3]);
  hr += a2R * gmm1_inv;
  double h = s * hl + c * hr;
  double a2 = gmm1 * (h - 0.5 * vel2);
  double a = sqrt(a2);

  double lambda[NV], eta[NV], Rc[NV][NV];
#pragma unroll
  for (int i = 0; i < NV; i++)
#pragma unroll
    for (int j = 0; j < NV; j++) Rc[i][j] = 0.0;
  lambda[0] = um[vn] - a;
  eta[0] = 0.5 / a2 * (dv[PRS] - dv[vn] * um[RHO] * a);
  Rc[RHO][0] = 1.0; Rc[vn][0] = um[vn] - a; Rc[t][0] = um[t]; Rc[b][0] = um[b];
  Rc[ENG][0] = h - um[vn] * a;
  lambda[1] = um[vn] + a;
  eta[1] = 0.5 / a2 * (dv[PRS] + dv[vn] * um[RHO] * a);
  Rc[RHO][1] = 1.0; Rc[vn][1] = um[vn] + a; Rc[t][1] = um[t]; Rc[b][1] = um[b];
  Rc[ENG][1] = h + um[vn] * a;
  lambda[2] = um[vn];
  eta[2] = dv[RHO] - dv[PRS] / a2;
  Rc[RHO][2] = 1.0; Rc[VX1][2] = um[VX1]; Rc[VX2][2] = um[VX2]; Rc[VX3][2] = um[VX3];
  Rc[ENG][2] = 0.5 * vel2;
  lambda[3] = um[vn];
  eta[3] = um[RHO] * dv[t];
  Rc[t][3] = 1.0; Rc[ENG][3] = um[t];
  lambda[4] = um[vn];
  eta[4] = um[RHO] * dv[b];
  Rc[b][4] = 1.0; Rc[ENG][4] = um[b];

  cmax = fabs(um[vn]) + a;

  if (NDIM > 1) {
    double scrh = fabs(ql[PRS] - qr[PRS]);
    scrh /= dmin(ql[PRS], qr[PRS]);
    if (scrh > 0.5 && (qr[vn] < ql[vn])) {
      double bmin = dmin(0.0, lambda[0]);
      double bmax = dmax(0.0, lambda[1]);
      double scrh1 = 1.0 / (bmax - bmin);
#pragma unroll
      for (int k = 0; k < NV; k++) {
        flux[k] = bmin * bmax * (uR[k] - uL[k]) + bmax * fL[k] - bmin * fR[k];
        flux[k] *= scrh1;
      }
      press = (bmax * ql[PRS] - bmin * qr[PRS]) * scrh1;
      return;
    }
  }
  double alambda[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) alambda[k] = fabs(lambda[k]);
  if (alambda[0] <= delta) alambda[0] = 0.5 * lambda[0] * lambda[0] / delta + 0.5 * delta;
  if (alambda[1] <= delta) alambda[1] = 0.5 * lambda[1] * lambda[1] / delta + 0.5 * delta;
#pragma unroll
  for (int nv = 0; nv < NV; nv++) {
    double f = fL[nv] + fR[nv];
#pragma unroll
    for (int k = NV - 1; k >= 0; k--) f -= alambda[k] * eta[k] * Rc[nv][k];
    flux[nv] = f * 0.5;
  }
  press = 0.5 * (ql[PRS] + qr[PRS]);
}

template <int SOLVER, int DIR, int NDIM>
__device__ __forceinline__ void riemann(const double* vL, const double* vR, double gamma,
                                        double gmm1, double* flux, double& press, double& cmax) {
  if (SOLVER == SOL_ROE) roe<DIR, NDIM>(vL, vR, gamma, gmm1, flux, press, cmax);
  else hllc<DIR, true>(vL, vR, gamma, gmm1, flux, press, cmax);
}

}  // namespace pmx
