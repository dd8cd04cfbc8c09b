/* Oracle (test infrastructure only): CPU restatement of the reference's
 * HD method-of-lines step on Cartesian grids.  See oracle.h.
 *
 * Structure follows the reference's call graph (one 1-D pencil at a time):
 *   run()            Src/main.c:215-337   (main loop, NextTimeStep :521-697)
 *   advance_step()   Src/Time_Stepping/rk_step.c:29-384
 *   update_stage()   Src/Time_Stepping/update_stage.c:35-410
 *   states_plm()     Src/States/plm_states.c:83-338, plm_coeffs.h:71-180
 *   states_ppm()     Src/States/ppm_states.c:68-270 (PPM_ORDER 4)
 *   hllc()           Src/HD/hllc.c:28-180, hll_speed.c:24-91, fluxes.c:23-51,
 *                    eigenv.c:34-66, EOS/Ideal/eos.c:16-42
 *   roe()            Src/HD/roe.c:42-346 (ROE_AVERAGE YES)
 *   rhs_cartesian()  Src/MHD/rhs.c:188-209
 *   boundary()       Src/boundary.c:56-330, 617-763
 *   p2c()/c2p()      Src/HD/mappers.c:26-96, 98-281; Src/mappers3D.c
 * The floating-point operations are written in the same order as the
 * reference so that results are identical bit for bit on x86-64 (no FMA
 * contraction: the Makefile builds with -ffp-contract=off).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#include "hd_internal.h"

#define NV 5
#define RHO 0
#define VX1 1
#define VX2 2
#define VX3 3
#define PRS 4
#define ENG 4

#define MAXD(a, b) ((a) >= (b) ? (a) : (b))   /* Src/macros.h:147 */
#define MIND(a, b) ((a) <= (b) ? (a) : (b))   /* Src/macros.h:150 */
#define ABSMIN(a, b) (fabs(a) < fabs(b) ? (a) : (b))       /* macros.h:153 */
#define MINMOD(a, b) ((a)*(b) > 0.0 ? (fabs(a) < fabs(b) ? (a) : (b)) : 0.0)

static const double SMALL_DENSITY = 1.e-12;   /* Src/globals.h:115 */
static const double SMALL_PRESSURE = 1.e-12;  /* Src/globals.h:116 */

typedef struct {
  const orc_cfg *c;
  int ng;            /* ghost zones of active directions            */
  int n[3];          /* interior cells                              */
  int nt[3];         /* total cells incl. ghosts                    */
  int beg[3];        /* first interior index                        */
  int ndim;
  double *V;         /* [nv][k][j][i] with ghosts                   */
  double *U;         /* [k][j][i][nv] with ghosts (Uc layout)       */
  double *U0;
  double *Cdt;       /* [k][j][i]                                   */
  double *dx[3], *inv_dx[3];
  double invDt_hyp;
  double maxMach;
  int stage;
} St;

#define VIDX(s, nv, k, j, i) \
  ((((size_t)(nv)*(s)->nt[2] + (k))*(s)->nt[1] + (j))*(s)->nt[0] + (i))
#define CIDX(s, k, j, i) ((((size_t)(k))*(s)->nt[1] + (j))*(s)->nt[0] + (i))

/* ------------------------------------------------------------------ */
/* prim <-> cons, Src/HD/mappers.c:26-96, 98-281                         */
/* ------------------------------------------------------------------ */
void orc_prim_to_cons(double gamma, const double *v, double *u)
{
  double gmm1 = gamma - 1.0;
  double rho = v[RHO];
  u[RHO] = rho;
  u[VX1] = rho*v[VX1];
  u[VX2] = rho*v[VX2];
  u[VX3] = rho*v[VX3];
  u[ENG] = v[VX1]*v[VX1] + v[VX2]*v[VX2] + v[VX3]*v[VX3];
  u[ENG] = 0.5*rho*u[ENG] + v[PRS]/gmm1;
}

int orc_cons_to_prim(double gamma, double *u, double *v)
{
  double gmm1 = gamma - 1.0, m2, tau, rho, kin;
  int fail = 0;
  m2 = u[VX1]*u[VX1] + u[VX2]*u[VX2] + u[VX3]*u[VX3];
  if (u[RHO] < 0.0) { u[RHO] = SMALL_DENSITY; fail = 1; }
  v[RHO] = rho = u[RHO];
  tau = 1.0/u[RHO];
  v[VX1] = u[VX1]*tau;
  v[VX2] = u[VX2]*tau;
  v[VX3] = u[VX3]*tau;
  kin = 0.5*m2/u[RHO];
  if (u[ENG] < 0.0) { u[ENG] = SMALL_PRESSURE/gmm1 + kin; fail |= 2; }
  v[PRS] = gmm1*(u[ENG] - kin);
  if (v[PRS] < 0.0) {
    v[PRS] = SMALL_PRESSURE;
    u[ENG] = v[PRS]/gmm1 + kin;
    fail |= 4;
  }
  (void)rho;
  return fail;
}

/* ------------------------------------------------------------------ */
/* Riemann solvers on one interface                                    */
/* ------------------------------------------------------------------ */
void orc_phys_flux(int n, const double *v, const double *u, double *f)
{
  /* Src/HD/fluxes.c:36-44: normal momentum flux carries no pressure */
  f[RHO] = u[VX1 + n];
  f[VX1] = u[VX1]*v[VX1 + n];
  f[VX2] = u[VX2]*v[VX1 + n];
  f[VX3] = u[VX3]*v[VX1 + n];
  f[ENG] = (u[ENG] + v[PRS])*v[VX1 + n];
}

void orc_hllc(double gamma, int n, const double *vL, const double *uL,
                 const double *vR, const double *uR, double *flux,
                 double *press, double *cmax, double *mach)
{
  /* Src/HD/hllc.c:57-179 with HLL_Speed (DAVIS) hll_speed.c:76-91 */
  double a2L = gamma*vL[PRS]/vL[RHO];
  double a2R = gamma*vR[PRS]/vR[RHO];
  double fL[NV], fR[NV], us[NV];
  double aL, aR, SL, SR, scrh, vxl, vxr, qL, qR, wL, wR, vs;
  int nv;

  orc_phys_flux(n, vL, uL, fL);
  orc_phys_flux(n, vR, uR, fR);

  aL = sqrt(a2L);
  aR = sqrt(a2R);
  {
    double sl_min = vL[VX1 + n] - aL, sl_max = vL[VX1 + n] + aL;
    double sr_min = vR[VX1 + n] - aR, sr_max = vR[VX1 + n] + aR;
    SL = MIND(sl_min, sr_min);
    SR = MAXD(sl_max, sr_max);
  }
  scrh = fabs(vL[VX1 + n]) + fabs(vR[VX1 + n]);
  scrh /= aL + aR;
  *mach = MAXD(scrh, *mach);

  *cmax = MAXD(fabs(SL), fabs(SR));

  if (SL > 0.0) {
    for (nv = 0; nv < NV; nv++) flux[nv] = fL[nv];
    *press = vL[PRS];
    return;
  }
  if (SR < 0.0) {
    for (nv = 0; nv < NV; nv++) flux[nv] = fR[nv];
    *press = vR[PRS];
    return;
  }
  vxr = vR[VX1 + n];
  vxl = vL[VX1 + n];
  qL = vL[PRS] + uL[VX1 + n]*(vL[VX1 + n] - SL);
  qR = vR[PRS] + uR[VX1 + n]*(vR[VX1 + n] - SR);
  wL = vL[RHO]*(vL[VX1 + n] - SL);
  wR = vR[RHO]*(vR[VX1 + n] - SR);
  vs = (qR - qL)/(wR - wL);

  if (vs >= 0.0) {
    us[RHO] = uL[RHO]*(SL - vxl)/(SL - vs);
    for (nv = VX1; nv <= VX3; nv++)
      us[nv] = us[RHO]*(nv == VX1 + n ? vs : vL[nv]);
    us[ENG] = uL[ENG]/vL[RHO]
              + (vs - vxl)*(vs + vL[PRS]/(vL[RHO]*(SL - vxl)));
    us[ENG] *= us[RHO];
    for (nv = 0; nv < NV; nv++) flux[nv] = fL[nv] + SL*(us[nv] - uL[nv]);
    *press = vL[PRS];
  } else {
    us[RHO] = uR[RHO]*(SR - vxr)/(SR - vs);
    for (nv = VX1; nv <= VX3; nv++)
      us[nv] = us[RHO]*(nv == VX1 + n ? vs : vR[nv]);
    us[ENG] = uR[ENG]/vR[RHO]
              + (vs - vxr)*(vs + vR[PRS]/(vR[RHO]*(SR - vxr)));
    us[ENG] *= us[RHO];
    for (nv = 0; nv < NV; nv++) flux[nv] = fR[nv] + SR*(us[nv] - uR[nv]);
    *press = vR[PRS];
  }
}

void orc_roe(double gamma, int ndim, int n, const double *ql,
                const double *uL, const double *qr, const double *uR,
                double *flux, double *press, double *cmax, double *mach)
{
  /* Src/HD/roe.c:86-345, ROE_AVERAGE == YES, EOS IDEAL */
  const double delta = 1.e-7;
  double gmm1 = gamma - 1.0, gmm1_inv = 1.0/gmm1;
  double a2L = gamma*ql[PRS]/ql[RHO];
  double a2R = gamma*qr[PRS]/qr[RHO];
  double fL[NV], fR[NV], dv[NV], um[NV], eta[NV], lambda[NV], alambda[NV];
  double Rc[NV][NV];
  double s, c, hl, hr, h, vel2, a2, a, scrh;
  int nv, k, t, b;
  int vn = VX1 + n;

  t = VX1 + (n + 1)%3;                 /* Src/set_indexes.c:26-80: cyclic */
  b = VX1 + (n + 2)%3;                 /* (n,t,b) = x-y-z, y-z-x, z-x-y   */

  memset(Rc, 0, sizeof(Rc));
  orc_phys_flux(n, ql, uL, fL);
  orc_phys_flux(n, qr, uR, fR);

  for (nv = NV; nv--;) dv[nv] = qr[nv] - ql[nv];
  s = sqrt(qr[RHO]/ql[RHO]);
  um[RHO] = ql[RHO]*s;
  s = 1.0/(1.0 + s);
  c = 1.0 - s;
  um[VX1] = s*ql[VX1] + c*qr[VX1];
  um[VX2] = s*ql[VX2] + c*qr[VX2];
  um[VX3] = s*ql[VX3] + c*qr[VX3];
  vel2 = um[VX1]*um[VX1] + um[VX2]*um[VX2] + um[VX3]*um[VX3];
  hl = 0.5*(ql[VX1]*ql[VX1] + ql[VX2]*ql[VX2] + ql[VX3]*ql[VX3]);
  hl += a2L*gmm1_inv;
  hr = 0.5*(qr[VX1]*qr[VX1] + qr[VX2]*qr[VX2] + qr[VX3]*qr[VX3]);
  hr += a2R*gmm1_inv;
  h = s*hl + c*hr;
  a2 = gmm1*(h - 0.5*vel2);
  a = sqrt(a2);

  lambda[0] = um[vn] - a;
  eta[0] = 0.5/a2*(dv[PRS] - dv[vn]*um[RHO]*a);
  Rc[RHO][0] = 1.0; Rc[vn][0] = um[vn] - a; Rc[t][0] = um[t]; Rc[b][0] = um[b];
  Rc[ENG][0] = h - um[vn]*a;

  lambda[1] = um[vn] + a;
  eta[1] = 0.5/a2*(dv[PRS] + dv[vn]*um[RHO]*a);
  Rc[RHO][1] = 1.0; Rc[vn][1] = um[vn] + a; Rc[t][1] = um[t]; Rc[b][1] = um[b];
  Rc[ENG][1] = h + um[vn]*a;

  lambda[2] = um[vn];
  eta[2] = dv[RHO] - dv[PRS]/a2;
  Rc[RHO][2] = 1.0; Rc[VX1][2] = um[VX1]; Rc[VX2][2] = um[VX2];
  Rc[VX3][2] = um[VX3]; Rc[ENG][2] = 0.5*vel2;

  lambda[3] = um[vn];
  eta[3] = um[RHO]*dv[t];
  Rc[t][3] = 1.0; Rc[ENG][3] = um[t];

  lambda[4] = um[vn];
  eta[4] = um[RHO]*dv[b];
  Rc[b][4] = 1.0; Rc[ENG][4] = um[b];

  *cmax = fabs(um[vn]) + a;
  *mach = MAXD(fabs(um[vn]/a), *mach);

  if (ndim > 1) {
    scrh = fabs(ql[PRS] - qr[PRS]);
    scrh /= MIND(ql[PRS], qr[PRS]);
    if (scrh > 0.5 && (qr[vn] < ql[vn])) {
      double bmin = MIND(0.0, lambda[0]);
      double bmax = MAXD(0.0, lambda[1]);
      double scrh1 = 1.0/(bmax - bmin);
      for (nv = NV; nv--;) {
        flux[nv] = bmin*bmax*(uR[nv] - uL[nv]) + bmax*fL[nv] - bmin*fR[nv];
        flux[nv] *= scrh1;
      }
      *press = (bmax*ql[PRS] - bmin*qr[PRS])*scrh1;
      return;
    }
  }

  for (nv = NV; nv--;) alambda[nv] = fabs(lambda[nv]);
  if (alambda[0] <= delta) alambda[0] = 0.5*lambda[0]*lambda[0]/delta + 0.5*delta;
  if (alambda[1] <= delta) alambda[1] = 0.5*lambda[1]*lambda[1]/delta + 0.5*delta;
  for (nv = NV; nv--;) {
    flux[nv] = fL[nv] + fR[nv];
    for (k = NV; k--;) flux[nv] -= alambda[k]*eta[k]*Rc[nv][k];
    flux[nv] *= 0.5;
  }
  *press = 0.5*(ql[PRS] + qr[PRS]);
}

void orc_riemann(int solver, double gamma, int dir, const double *vL,
                 const double *vR, double *flux, double *press, double *cmax)
{
  double uL[NV], uR[NV], mach = 0.0;
  orc_prim_to_cons(gamma, vL, uL);
  orc_prim_to_cons(gamma, vR, uR);
  if (solver == ORC_ROE)
    orc_roe(gamma, 3, dir, vL, uL, vR, uR, flux, press, cmax, &mach);
  else
    orc_hllc(gamma, dir, vL, uL, vR, uR, flux, press, cmax, &mach);
}

/* ------------------------------------------------------------------ */
/* Reconstruction on a pencil v[0..ntot-1][NV]                          */
/* vp[i] = state at the right face of cell i, vm[i] = left face          */
/* ------------------------------------------------------------------ */
static double lim_mm(double dvp, double dvm)
{
  return (dvp*dvm > 0.0 ? ABSMIN(dvp, dvm) : 0.0);   /* plm_coeffs.h:77 */
}
static double lim_vl(double dvp, double dvm)
{
  return (dvp*dvm > 0.0 ? 2.0*dvp*dvm/(dvp + dvm) : 0.0); /* plm_coeffs.h:114 */
}
static double lim_mc(double dvp, double dvm)
{
  double qc, scrh;                                       /* plm_coeffs.h:119 */
  if (dvp*dvm > 0.0) {
    qc = 0.5*(dvm + dvp);
    scrh = 2.0*ABSMIN(dvp, dvm);
    return ABSMIN(qc, scrh);
  }
  return 0.0;
}

static void states_plm(int limiter, double (*v)[NV], double (*vp)[NV],
                       double (*vm)[NV], int beg, int end)
{
  /* Src/States/plm_states.c:141-258, UNIFORM_CARTESIAN_GRID branch */
  int i, nv;
  for (i = beg; i <= end; i++) {
    for (nv = 0; nv < NV; nv++) {
      double dvp = v[i + 1][nv] - v[i][nv];
      double dvm = v[i][nv] - v[i - 1][nv];
      double dl;
      int lim = limiter;
      if (lim == ORC_LIM_DEFAULT) {
        lim = (nv == RHO) ? ORC_LIM_MC : (nv == PRS ? ORC_LIM_MINMOD : ORC_LIM_VANLEER);
      }
      switch (lim) {
        case ORC_LIM_VANLEER: dl = lim_vl(dvp, dvm); break;
        case ORC_LIM_MC:      dl = lim_mc(dvp, dvm); break;
        case ORC_LIM_MINMOD:  dl = lim_mm(dvp, dvm); break;
        default:              dl = 0.0; break;
      }
      vp[i][nv] = v[i][nv] + dl*0.5;
      vm[i][nv] = v[i][nv] - dl*0.5;
    }
  }
}

static void states_ppm(double (*v)[NV], double (*vp)[NV], double (*vm)[NV],
                       int beg, int end)
{
  /* Src/States/ppm_states.c:132-221 with PPM_ORDER 4 and the uniform
   * Cartesian weights of ppm_coeffs.c:468-494 (hp = hm = 3, so cp=cm=2). */
  const double w0 = -1.0/12.0, w1 = 7.0/12.0, w2 = 7.0/12.0, w3 = -1.0/12.0;
  const double cm = (3.0 + 1.0)/(3.0 - 1.0), cp = (3.0 + 1.0)/(3.0 - 1.0);
  int i, nv;
  for (i = beg - 1; i <= end; i++) {
    for (nv = 0; nv < NV; nv++) {
      double q = w0*v[i - 1][nv] + w1*v[i][nv] + w2*v[i + 1][nv] + w3*v[i + 2][nv];
      double dv = v[i + 1][nv] - v[i][nv];
      double dvp = q - v[i][nv];
      vp[i][nv] = v[i][nv] + MINMOD(dvp, dv);
    }
  }
  for (i = beg; i <= end; i++)
    for (nv = 0; nv < NV; nv++) vm[i][nv] = vp[i - 1][nv];
  for (i = beg; i <= end; i++) {
    for (nv = 0; nv < NV; nv++) {
      double dvp = vp[i][nv] - v[i][nv];
      double dvm = vm[i][nv] - v[i][nv];
      if (dvp*dvm >= 0.0) dvp = dvm = 0.0;
      else {
        if (fabs(dvp) >= cm*fabs(dvm)) dvp = -cm*dvm;
        else if (fabs(dvm) >= cp*fabs(dvp)) dvm = -cp*dvp;
      }
      vp[i][nv] = v[i][nv] + dvp;
      vm[i][nv] = v[i][nv] + dvm;
    }
  }
}

/* ------------------------------------------------------------------ */
/* Boundary conditions, Src/boundary.c:56-330 (sides in order)         */
/* ------------------------------------------------------------------ */
static void boundary(St *s)
{
  const orc_cfg *c = s->c;
  int side, nv, i, j, k;
  for (side = 0; side < 2*s->ndim; side++) {
    int d = side/2, left = (side%2 == 0);
    int lo[3] = {0, 0, 0}, hi[3];
    int b = s->beg[d], e = s->beg[d] + s->n[d] - 1;
    for (i = 0; i < 3; i++) hi[i] = s->nt[i] - 1;
    if (left) { lo[d] = 0; hi[d] = b - 1; }
    else      { lo[d] = e + 1; hi[d] = s->nt[d] - 1; }
    for (nv = 0; nv < NV; nv++) {
      double sgn = (c->bc[side] == ORC_BC_REFLECTIVE && nv == VX1 + d) ? -1.0 : 1.0;
      for (k = lo[2]; k <= hi[2]; k++)
      for (j = lo[1]; j <= hi[1]; j++)
      for (i = lo[0]; i <= hi[0]; i++) {
        int idx[3] = {i, j, k}, src[3] = {i, j, k};
        int g = idx[d];
        switch (c->bc[side]) {
          case ORC_BC_PERIODIC:   src[d] = left ? g + s->n[d] : g - s->n[d]; break;
          case ORC_BC_OUTFLOW:    src[d] = left ? b : e; break;
          case ORC_BC_REFLECTIVE: src[d] = left ? (2*b - 1 - g) : (2*e + 1 - g); break;
        }
        s->V[VIDX(s, nv, k, j, i)] = sgn*s->V[VIDX(s, nv, src[2], src[1], src[0])];
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* One stage: Src/Time_Stepping/update_stage.c:142-316                  */
/* ------------------------------------------------------------------ */
static void update_stage(St *s, double dt)
{
  const orc_cfg *c = s->c;
  int d, nv, ip, maxn = 0;
  double (*v)[NV], (*vp)[NV], (*vm)[NV], (*flux)[NV], (*rhs)[NV];
  double *press, *cmax;
  for (d = 0; d < 3; d++) if (s->nt[d] > maxn) maxn = s->nt[d];
  v = malloc(sizeof(double)*NV*maxn);
  vp = malloc(sizeof(double)*NV*maxn);
  vm = malloc(sizeof(double)*NV*maxn);
  flux = malloc(sizeof(double)*NV*maxn);
  rhs = malloc(sizeof(double)*NV*maxn);
  press = malloc(sizeof(double)*maxn);
  cmax = malloc(sizeof(double)*maxn);

  if (s->stage == 1) {
    size_t ncell = (size_t)s->nt[0]*s->nt[1]*s->nt[2];
    memset(s->Cdt, 0, sizeof(double)*ncell);
  }

  for (d = 0; d < s->ndim; d++) {
    int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;   /* transverse dirs */
    int ntot = s->nt[d], nbeg = s->beg[d], nend = s->beg[d] + s->n[d] - 1;
    int a, bb;
    for (bb = s->beg[t2]; bb < s->beg[t2] + s->n[t2]; bb++)
    for (a = s->beg[t1]; a < s->beg[t1] + s->n[t1]; a++) {
      int ijk[3];
      ijk[t1] = a; ijk[t2] = bb;
      for (ip = 0; ip < ntot; ip++) {
        ijk[d] = ip;
        for (nv = 0; nv < NV; nv++) v[ip][nv] = s->V[VIDX(s, nv, ijk[2], ijk[1], ijk[0])];
      }
      if (c->recon == ORC_PPM) states_ppm(v, vp, vm, nbeg - 1, nend + 1);
      else states_plm(c->limiter, v, vp, vm, nbeg - 1, nend + 1);

      for (ip = nbeg - 1; ip <= nend; ip++) {
        double uL[NV], uR[NV];
        orc_prim_to_cons(c->gamma, vp[ip], uL);
        orc_prim_to_cons(c->gamma, vm[ip + 1], uR);
        if (c->solver == ORC_ROE)
          orc_roe(c->gamma, s->ndim, d, vp[ip], uL, vm[ip + 1], uR, flux[ip],
              &press[ip], &cmax[ip], &s->maxMach);
        else
          orc_hllc(c->gamma, d, vp[ip], uL, vm[ip + 1], uR, flux[ip], &press[ip],
               &cmax[ip], &s->maxMach);
      }
      for (ip = nbeg; ip <= nend; ip++) {           /* rhs.c:188-209 */
        double scrh = dt/s->dx[d][ip];
        for (nv = 0; nv < NV; nv++) rhs[ip][nv] = -scrh*(flux[ip][nv] - flux[ip - 1][nv]);
        rhs[ip][VX1 + d] -= scrh*(press[ip] - press[ip - 1]);
      }
      for (ip = nbeg; ip <= nend; ip++) {
        double *u;
        ijk[d] = ip;
        u = s->U + NV*CIDX(s, ijk[2], ijk[1], ijk[0]);
        for (nv = 0; nv < NV; nv++) u[nv] += rhs[ip][nv];
      }
      if (s->stage == 1) {                           /* :303-316 */
        for (ip = nbeg; ip <= nend; ip++) {
          ijk[d] = ip;
          s->Cdt[CIDX(s, ijk[2], ijk[1], ijk[0])] +=
              0.5*(cmax[ip - 1] + cmax[ip])*s->inv_dx[d][ip]*1.0;
        }
      }
    }
  }
  if (s->stage == 1) {                               /* :389-394 */
    int i, j, k;
    for (k = s->beg[2]; k < s->beg[2] + s->n[2]; k++)
    for (j = s->beg[1]; j < s->beg[1] + s->n[1]; j++)
    for (i = s->beg[0]; i < s->beg[0] + s->n[0]; i++)
      s->invDt_hyp = MAXD(s->invDt_hyp, s->Cdt[CIDX(s, k, j, i)]);
    s->invDt_hyp /= (double)s->ndim;
  }
  free(v); free(vp); free(vm); free(flux); free(rhs); free(press); free(cmax);
}

static void interior_p2c(St *s, double *U)
{
  int i, j, k, nv;
  for (k = s->beg[2]; k < s->beg[2] + s->n[2]; k++)
  for (j = s->beg[1]; j < s->beg[1] + s->n[1]; j++)
  for (i = s->beg[0]; i < s->beg[0] + s->n[0]; i++) {
    double v[NV];
    for (nv = 0; nv < NV; nv++) v[nv] = s->V[VIDX(s, nv, k, j, i)];
    orc_prim_to_cons(s->c->gamma, v, U + NV*CIDX(s, k, j, i));
  }
}

static int interior_c2p(St *s)
{
  int i, j, k, nv, err = 0;
  for (k = s->beg[2]; k < s->beg[2] + s->n[2]; k++)
  for (j = s->beg[1]; j < s->beg[1] + s->n[1]; j++)
  for (i = s->beg[0]; i < s->beg[0] + s->n[0]; i++) {
    double v[NV];
    if (orc_cons_to_prim(s->c->gamma, s->U + NV*CIDX(s, k, j, i), v) & 1) err = 1;
    for (nv = 0; nv < NV; nv++) s->V[VIDX(s, nv, k, j, i)] = v[nv];
  }
  return err;
}

static void rk_combine(St *s, double w0, double wc, int third)
{
  int i, j, k, nv;
  const double one_third = 1.0/3.0;
  for (k = s->beg[2]; k < s->beg[2] + s->n[2]; k++)
  for (j = s->beg[1]; j < s->beg[1] + s->n[1]; j++)
  for (i = s->beg[0]; i < s->beg[0] + s->n[0]; i++) {
    double *u = s->U + NV*CIDX(s, k, j, i), *u0 = s->U0 + NV*CIDX(s, k, j, i);
    for (nv = 0; nv < NV; nv++) {
      if (third) u[nv] = one_third*(u0[nv] + 2.0*u[nv]);
      else       u[nv] = w0*u0[nv] + wc*u[nv];
    }
  }
}

static int advance_step(St *s, double dt)
{
  /* Src/Time_Stepping/rk_step.c:117-330 */
  int err = 0;
  size_t ncons = (size_t)NV*s->nt[0]*s->nt[1]*s->nt[2];
  s->stage = 1;
  boundary(s);
  interior_p2c(s, s->U);
  memcpy(s->U0, s->U, sizeof(double)*ncons);
  update_stage(s, dt);
  err |= interior_c2p(s);

  s->stage = 2;
  boundary(s);
  update_stage(s, dt);
  if (s->c->rk == 3) rk_combine(s, 0.75, 0.25, 0);
  else rk_combine(s, 0.5, 0.5, 0);
  err |= interior_c2p(s);

  if (s->c->rk == 3) {
    s->stage = 3;
    boundary(s);
    update_stage(s, dt);
    rk_combine(s, 0.0, 0.0, 1);
    err |= interior_c2p(s);
  }
  return err;
}

int orc_nghost(const orc_cfg *c)
{
  return (c->recon == ORC_PPM) ? 3 : 2;   /* Src/get_nghost.c:19-42 */
}

int orc_run(const orc_cfg *c, double *Vin, int nsteps, double *dt, double *t,
            double *U_out, double *dt_hist)
{
  St s;
  int d, nv, i, j, k, step;
  size_t ncell;
  if (c->dims < 2 || c->dims > 3) return -1;
  memset(&s, 0, sizeof(s));
  s.c = c;
  s.ndim = c->dims;
  s.ng = orc_nghost(c);
  for (d = 0; d < 3; d++) {
    int active = d < c->dims;
    s.n[d] = active ? c->nx[d] : 1;
    s.beg[d] = active ? s.ng : 0;
    s.nt[d] = s.n[d] + 2*s.beg[d];
    s.dx[d] = malloc(sizeof(double)*s.nt[d]);
    s.inv_dx[d] = malloc(sizeof(double)*s.nt[d]);
    /* uniform grid: Src/set_grid.c:396-403 (interior) and :190-205 (ghosts
     * copy the adjacent spacing), so every dx equals (xR - xL)/N */
    for (i = 0; i < s.nt[d]; i++) {
      s.dx[d][i] = (c->xend[d] - c->xbeg[d])/(double)(s.n[d]);
      s.inv_dx[d][i] = 1.0/s.dx[d][i];               /* set_geometry.c:305 */
    }
  }
  ncell = (size_t)s.nt[0]*s.nt[1]*s.nt[2];
  s.V = calloc(NV*ncell, sizeof(double));
  s.U = calloc(NV*ncell, sizeof(double));
  s.U0 = calloc(NV*ncell, sizeof(double));
  s.Cdt = calloc(ncell, sizeof(double));

  for (nv = 0; nv < NV; nv++)
  for (k = 0; k < s.n[2]; k++)
  for (j = 0; j < s.n[1]; j++)
  for (i = 0; i < s.n[0]; i++)
    s.V[VIDX(&s, nv, k + s.beg[2], j + s.beg[1], i + s.beg[0])] =
        Vin[(((size_t)nv*s.n[2] + k)*s.n[1] + j)*s.n[0] + i];

  orc_set_quit(ORC_QUIT_NONE);
  for (step = 0; step < nsteps; step++) {            /* Src/main.c:215-337 */
    double dtn, dt_hyp;
    int last = 0;
    if ((*t + *dt) >= c->tstop*(1.0 - 1.e-8)) {     /* main.c:226-229 */
      *dt = (c->tstop - *t);
      last = 1;
    }
    s.invDt_hyp = 0.0;                               /* main.c:406-415 */
    s.maxMach = 0.0;
    if (dt_hist) dt_hist[step] = *dt;
    advance_step(&s, *dt);
    *t += *dt;
    dt_hyp = 1.0/s.invDt_hyp;                        /* main.c:601-609 */
    dt_hyp *= c->cfl;
    dtn = dt_hyp;
    dtn = MIND(dtn, c->cfl_max_var*(*dt));           /* main.c:670 */
    if (c->cfl_max_var != 1.0) *dt = dtn;            /* main.c:695-696 */
    if (dtn < c->first_dt*1.e-9) {                   /* main.c:672-683 */
      orc_set_quit(ORC_QUIT_DT_TOO_SMALL);
      step++;
      break;
    }
    if (last) {
      orc_set_quit(ORC_STOP_TSTOP);
      step++;
      break;
    }
  }

  for (nv = 0; nv < NV; nv++)
  for (k = 0; k < s.n[2]; k++)
  for (j = 0; j < s.n[1]; j++)
  for (i = 0; i < s.n[0]; i++) {
    size_t o = (((size_t)nv*s.n[2] + k)*s.n[1] + j)*s.n[0] + i;
    Vin[o] = s.V[VIDX(&s, nv, k + s.beg[2], j + s.beg[1], i + s.beg[0])];
    if (U_out) U_out[o] = s.U[NV*CIDX(&s, k + s.beg[2], j + s.beg[1], i + s.beg[0]) + nv];
  }
  for (d = 0; d < 3; d++) { free(s.dx[d]); free(s.inv_dx[d]); }
  free(s.V); free(s.U); free(s.U0); free(s.Cdt);
  return step;
}

static int g_quit = ORC_QUIT_NONE;
void orc_set_quit(int why) { g_quit = why; }
int orc_last_quit(void) { return g_quit; }
