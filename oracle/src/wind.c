/* Oracle (test infrastructure only): CPU restatement of the reference's
 * line-driven disk-wind step on a 2-D spherical (r, theta) grid.
 *
 *   geometry()        Src/set_geometry.c:16-309 (SPHERICAL, 2-D)
 *   plm_coeffs()      Src/States/plm_coeffs.c:30-91
 *   states()          Src/States/plm_states.c:83-338 (general grid, VANLEER_LIM)
 *   riemann + tracer  Src/HD/hllc.c, Src/adv_flux.c:42-52
 *   rhs()             Src/MHD/rhs.c:234-351 (+ TotalFlux :466-560)
 *   sources           Src/MHD/rhs_source.c:181-399 (geometry, BodyForceVector,
 *                     LineForce)
 *   vgrad_setup/vgrad Src/LineDriven/line_connect.c:504-742 (VGradCalc),
 *                     :746-777 (bilinear)
 *   line_force()      Src/LineDriven/line_connect.c:815-903 (k, alpha mode;
 *                     tabulated M(t) with linterp(), :781-811, for
 *                     krad = alpharad = 999)
 *   blondin()         Src/Cooling/BLONDIN/cooling.c:50-319
 *   cooling_plain()   Test_Problems/LineDrivenWind/cv_idl/cooling_source.c
 *                     (CoolingSource, per-cell prefactors :101-105) with
 *                     cv_idl/radiat.c (Radiat, :118-123) and
 *                     Src/Cooling/cooling_ode_solver.c (RKF12, CK45)
 *   run()             Src/main.c:215-337, :383-485 (Strang order, cooling
 *                     dt every second step), :521-697
 * Operation order follows the reference so results agree bit for bit with
 * the compiled reference on the same host (same libm).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#include "hd_internal.h"

#define NVW 6
#define RHO 0
#define VX1 1
#define VX2 2
#define VX3 3
#define PRS 4
#define ENG 4
#define TRC 5
#define RHOE 6            /* NVAR_COOLING slot (Src/Cooling/cooling.h:20) */
#define NVC 7
#define NG 2

/* path counters, for tests that check which branches a case exercises */
enum { CNT_RKF12_OK, CNT_CK45_CALLS, CNT_CK45_SUB, CNT_CK45_REJECT, CNT_BL_SOLVE,
       CNT_BL_WIDEN, CNT_BL_SECOND, CNT_BELOW_MINT, CNT_DT_CFL, CNT_DT_COOL,
       CNT_DT_MAXVAR, CNT_C2P_FAIL, CNT_FLOOR_T, CNT_FLOOR_D, CNT_SHOCK, CNT_MIDPLANE,
       CNT_N };
static long cnt[CNT_N];

int orc_wind_counters(long *out, int n)
{
  int k;
  for (k = 0; k < n && k < CNT_N; k++) out[k] = cnt[k];
  return CNT_N;
}

#define MAXD(a, b) ((a) >= (b) ? (a) : (b))
#define MIND(a, b) ((a) <= (b) ? (a) : (b))

/* Src/pluto.h:296-321 */
#define C_AH 1.008
#define C_AHe 4.004
#define C_AZ 30.0
#define C_amu 1.66053886e-24
#define C_c 2.99792458e10
#define C_G 6.6726e-8
#define C_kB 1.3806505e-16
#define C_mp 1.67262171e-24
#define C_PI 3.14159265358979
#define C_sigma 5.67051e-5
#define C_sigmaT 6.6524e-25

typedef struct {
  const orc_wind_cfg *c;
  int n1, n2, t1, t2;                 /* interior and total sizes           */
  double *V, *U, *U0, *Cdt;           /* V[nv][j][i]; U[j][i][nv]           */
  const double *x1, *x1l, *x1r, *dx1, *x2, *x2l, *x2r, *dx2;
  double *xgc1, *xgc2, *rt, *s, *sp, *inv_dx1, *inv_dx2;
  double *dV, *A1, *A2;               /* [j][i]                             */
  double *pc[2][6];                   /* PLM cp, cm, wp, wm, dp, dm per dir */
  /* line force */
  int na;
  double *fr, *ft, *fp;               /* [a][j][i] (total layout)           */
  double *x11_0, *x11_1, *x22_0, *x22_1;
  double *roff, *toff, *moff, *dvds;  /* [a][j][i]                          */
  int vgrad_setup;
  /* cooling */
  double *xi_s, *tr_s;                /* sirocco_xi, sirocco_t_r            */
  /* units and derived constants */
  double UD, UL, UV, UT, UP, UA, KELVIN;
  double invDt_hyp, dt_cool, t;
  int stage;
  unsigned char *flag;                /* zone flags [j][i] (FLAG_MINMOD, FLAG_HLL) */
  /* ENTROPY_SWITCH: Vc[ENTR] = p / rho^gamma and Uc[ENTR] = rho Vc[ENTR]
   * (pluto.h:737-746), kept beside V and U, [j][i] */
  double *S, *Ue, *U0e;
  /* tabulated force multiplier: log10 t [fm_points], log10 M [p][j][i]
     (interior layout), as read_sirocco_fluxes keeps them */
  double *fm_lt, *fm_lm;
  int mid_row;                        /* the cv_idl midplane row, or -1     */
} W;

enum { FLAG_MINMOD = 1, FLAG_HLL = 4 };   /* Src/pluto.h:212-214 */

#define VI(w, nv, j, i) (((size_t)(nv)*(w)->t2 + (j))*(w)->t1 + (i))
#define CI(w, j, i) ((size_t)(j)*(w)->t1 + (i))
#define AI(w, a, j, i) (((size_t)(a)*(w)->t2 + (j))*(w)->t1 + (i))

/* the five prefactors of cell (j, i) in the order comp_h, comp_c, xray_h,
   brem_c, line_c (cv_idl/cooling_source.c:101-105) */
static void prefactors(const W *w, int j, int i, double *pre)
{
  int p;
  for (p = 0; p < 5; p++)
    pre[p] = w->c->pre ? w->c->pre[((size_t)p*w->n2 + (j - NG))*w->n1 + (i - NG)] : 1.0;
}

/* ------------------------------------------------------------------ */
static void p2c(double gamma, const double *v, double *u)
{
  double gmm1 = gamma - 1.0, rho = v[RHO];
  u[RHO] = rho;
  u[VX1] = rho*v[VX1];
  u[VX2] = rho*v[VX2];
  u[VX3] = rho*v[VX3];
  u[ENG] = v[VX1]*v[VX1] + v[VX2]*v[VX2] + v[VX3]*v[VX3];
  u[ENG] = 0.5*rho*u[ENG] + v[PRS]/gmm1;
  u[TRC] = rho*v[TRC];                         /* mappers.c:65-67 */
}

static int c2p(double gamma, double *u, double *v)
{
  double gmm1 = gamma - 1.0, m2, tau, kin;
  int fail = 0;
  m2 = u[VX1]*u[VX1] + u[VX2]*u[VX2] + u[VX3]*u[VX3];
  if (u[RHO] < 0.0) { u[RHO] = 1.e-12; fail = 1; }
  v[RHO] = u[RHO];
  tau = 1.0/u[RHO];
  v[VX1] = u[VX1]*tau;
  v[VX2] = u[VX2]*tau;
  v[VX3] = u[VX3]*tau;
  kin = 0.5*m2/u[RHO];
  if (u[ENG] < 0.0) { u[ENG] = 1.e-12/gmm1 + kin; fail |= 2; }
  v[PRS] = gmm1*(u[ENG] - kin);
  if (v[PRS] < 0.0) { v[PRS] = 1.e-12; u[ENG] = v[PRS]/gmm1 + kin; fail |= 4; }
  v[TRC] = u[TRC]*tau;                         /* mappers.c:247-249 */
  return fail;
}

/* ConsToPrim with FLAG_ENTROPY (ENTROPY_SWITCH ALWAYS flags every zone,
 * flag_shock.c:140-143): the pressure from the entropy, the energy
 * redefined from it (Src/HD/mappers.c:129-204, :247-249). */
static int c2p_entropy(double gamma, double *u, double *ue, double *v, double *s)
{
  double gmm1 = gamma - 1.0, m2, tau, kin, rhog1;
  int fail = 0;
  m2 = u[VX1]*u[VX1] + u[VX2]*u[VX2] + u[VX3]*u[VX3];
  if (u[RHO] < 0.0) { u[RHO] = 1.e-12; fail = 1; }
  v[RHO] = u[RHO];
  tau = 1.0/u[RHO];
  v[VX1] = u[VX1]*tau;
  v[VX2] = u[VX2]*tau;
  v[VX3] = u[VX3]*tau;
  kin = 0.5*m2/u[RHO];
  if (u[ENG] < 0.0) { u[ENG] = 1.e-12/gmm1 + kin; fail |= 2; }
  rhog1 = pow(v[RHO], gmm1);
  v[PRS] = (*ue)*rhog1;
  if (v[PRS] < 0.0) { v[PRS] = 1.e-12; fail |= 4; }
  u[ENG] = v[PRS]/gmm1 + kin;
  v[TRC] = u[TRC]*tau;
  *s = (*ue)*tau;
  return fail;
}

/* ------------------------------------------------------------------ */
/* grid geometry and reconstruction weights                            */
/* ------------------------------------------------------------------ */
static void geometry(W *w)
{
  int i, j, d;
  for (i = 0; i < w->t1; i++) {
    double xL = w->x1l[i], xR = w->x1r[i], x = w->x1[i], dx = w->dx1[i];
    w->xgc1[i] = x + 2.0*x*dx*dx/(12.0*x*x + dx*dx);
    w->rt[i] = (xR*xR*xR - xL*xL*xL)/(xR*xR - xL*xL)/1.5;
    w->inv_dx1[i] = 1.0/dx;
  }
  for (j = 0; j < w->t2; j++) {
    double xL = w->x2l[j], xR = w->x2r[j];
    w->xgc2[j] = sin(xR) - sin(xL) + xL*cos(xL) - xR*cos(xR);
    w->xgc2[j] /= cos(xL) - cos(xR);
    w->sp[j] = fabs(sin(xR));
    w->s[j] = fabs(sin(w->x2[j]));
    w->inv_dx2[j] = 1.0/w->dx2[j];
  }
  for (j = 0; j < w->t2; j++) {
    for (i = 0; i < w->t1; i++) {
      double x1p = w->x1r[i], x1m = w->x1l[i];
      double dVr = fabs(x1p*x1p*x1p - x1m*x1m*x1m)/3.0;
      double dmu = fabs(cos(w->x2l[j]) - cos(w->x2r[j]));
      w->dV[CI(w, j, i)] = dVr*dmu;
      w->A1[CI(w, j, i)] = x1p*x1p*dmu;                      /* face i+1/2 */
      w->A2[CI(w, j, i)] = fabs(w->x1[i])*w->dx1[i]*fabs(sin(w->x2r[j]));  /* j+1/2 */
    }
  }
  for (d = 0; d < 2; d++) {                    /* Src/States/plm_coeffs.c:65-84 */
    int n = d ? w->t2 : w->t1;
    const double *dx = d ? w->dx2 : w->dx1, *xr = d ? w->x2r : w->x1r;
    const double *xgc = d ? w->xgc2 : w->xgc1;
    for (i = 1; i <= n - 2; i++) {
      w->pc[d][2][i] = dx[i]/(xgc[i + 1] - xgc[i]);
      w->pc[d][3][i] = dx[i]/(xgc[i] - xgc[i - 1]);
      w->pc[d][0][i] = (xgc[i + 1] - xgc[i])/(xr[i] - xgc[i]);
      w->pc[d][1][i] = (xgc[i] - xgc[i - 1])/(xgc[i] - xr[i - 1]);
      w->pc[d][4][i] = (xr[i] - xgc[i])/dx[i];
      w->pc[d][5][i] = (xgc[i] - xr[i - 1])/dx[i];
    }
  }
}

/* ------------------------------------------------------------------ */
/* cv_idl UserDefBoundary(side 0), Test_Problems/LineDrivenWind/cv_idl/  */
/* init.c:196-273: density, pressure and temperature floors on every    */
/* cell (TOT_LOOP), momentum kept by scaling the velocities, the sound   */
/* speed kept by the new pressure; PrimToCons of the changed cells.     */
/* Then the pass's midplane row (:276-311): see midplane_cell.          */
/* ------------------------------------------------------------------ */
static int floor_cell(const W *w, double *v)
{
  const orc_wind_cfg *c = w->c;
  const double dfloor = c->dfloor/w->UD, tfloor = 5.e2;
  const double pfloor = dfloor*tfloor/(w->KELVIN*c->mu);
  int changed = 0;
  if (v[RHO] < dfloor) {
    double cs, dfact, temp;
    if (v[RHO] < 0.0) v[RHO] = dfloor;
    cs = sqrt(c->gamma*v[PRS]/v[RHO]);
    dfact = v[RHO]/dfloor;
    v[RHO] = dfloor;
    v[VX1] = dfact*v[VX1];
    v[VX2] = dfact*v[VX2];
    v[VX3] = dfact*v[VX3];
    v[PRS] = pow(cs, 2)*v[RHO]/c->gamma;
    temp = v[PRS]/v[RHO]*w->KELVIN*c->mu;
    if (temp < tfloor) {
      temp = tfloor;
      v[PRS] = v[RHO]*temp/(w->KELVIN*c->mu);
    }
    v[TRC] = 0.0;
    changed = 1;
  }
  if (v[PRS] < pfloor) {
    v[PRS] = pfloor;
    changed = 1;
  }
  return changed;
}

/* The midplane row of the same pass (init.c:276-311).  The pass runs
 * over every zone (TOT_LOOP) and resets those of row j == np_int + 2 when
 * that row's centre x_glob and its volume centroid xgc agree to 1e-6
 * (:280-281), which they do from about 48 polar zones on.  With 2 ghost
 * zones that row is the first ghost row beyond theta_max, and the X2-end
 * fill that follows overwrites it; with SHOCK_FLATTENING MULTID the
 * reference runs with 3 (Src/get_nghost.c:48-57) and it is the LAST
 * INTERIOR row: its density is set to the disk's rho_0 (r / r_in)^-alpha
 * with the polar momentum kept, v_r to 0, v_phi to the Keplerian speed,
 * the pressure to the Shakura-Sunyaev disk temperature's, the tracer to 1,
 * after the floors and without a new Uc (PrimToCons ran before it).  r and
 * theta are the zone's volume centroids (xgc). */
static int midplane_row(const W *w)
{
  const orc_wind_cfg *c = w->c;
  int j = NG + w->n2 - 1;
  if (c->dfloor <= 0.0 || !c->shock_flattening) return -1;
  return (fabs(w->x2[j] - w->xgc2[j])/w->x2[j]) < 1e-6 ? j : -1;
}

static void midplane_cell(const W *w, int i, double *v)
{
  const orc_wind_cfg *c = w->c;
  const double rho_0 = c->rho_0/w->UD, rho_alpha = c->rho_alpha, disk_mdot = c->disk_mdot;
  const double r_WD = w->x1l[NG];                  /* g_domBeg[IDIR] */
  const double gm_cgs = C_G*c->cent_mass, gm_code = gm_cgs/(w->UL*w->UV*w->UV);
  double r = w->xgc1[i], theta = w->xgc2[w->mid_row], rcyl = r*sin(theta);
  double rho_mid = rho_0*pow((r/r_WD), -1.0*rho_alpha), teff, temp;
  v[VX2] = (v[RHO]*v[VX2])/rho_mid;
  v[RHO] = rho_mid;
  v[VX1] = 0.0;
  v[VX3] = sqrt(gm_code/r)*sin(theta);
  teff = pow(3.0*gm_cgs*disk_mdot/(8.0*C_PI*C_sigma), 0.25);
  teff *= pow(r_WD*w->UL, -0.75);
  temp = teff*pow(r_WD/rcyl, 0.75)*pow(1.0 - sqrt(r_WD/rcyl), 0.25);
  v[PRS] = rho_mid*temp/(w->KELVIN*c->mu);
  v[TRC] = 1.0;
}

static void internal_floors(W *w)
{
  int i, j, nv;
  if (w->c->dfloor <= 0.0) return;
  for (j = 0; j < w->t2; j++)
    for (i = 0; i < w->t1; i++) {
      double v[NVW];
      for (nv = 0; nv < NVW; nv++) v[nv] = w->V[VI(w, nv, j, i)];
      if (floor_cell(w, v)) {
        for (nv = 0; nv < NVW; nv++) w->V[VI(w, nv, j, i)] = v[nv];
        p2c(w->c->gamma, v, w->U + NVW*CI(w, j, i));
        /* PrimToCons3D of the changed zone also forms Uc[ENTR] = rho
         * Vc[ENTR], from the entropy the last cons->prim left */
        if (w->c->entropy_switch) w->Ue[CI(w, j, i)] = v[RHO]*w->S[CI(w, j, i)];
        cnt[CNT_FLOOR_D]++;
      }
      if (j == w->mid_row) {
        midplane_cell(w, i, v);
        for (nv = 0; nv < NVW; nv++) w->V[VI(w, nv, j, i)] = v[nv];
        cnt[CNT_MIDPLANE]++;
      }
    }
}

/* ------------------------------------------------------------------ */
/* FlagShock with SHOCK_FLATTENING MULTID (Src/flag_shock.c:81-253),    */
/* 2-D spherical: a zone with div v < 0 and a total pressure jump       */
/* |dp_1| + |dp_2| > 5 min(p of the zone and its 4 neighbours) gets     */
/* FLAG_HLL and FLAG_MINMOD, its 4 neighbours FLAG_MINMOD.  MULTID      */
/* raises nghost to 3 (Src/get_nghost.c:48-57) and the test runs from   */
/* 1 to NX_TOT - 2, i.e. over both of the ghost layers this grid        */
/* stores; their outer neighbours (the third ghost layer) come from the */
/* same boundary rules, through ghost3().                               */
/* ------------------------------------------------------------------ */

/* Variable nv of zone (j, i) for j in [-1, t2], i in [-1, t1] (at most one
 * index outside the stored range): the boundary rule of that side. */
static double ghost3(const W *w, int nv, int j, int i)
{
  int d = (i < 0 || i >= w->t1) ? 0 : ((j < 0 || j >= w->t2) ? 1 : -1);
  if (d < 0) return w->V[VI(w, nv, j, i)];
  {
    int g = d ? j : i, n = d ? w->n2 : w->n1, b = NG, e = NG + n - 1, left = g < 0;
    int type = w->c->bc[2*d + (left ? 0 : 1)], src;
    double sgn = 1.0, val;
    if (type == ORC_BC_PERIODIC) src = left ? g + n : g - n;
    else if (type == ORC_BC_OUTFLOW || type == ORC_BC_WIND_INNER) src = left ? b : e;
    else if (type == ORC_BC_WIND_POLE && (nv == RHO || nv == PRS)) src = left ? b : e;
    else {
      src = left ? 2*b - 1 - g : 2*e + 1 - g;
      if (nv == VX1 + d) sgn = -1.0;
    }
    val = sgn*(d ? w->V[VI(w, nv, src, i)] : w->V[VI(w, nv, j, src)]);
    if (type == ORC_BC_WIND_INNER && nv == VX1) val = MIND(val, 0.0);
    return val;
  }
}

static void flag_shock(W *w)
{
  int i, j;
  const double eps = 5.0;                  /* EPS_PSHOCK_FLATTEN */
  for (j = 0; j < w->t2; j++)
    for (i = 0; i < w->t1; i++) {
      size_t c = CI(w, j, i);
      /* face areas left / below the zone: set_geometry.c:141-186 (i = -1,
       * j = -1 use the left edge of the first stored zone) */
      double dmu = fabs(cos(w->x2l[j]) - cos(w->x2r[j]));
      double A1m = i > 0 ? w->A1[c - 1] : w->x1l[0]*w->x1l[0]*dmu;
      double A2m = j > 0 ? w->A2[c - w->t1]
                         : fabs(w->x1[i])*w->dx1[i]*fabs(sin(w->x2l[0]));
      double dvx1, dvx2, divv;
      dvx1 = w->A1[c]*(ghost3(w, VX1, j, i + 1) + ghost3(w, VX1, j, i))
           - A1m*(ghost3(w, VX1, j, i - 1) + ghost3(w, VX1, j, i));
      dvx2 = w->A2[c]*(ghost3(w, VX2, j + 1, i) + ghost3(w, VX2, j, i))
           - A2m*(ghost3(w, VX2, j - 1, i) + ghost3(w, VX2, j, i));
      divv = (dvx1 + dvx2)/w->dV[c];
      if (divv < 0.0) {
        double pc = ghost3(w, PRS, j, i);
        double pe = ghost3(w, PRS, j, i + 1), pw = ghost3(w, PRS, j, i - 1);
        double pn = ghost3(w, PRS, j + 1, i), ps = ghost3(w, PRS, j - 1, i);
        double pt_min = pc, pt_min1, pt_min2, dpx1, dpx2, gradp;
        pt_min1 = MIND(pe, pw);
        pt_min2 = MIND(pn, ps);
        pt_min = MIND(pt_min, pt_min1);
        pt_min = MIND(pt_min, pt_min2);
        dpx1 = fabs(pe - pw);
        dpx2 = fabs(pn - ps);
        gradp = dpx1 + dpx2;
        if (gradp > eps*pt_min) {
          w->flag[c] |= FLAG_HLL | FLAG_MINMOD;
          if (i + 1 < w->t1) w->flag[c + 1] |= FLAG_MINMOD;
          if (i > 0) w->flag[c - 1] |= FLAG_MINMOD;
          if (j > 0) w->flag[c - w->t1] |= FLAG_MINMOD;
          if (j + 1 < w->t2) w->flag[c + w->t1] |= FLAG_MINMOD;
          cnt[CNT_SHOCK]++;
        }
      }
    }
}

/* PLM on characteristic variables, CHAR_LIMITING YES (Src/States/
 * plm_states.c:479-676) with the HD primitive eigenvectors (Src/HD/
 * eigenv.c:92-189, the adiabatic case) and PrimToChar (eigenv.c:575-605),
 * for one zone of a pencil in direction d: v the zone, dvp / dvm the
 * weighted forward / backward differences, flag its zone flags.  The
 * eigenvector arrays are cleared per sweep (ResetState), so only the
 * entries below are non-zero.  The tracer is limited in primitive form. */
static void plm_char(double gamma, int d, const double *v, const double *dvp,
                     const double *dvm, double cp, double cm, double dp, double dm, int mm,
                     double *vp, double *vm)
{
  const int VXn = VX1 + d, VXt = d ? VX3 : VX2, VXb = d ? VX1 : VX3;   /* set_indexes.c */
  double R[5][5] = {{0}}, L[5][5] = {{0}};
  double dwp[5], dwm[5], dw_lim[5], dv_lim[NVW];
  double cs2 = gamma*v[PRS]/v[RHO];                      /* SoundSpeed2, eos.c:33 */
  double cs = sqrt(cs2), rhocs = v[RHO]*cs, rho_cs = v[RHO]/cs;
  int k, nv;
  R[RHO][0] = 0.5*rho_cs; R[VXn][0] = -0.5; R[PRS][0] = 0.5*rhocs;
  R[RHO][1] = 0.5*rho_cs; R[VXn][1] = 0.5;  R[PRS][1] = 0.5*rhocs;
  R[RHO][2] = 1.0; R[VXt][3] = 1.0; R[VXb][4] = 1.0;
  L[0][VXn] = -1.0; L[0][PRS] = 1.0/rhocs;
  L[1][VXn] = 1.0;  L[1][PRS] = 1.0/rhocs;
  L[2][RHO] = 1.0; L[3][VXt] = 1.0; L[4][VXb] = 1.0;
  L[2][PRS] = -1.0/cs2;
#define PRIM_TO_CHAR(vv, w) do { \
    w[0] = L[0][VXn]*vv[VXn] + L[0][PRS]*vv[PRS]; \
    w[1] = L[1][VXn]*vv[VXn] + L[1][PRS]*vv[PRS]; \
    w[2] = vv[RHO] + L[2][PRS]*vv[PRS]; \
    w[3] = vv[VXt]; \
    w[4] = vv[VXb]; } while (0)
  PRIM_TO_CHAR(dvm, dwm);
  PRIM_TO_CHAR(dvp, dwp);
#undef PRIM_TO_CHAR
  for (k = 5; k--; ) {
    if (mm) dw_lim[k] = (dwp[k]*dwm[k] > 0.0 ? (fabs(dwp[k]) < fabs(dwm[k]) ? dwp[k] : dwm[k])
                                              : 0.0);
    else dw_lim[k] = (dwp[k]*dwm[k] > 0.0 ? (dwp[k])*(dwm[k])*(cp*(dwm[k]) + cm*(dwp[k]))
                      /((dwp[k])*(dwp[k]) + (dwm[k])*(dwm[k]) + (cp + cm - 2.0)*(dwp[k])*(dwm[k]))
                      : 0.0);
  }
  for (nv = 5; nv--; ) {
    double dc = 0.0, d2v;
    for (k = 0; k < 5; k++) dc += dw_lim[k]*R[nv][k];
    if (dvp[nv]*dvm[nv] > 0.0) {
      d2v = (fabs(cp*dvp[nv]) < fabs(cm*dvm[nv]) ? cp*dvp[nv] : cm*dvm[nv]);  /* ABS_MIN */
      dv_lim[nv] = (d2v*dc > 0.0 ? (fabs(d2v) < fabs(dc) ? d2v : dc) : 0.0); /* MINMOD_LIMITER */
    } else dv_lim[nv] = 0.0;
  }
  /* the tracer: SET_LIMITER (van Leer, general grid) */
  nv = TRC;
  dv_lim[nv] = (dvp[nv]*dvm[nv] > 0.0 ? (dvp[nv])*(dvm[nv])*(cp*(dvm[nv]) + cm*(dvp[nv]))
                /((dvp[nv])*(dvp[nv]) + (dvm[nv])*(dvm[nv]) + (cp + cm - 2.0)*(dvp[nv])*(dvm[nv]))
                : 0.0);
  for (nv = NVW; nv--; ) {
    vp[nv] = v[nv] + dv_lim[nv]*dp;
    vm[nv] = v[nv] - dv_lim[nv]*dm;
  }
}

/* Roe's HLL switch next to a shocked zone (Src/HD/roe.c:105-131): Davis
 * speeds, bmin = min(0, SL), bmax = max(0, SR), the HLL formula for every
 * interface. */
static void roe_hll_face(double gamma, int n, const double *vL, const double *uL,
                         const double *vR, const double *uR, double *flux, double *press,
                         double *cmax)
{
  double fL[NVW], fR[NVW], aL, aR, bmin, bmax, scrh;
  int nv;
  orc_phys_flux(n, vL, uL, fL);
  orc_phys_flux(n, vR, uR, fR);
  aL = sqrt(gamma*vL[PRS]/vL[RHO]);
  aR = sqrt(gamma*vR[PRS]/vR[RHO]);
  bmin = MIND(vL[VX1 + n] - aL, vR[VX1 + n] - aR);         /* hll_speed.c:76-91 */
  bmax = MAXD(vL[VX1 + n] + aL, vR[VX1 + n] + aR);
  *cmax = MAXD(fabs(bmin), fabs(bmax));
  bmin = MIND(0.0, bmin);
  bmax = MAXD(0.0, bmax);
  scrh = 1.0/(bmax - bmin);
  for (nv = TRC; nv--; ) {
    flux[nv] = bmin*bmax*(uR[nv] - uL[nv]) + bmax*fL[nv] - bmin*fR[nv];
    flux[nv] *= scrh;
  }
  *press = (bmax*vL[PRS] - bmin*vR[PRS])*scrh;
}

/* The interface flux of the configured solver: HLLC with the FLAG_HLL
 * switch of SHOCK_FLATTENING MULTID (Src/HD/hllc.c:57-104: the same speeds
 * and supersonic cases, then the HLL flux for a subsonic interface; the
 * [Solver] hll is that branch everywhere), or Roe (Src/HD/roe.c:86-345,
 * with its own HLL branch at flagged interfaces). */
static void riemann_face(double gamma, int n, int solver, int flagged, const double *vL,
                         const double *uL, const double *vR, const double *uR, double *flux,
                         double *press, double *cmax)
{
  double mach = 0.0;
  const int hll = solver == 2 || flagged;
  if (solver == 1) {
    if (flagged) roe_hll_face(gamma, n, vL, uL, vR, uR, flux, press, cmax);
    else orc_roe(gamma, 2, n, vL, uL, vR, uR, flux, press, cmax, &mach);
    return;
  }
  if (!hll) {
    orc_hllc(gamma, n, vL, uL, vR, uR, flux, press, cmax, &mach);
    return;
  }
  {
    double a2L = gamma*vL[PRS]/vL[RHO], a2R = gamma*vR[PRS]/vR[RHO];
    double fL[NVW], fR[NVW], aL, aR, SL, SR, scrh;   /* the NFLX = 5 HD fluxes */
    int nv;
    orc_phys_flux(n, vL, uL, fL);
    orc_phys_flux(n, vR, uR, fR);
    aL = sqrt(a2L);
    aR = sqrt(a2R);
    SL = MIND(vL[VX1 + n] - aL, vR[VX1 + n] - aR);        /* hll_speed.c:76-91 */
    SR = MAXD(vL[VX1 + n] + aL, vR[VX1 + n] + aR);
    *cmax = MAXD(fabs(SL), fabs(SR));
    if (SL > 0.0) {
      for (nv = 0; nv < TRC; nv++) flux[nv] = fL[nv];
      *press = vL[PRS];
      return;
    }
    if (SR < 0.0) {
      for (nv = 0; nv < TRC; nv++) flux[nv] = fR[nv];
      *press = vR[PRS];
      return;
    }
    scrh = 1.0/(SR - SL);
    for (nv = 0; nv < TRC; nv++) {
      flux[nv] = SL*SR*(uR[nv] - uL[nv]) + SR*fL[nv] - SL*fR[nv];
      flux[nv] *= scrh;
    }
    *press = (SR*vL[PRS] - SL*vR[PRS])*scrh;
  }
}

/* ------------------------------------------------------------------ */
/* boundaries (Src/boundary.c), sides in order x1beg, x1end, x2beg, x2end */
/* ------------------------------------------------------------------ */
static void compute_entropy(W *w);

static void boundary(W *w)
{
  int side, nv;
  internal_floors(w);                 /* boundary.c:126-128, INTERNAL_BOUNDARY */
  for (side = 0; side < 4; side++) {
    int d = side/2, left = (side%2 == 0), type = w->c->bc[side];
    int n = d ? w->n2 : w->n1, b = NG, e = NG + n - 1;
    int lo = left ? 0 : e + 1, hi = left ? b - 1 : (d ? w->t2 : w->t1) - 1;
    for (nv = 0; nv < NVW; nv++) {
      double sgn = (type == ORC_BC_REFLECTIVE && nv == VX1 + d) ? -1.0 : 1.0;
      int a0 = 0, a1 = d ? w->t1 - 1 : w->t2 - 1, a, g;
      if (type == ORC_BC_WIND_POLE && nv == VX2) sgn = -1.0;
      for (a = a0; a <= a1; a++)
      for (g = lo; g <= hi; g++) {
        int src;
        if (type == ORC_BC_PERIODIC) src = left ? g + n : g - n;
        else if (type == ORC_BC_OUTFLOW || type == ORC_BC_WIND_INNER) src = left ? b : e;
        else if (type == ORC_BC_WIND_POLE && (nv == RHO || nv == PRS)) src = left ? b : e;
        else src = left ? 2*b - 1 - g : 2*e + 1 - g;
        if (d == 0) w->V[VI(w, nv, a, g)] = sgn*w->V[VI(w, nv, a, src)];
        else w->V[VI(w, nv, g, a)] = sgn*w->V[VI(w, nv, src, a)];
        /* cv_idl X1-beg: inflow only (v_r <= 0) */
        if (type == ORC_BC_WIND_INNER && nv == VX1) {
          size_t q = d == 0 ? VI(w, nv, a, g) : VI(w, nv, g, a);
          w->V[q] = MIND(w->V[q], 0.0);
        }
      }
    }
  }
  if (w->c->entropy_switch) compute_entropy(w);   /* boundary.c:491-493 */
}

/* ComputeEntropy (Src/entropy_switch.c:14-37, Entropy() of
 * Src/EOS/Ideal/eos.c:77-96): s = p / rho^gamma on every zone. */
static void compute_entropy(W *w)
{
  int i, j;
  for (j = 0; j < w->t2; j++)
    for (i = 0; i < w->t1; i++)
      w->S[CI(w, j, i)] = w->V[VI(w, PRS, j, i)]/pow(w->V[VI(w, RHO, j, i)], w->c->gamma);
}

/* ------------------------------------------------------------------ */
/* line force                                                          */
/* ------------------------------------------------------------------ */
static void vgrad(W *w)
{
  const double UL = w->UL, UV = w->UV;
  int a, i, j;
  double *V1 = w->V + VI(w, VX1, 0, 0), *V2 = w->V + VI(w, VX2, 0, 0);
#define P1(jj, ii) V1[CI(w, jj, ii)]
#define P2(jj, ii) V2[CI(w, jj, ii)]
  if (!w->vgrad_setup) {                        /* line_connect.c:534-625 */
    for (a = 0; a < w->na; a++)
    for (j = NG; j < NG + w->n2; j++)
    for (i = NG; i < NG + w->n1; i++) {
      size_t c = CI(w, j, i), q = AI(w, a, j, i);
      double maxds, mod_flux;
      w->x11_0[c] = (w->x1[i - 1] + w->x1[i])/2.0*UL;
      w->x11_1[c] = (w->x2[j - 1] + w->x2[j])/2.0;
      w->x22_0[c] = (w->x1[i + 1] + w->x1[i])/2.0*UL;
      w->x22_1[c] = (w->x2[j + 1] + w->x2[j])/2.0;
      maxds = fabs(w->x22_0[c] - w->x11_0[c]);
      if (maxds > (w->x1[i]*UL*fabs(w->x22_1[c] - w->x11_1[c])))
        maxds = w->x1[i]*UL*fabs(w->x22_1[c] - w->x11_1[c]);
      maxds /= 2.0;
      mod_flux = sqrt(pow(w->fr[q], 2) + pow(w->ft[q], 2) + pow(w->fp[q], 2));
      if (mod_flux == 0.0) {
        w->moff[q] = -999;
      } else {
        double x = w->x1[i]*sin(w->x2[j])*UL, z = w->x1[i]*cos(w->x2[j])*UL;
        double ta = (a + 0.5)*(2.0*C_PI)/36.0;
        double dx1 = maxds*sin(ta), dx2 = maxds*cos(ta);
        double ds = sqrt(dx1*dx1 + dx2*dx2);
        w->roff[q] = sqrt((x + dx1)*(x + dx1) + (z + dx2)*(z + dx2));
        w->toff[q] = atan((x + dx1)/(z + dx2));
        w->moff[q] = ds;
      }
    }
    w->vgrad_setup = 1;
  }
  for (a = 0; a < w->na; a++)                  /* line_connect.c:628-738 */
  for (j = NG; j < NG + w->n2; j++)
  for (i = NG; i < NG + w->n1; i++) {
    size_t c = CI(w, j, i), q = AI(w, a, j, i);
    double v11[2], v12[2], v22[2], v21[2], x11[2], x22[2], loc[2], ans1[2], ans2[2];
    double vx1, vz1, ds, f1, f2, t1, t2;
    v11[0] = (P1(j - 1, i - 1) + P1(j - 1, i) + P1(j, i - 1) + P1(j, i))/4.0;
    v11[1] = (P2(j - 1, i - 1) + P2(j - 1, i) + P2(j, i - 1) + P2(j, i))/4.0;
    v12[0] = (P1(j + 1, i - 1) + P1(j, i - 1) + P1(j + 1, i) + P1(j, i))/4.0;
    v12[1] = (P2(j + 1, i - 1) + P2(j, i - 1) + P2(j + 1, i) + P2(j, i))/4.0;
    v22[0] = (P1(j + 1, i) + P1(j + 1, i + 1) + P1(j, i + 1) + P1(j, i))/4.0;
    v22[1] = (P2(j + 1, i) + P2(j + 1, i + 1) + P2(j, i + 1) + P2(j, i))/4.0;
    v21[0] = (P1(j, i + 1) + P1(j - 1, i + 1) + P1(j - 1, i) + P1(j, i))/4.0;
    v21[1] = (P2(j, i + 1) + P2(j - 1, i + 1) + P2(j - 1, i) + P2(j, i))/4.0;
    x11[0] = w->x11_0[c]; x11[1] = w->x11_1[c];
    x22[0] = w->x22_0[c]; x22[1] = w->x22_1[c];
    loc[0] = w->x1[i]*UL;
    loc[1] = w->x2[j];
    /* bilinear(), line_connect.c:746-777 */
    f1 = (loc[0] - x11[0])/(x22[0] - x11[0]);
    f2 = (loc[1] - x11[1])/(x22[1] - x11[1]);
    t1 = (1.0 - f1)*v11[0] + f1*v21[0];
    t2 = (1.0 - f1)*v12[0] + f1*v22[0];
    ans1[0] = (1.0 - f2)*t1 + f2*t2;
    t1 = (1.0 - f1)*v11[1] + f1*v21[1];
    t2 = (1.0 - f1)*v12[1] + f1*v22[1];
    ans1[1] = (1.0 - f2)*t1 + f2*t2;
    vx1 = (ans1[0]*UV*sin(w->x2[j]) + ans1[1]*UV*cos(w->x2[j]));
    vz1 = (ans1[0]*UV*cos(w->x2[j]) - ans1[1]*UV*sin(w->x2[j]));
    loc[0] = w->roff[q];
    loc[1] = w->toff[q];
    ds = w->moff[q];
    if (ds == -999) {
      w->dvds[q] = -999;
    } else {
      double vx2, vz2, ta, v1, v2;
      f1 = (loc[0] - x11[0])/(x22[0] - x11[0]);
      f2 = (loc[1] - x11[1])/(x22[1] - x11[1]);
      t1 = (1.0 - f1)*v11[0] + f1*v21[0];
      t2 = (1.0 - f1)*v12[0] + f1*v22[0];
      ans2[0] = (1.0 - f2)*t1 + f2*t2;
      t1 = (1.0 - f1)*v11[1] + f1*v21[1];
      t2 = (1.0 - f1)*v12[1] + f1*v22[1];
      ans2[1] = (1.0 - f2)*t1 + f2*t2;
      vx2 = (ans2[0]*UV*sin(loc[1]) + ans2[1]*UV*cos(loc[1]));
      vz2 = (ans2[0]*UV*cos(loc[1]) - ans2[1]*UV*sin(loc[1]));
      ta = (a + 0.5)*(2.0*C_PI)/36.0;
      v1 = sin(ta)*vx1 + cos(ta)*vz1;
      v2 = sin(ta)*vx2 + cos(ta)*vz2;
      w->dvds[q] = fabs((v2 - v1)/ds);
    }
  }
#undef P1
#undef P2
}

/* linterp(), line_connect.c:781-811: the first fit point not below x by
   a linear scan, linear interpolation in log10 M, 10^ of the result; the
   end values outside the fit range; 0 for a NaN result. */
static double linterp(double x, const double *xa, const double *ya, size_t ystride, int n)
{
  int idx = 0;
  double result;
  while (idx < n && xa[idx] < x) idx++;
  if (idx == 0) {
    result = pow(10.0, ya[0]);
  } else if (idx >= n) {
    result = pow(10.0, ya[(size_t)(n - 1)*ystride]);
  } else {
    double x_low = xa[idx - 1], x_high = xa[idx];
    double y_low = ya[(size_t)(idx - 1)*ystride], y_high = ya[(size_t)idx*ystride];
    double slope = (y_high - y_low)/(x_high - x_low);
    result = pow(10.0, y_low + slope*(x - x_low));
  }
  if (isnan(result)) result = 0.0;
  return result;
}

/* LineForce(v, grad, ...) for cell (j, i), line_connect.c:815-903: the
   k, alpha mode, or the tabulated M(t) of the cell for krad = alpharad =
   999 */
static void line_force(W *w, const double *v, double *grad, int j, int i)
{
  const orc_wind_cfg *c = w->c;
  double sigma_e = C_sigmaT/C_amu/1.18;
  double rho = v[RHO]*w->UD;
  double krad = c->krad, alpharad = c->alpharad, M_max = 4400.;
  double T, v_th;
  int a, tab = (krad == 999 && alpharad == 999);
  const double *mcol = tab ? w->fm_lm + (size_t)(j - NG)*w->n1 + (i - NG) : NULL;
  T = v[PRS]/v[RHO]*w->KELVIN*c->mu;
  v_th = sqrt((2.0*C_kB*T)/C_mp);
  grad[0] = grad[1] = grad[2] = 0.0;
  for (a = 0; a < w->na; a++) {
    size_t q = AI(w, a, j, i);
    double flux_r = w->fr[q], flux_t = w->ft[q], M;
    if (w->dvds[q] > 0.0) {
      double t_UV = sigma_e*rho*v_th/w->dvds[q];
      if (tab) {
        double log10_t_UV = log10(t_UV);
        M = linterp(log10_t_UV, w->fm_lt, mcol, (size_t)w->n1*w->n2, c->fm_points);
      } else {
        M = krad*pow(t_UV, alpharad);
      }
      if (M > M_max) M = M_max;
    } else {
      M = 0.0;
    }
    grad[0] += ((1.0 + M)*sigma_e*flux_r/C_c)/w->UA;
    grad[1] += ((1.0 + M)*sigma_e*flux_t/C_c)/w->UA;
  }
}

/* ------------------------------------------------------------------ */
/* one stage (UpdateStage) on the 2-D spherical grid                   */
/* ------------------------------------------------------------------ */
static void update_stage(W *w, double dt)
{
  const orc_wind_cfg *c = w->c;
  double gamma = c->gamma;
  double gm_code = C_G*c->cent_mass/(w->UL*w->UV*w->UV);
  int d, ip, nv, maxn = MAXD(w->t1, w->t2);
  double (*v)[NVW] = malloc(sizeof(double)*NVW*maxn);
  double (*vp)[NVW] = malloc(sizeof(double)*NVW*maxn);
  double (*vm)[NVW] = malloc(sizeof(double)*NVW*maxn);
  double (*flux)[NVW] = malloc(sizeof(double)*NVW*maxn);
  double (*fA)[NVW] = malloc(sizeof(double)*NVW*maxn);
  double (*rhs)[NVW] = malloc(sizeof(double)*NVW*maxn);
  double *press = malloc(sizeof(double)*maxn), *cmax = malloc(sizeof(double)*maxn);
  double *dv = malloc(sizeof(double)*NVW*maxn);
  /* the entropy scalar: zone values, face states, flux, F.A, rhs */
  double *se = malloc(sizeof(double)*maxn), *sep = malloc(sizeof(double)*maxn);
  double *sem = malloc(sizeof(double)*maxn), *fe = malloc(sizeof(double)*maxn);
  double *fAe = malloc(sizeof(double)*maxn), *rhse = malloc(sizeof(double)*maxn);
  const int ent = w->c->entropy_switch;

  if (w->stage == 1) memset(w->Cdt, 0, sizeof(double)*w->t1*w->t2);
  vgrad(w);

  for (d = 0; d < 2; d++) {
    int ntot = d ? w->t2 : w->t1, nbeg = NG, nend = NG + (d ? w->n2 : w->n1) - 1;
    int other, ob = NG, oe = NG + (d ? w->n1 : w->n2) - 1;
    double *cp = w->pc[d][0], *cm = w->pc[d][1], *wp = w->pc[d][2], *wm = w->pc[d][3];
    double *dp = w->pc[d][4], *dm = w->pc[d][5];
    for (other = ob; other <= oe; other++) {
      int i = d ? other : 0, j = d ? 0 : other;
      for (ip = 0; ip < ntot; ip++)
        for (nv = 0; nv < NVW; nv++)
          v[ip][nv] = d ? w->V[VI(w, nv, ip, i)] : w->V[VI(w, nv, j, ip)];
      if (ent)
        for (ip = 0; ip < ntot; ip++) se[ip] = d ? w->S[CI(w, ip, i)] : w->S[CI(w, j, ip)];

      /* States, plm_states.c:141-258 (general grid, van Leer) */
      for (ip = nbeg - 2; ip <= nend + 1; ip++)
        for (nv = 0; nv < NVW; nv++) dv[ip*NVW + nv] = v[ip + 1][nv] - v[ip][nv];
      for (ip = nbeg - 1; ip <= nend + 1; ip++) {
        /* plm_states.c:177-190: a FLAG_MINMOD zone takes the minmod slope */
        int mm = (d ? w->flag[CI(w, ip, i)] : w->flag[CI(w, j, ip)]) & FLAG_MINMOD;
        if (w->c->char_limiting) {
          double dvp[NVW], dvm[NVW];
          for (nv = 0; nv < NVW; nv++) {
            dvp[nv] = dv[ip*NVW + nv]*wp[ip];
            dvm[nv] = dv[(ip - 1)*NVW + nv]*wm[ip];
          }
          plm_char(gamma, d, v[ip], dvp, dvm, cp[ip], cm[ip], dp[ip], dm[ip], mm, vp[ip], vm[ip]);
          continue;
        }
        for (nv = 0; nv < NVW; nv++) {
          double dvp = dv[ip*NVW + nv]*wp[ip], dvm = dv[(ip - 1)*NVW + nv]*wm[ip];
          double dl;
          if (mm) dl = (dvp*dvm > 0.0 ? (fabs(dvp) < fabs(dvm) ? dvp : dvm) : 0.0);
          else dl = (dvp*dvm > 0.0 ? (dvp)*(dvm)*(cp[ip]*(dvm) + cm[ip]*(dvp))
                     /((dvp)*(dvp) + (dvm)*(dvm) + (cp[ip] + cm[ip] - 2.0)*(dvp)*(dvm)) : 0.0);
          vp[ip][nv] = v[ip][nv] + dl*dp[ip];
          vm[ip][nv] = v[ip][nv] - dl*dm[ip];
        }
      }
      /* the entropy scalar: limited in primitive form like the tracer
       * (van Leer; minmod in FLAG_MINMOD zones unless CHAR_LIMITING,
       * whose scalars always take the van Leer limiter) */
      if (ent)
        for (ip = nbeg - 1; ip <= nend + 1; ip++) {
          int mm = !w->c->char_limiting &&
                   ((d ? w->flag[CI(w, ip, i)] : w->flag[CI(w, j, ip)]) & FLAG_MINMOD);
          double dvp = (se[ip + 1] - se[ip])*wp[ip], dvm = (se[ip] - se[ip - 1])*wm[ip], dl;
          if (mm) dl = (dvp*dvm > 0.0 ? (fabs(dvp) < fabs(dvm) ? dvp : dvm) : 0.0);
          else dl = (dvp*dvm > 0.0 ? (dvp)*(dvm)*(cp[ip]*(dvm) + cm[ip]*(dvp))
                     /((dvp)*(dvp) + (dvm)*(dvm) + (cp[ip] + cm[ip] - 2.0)*(dvp)*(dvm)) : 0.0);
          sep[ip] = se[ip] + dl*dp[ip];
          sem[ip] = se[ip] - dl*dm[ip];
        }
      /* Riemann + tracer flux (adv_flux.c) */
      for (ip = nbeg - 1; ip <= nend; ip++) {
        double uL[NVW], uR[NVW];
        int f0 = d ? w->flag[CI(w, ip, i)] : w->flag[CI(w, j, ip)];
        int f1 = d ? w->flag[CI(w, ip + 1, i)] : w->flag[CI(w, j, ip + 1)];
        p2c(gamma, vp[ip], uL);
        p2c(gamma, vm[ip + 1], uR);
        riemann_face(gamma, d, w->c->solver, ((f0 | f1) & FLAG_HLL) != 0, vp[ip], uL,
                     vm[ip + 1], uR, flux[ip], &press[ip], &cmax[ip]);
        flux[ip][TRC] = flux[ip][RHO]*(flux[ip][RHO] > 0.0 ? vp[ip][TRC] : vm[ip + 1][TRC]);
        if (ent) {                                /* adv_flux.c:131-134 */
          if (flux[ip][RHO] >= 0.0) fe[ip] = sep[ip]*flux[ip][RHO];
          else fe[ip] = sem[ip + 1]*flux[ip][RHO];
        }
      }
      /* TotalFlux (rhs.c:466-560) */
      for (ip = nbeg - 1; ip <= nend; ip++) {
        double A = d ? w->A2[CI(w, ip, i)] : w->A1[CI(w, j, ip)];
        for (nv = 0; nv < NVW; nv++) fA[ip][nv] = flux[ip][nv]*A;
        if (ent) fAe[ip] = fe[ip]*A;
        if (d == 0) fA[ip][VX3] *= fabs(w->x1r[ip]);
        else fA[ip][VX3] *= fabs(w->sp[ip]);
      }
      /* RightHandSide (rhs.c:234-351) and RightHandSideSource */
      for (ip = nbeg; ip <= nend; ip++) {
        int ci = d ? i : ip, cj = d ? ip : j;
        double dtdV = dt/w->dV[CI(w, cj, ci)];
        double dtdl = d ? dt/w->dx2[ip]*(1.0/w->rt[i]) : dt/w->dx1[ip];
        double vc[NVW], g[3], grad[3], Sm;
        for (nv = 0; nv < NVW; nv++) rhs[ip][nv] = -dtdV*(fA[ip][nv] - fA[ip - 1][nv]);
        if (ent) rhse[ip] = -dtdV*(fAe[ip] - fAe[ip - 1]);
        rhs[ip][VX1 + d] -= dtdl*(press[ip] - press[ip - 1]);
        if (d == 0) rhs[ip][VX3] /= fabs(w->x1[ip]);
        else rhs[ip][VX3] /= fabs(w->s[ip]);

        if (d == 0) {                           /* rhs_source.c:227-297 */
          double r_1 = 1.0/w->x1[ip], vphi;
          for (nv = 0; nv < NVW; nv++) vc[nv] = 0.5*(vp[ip][nv] + vm[ip][nv]);
          vphi = vc[VX3];
          Sm = vc[RHO]*(vc[VX2]*vc[VX2] + vphi*vphi);
          Sm += - 0.0 - 0.0;
          rhs[ip][VX1] += dt*Sm*r_1;
          g[0] = -1.0*gm_code/(w->x1[ip]*w->x1[ip]); g[1] = 0.0; g[2] = 0.0;
          rhs[ip][VX1] += dt*vc[RHO]*g[0];
          rhs[ip][ENG] += dt*0.5*(flux[ip][RHO] + flux[ip - 1][RHO])*g[0];
          line_force(w, vc, grad, cj, ci);
          rhs[ip][VX1] += dt*vc[RHO]*grad[0];
          rhs[ip][ENG] += dt*0.5*(flux[ip][RHO] + flux[ip - 1][RHO])*grad[0];
        } else {                                /* rhs_source.c:313-396 */
          double r_1 = 1.0/w->rt[i], ct = 1.0/tan(w->x2[ip]), vphi;
          for (nv = 0; nv < NVW; nv++) vc[nv] = v[ip][nv];
          vphi = vc[VX3];
          Sm = vc[RHO]*(- vc[VX2]*vc[VX1] + ct*vphi*vphi);
          Sm += 0.0 - ct*0.0;
          rhs[ip][VX2] += dt*Sm*r_1;
          g[0] = -1.0*gm_code/(w->x1[i]*w->x1[i]); g[1] = 0.0; g[2] = 0.0;
          rhs[ip][VX2] += dt*vc[RHO]*g[1];
          rhs[ip][ENG] += dt*0.5*(flux[ip][RHO] + flux[ip - 1][RHO])*g[1];
          rhs[ip][VX3] += dt*vc[RHO]*g[2];
          rhs[ip][ENG] += dt*vc[RHO]*vc[VX3]*g[2];
          line_force(w, vc, grad, cj, ci);
          rhs[ip][VX2] += dt*vc[RHO]*grad[1];
          rhs[ip][ENG] += dt*0.5*(flux[ip][RHO] + flux[ip - 1][RHO])*grad[1];
          rhs[ip][VX3] += dt*vc[RHO]*grad[2];
          rhs[ip][ENG] += dt*vc[RHO]*vc[VX3]*grad[2];
        }
      }
      for (ip = nbeg; ip <= nend; ip++) {
        double *u = w->U + NVW*(d ? CI(w, ip, i) : CI(w, j, ip));
        for (nv = 0; nv < NVW; nv++) u[nv] += rhs[ip][nv];
        if (ent) w->Ue[d ? CI(w, ip, i) : CI(w, j, ip)] += rhse[ip];
      }
      if (w->stage == 1) {                      /* update_stage.c:303-316 */
        for (ip = nbeg; ip <= nend; ip++) {
          double inv_dl = d ? w->inv_dx2[ip]*(1.0/w->x1[i]) : w->inv_dx1[ip];
          w->Cdt[d ? CI(w, ip, i) : CI(w, j, ip)] += 0.5*(cmax[ip - 1] + cmax[ip])*inv_dl*1.0;
        }
      }
    }
  }
  if (w->stage == 1) {
    int i, j;
    for (j = NG; j < NG + w->n2; j++)
      for (i = NG; i < NG + w->n1; i++) w->invDt_hyp = MAXD(w->invDt_hyp, w->Cdt[CI(w, j, i)]);
    w->invDt_hyp /= (double)2;
  }
  free(v); free(vp); free(vm); free(flux); free(fA); free(rhs); free(press); free(cmax);
  free(dv); free(se); free(sep); free(sem); free(fe); free(fAe); free(rhse);
}

static void advance_step(W *w, double dt)
{
  int i, j, nv;
  size_t nc = (size_t)w->t1*w->t2;
  memset(w->flag, 0, nc);             /* main.c:258-261, once per step */
  w->stage = 1;
  boundary(w);
  if (w->c->shock_flattening) flag_shock(w);   /* rk_step.c:123-125 */
  for (j = NG; j < NG + w->n2; j++)
    for (i = NG; i < NG + w->n1; i++) {
      double v[NVW];
      for (nv = 0; nv < NVW; nv++) v[nv] = w->V[VI(w, nv, j, i)];
      p2c(w->c->gamma, v, w->U + NVW*CI(w, j, i));
      if (w->c->entropy_switch) w->Ue[CI(w, j, i)] = v[RHO]*w->S[CI(w, j, i)];
    }
  memcpy(w->U0, w->U, sizeof(double)*NVW*nc);
  if (w->c->entropy_switch) memcpy(w->U0e, w->Ue, sizeof(double)*nc);
  update_stage(w, dt);
  for (j = NG; j < NG + w->n2; j++)
    for (i = NG; i < NG + w->n1; i++) {
      double v[NVW];
      size_t c = CI(w, j, i);
      int f = w->c->entropy_switch ? c2p_entropy(w->c->gamma, w->U + NVW*c, &w->Ue[c], v, &w->S[c])
                                   : c2p(w->c->gamma, w->U + NVW*c, v);
      if (f) cnt[CNT_C2P_FAIL]++;
      for (nv = 0; nv < NVW; nv++) w->V[VI(w, nv, j, i)] = v[nv];
    }
  w->stage = 2;
  boundary(w);
  update_stage(w, dt);
  for (j = NG; j < NG + w->n2; j++)
    for (i = NG; i < NG + w->n1; i++) {
      size_t c = CI(w, j, i);
      double v[NVW], *u = w->U + NVW*c, *u0 = w->U0 + NVW*c;
      int f;
      for (nv = 0; nv < NVW; nv++) u[nv] = 0.5*u0[nv] + 0.5*u[nv];
      if (w->c->entropy_switch) {
        w->Ue[c] = 0.5*w->U0e[c] + 0.5*w->Ue[c];
        f = c2p_entropy(w->c->gamma, u, &w->Ue[c], v, &w->S[c]);
      } else {
        f = c2p(w->c->gamma, u, v);
      }
      if (f) cnt[CNT_C2P_FAIL]++;
      for (nv = 0; nv < NVW; nv++) w->V[VI(w, nv, j, i)] = v[nv];
    }
}

/* ------------------------------------------------------------------ */
/* BLONDIN cooling, Src/Cooling/BLONDIN/cooling.c                      */
/* ------------------------------------------------------------------ */
typedef struct {
  double xi, sqxi, sqsqxi, ne, nH, n, tx, E, hc_init, dt_share;
  double comp_c_pre, comp_h_pre, line_c_pre, brem_c_pre, xray_h_pre;
} Bl;

static double ne_rat(double T)
{
  double ne_ratio;
  if (T < 1.5e4) {
    ne_ratio = 1e-2 + pow(10, (-51.59417133 + 12.27740153*log10(T)));
  } else if (T >= 1.5e4 && T < 3.3e4) {
    ne_ratio = pow(10, (-3.80749689 + 0.86092628*log10(T)));
  } else {
    ne_ratio = 1.21;
  }
  return ne_ratio;
}

static double heatcool(Bl *b, double T)
{
  double lambda, sqT = sqrt(T), comp_heat, comp_cool, xray_heat, line_cool, brem_cool;
  b->ne = b->nH*ne_rat(T);
  comp_heat = b->comp_h_pre*(8.9e-36*b->xi*b->tx);
  comp_cool = b->comp_c_pre*(8.9e-36*b->xi*(4.0*T));
  xray_heat = b->xray_h_pre*(1.5e-21*(b->sqsqxi/sqT));
  line_cool = b->line_c_pre*((1e-16*exp(-1.3e5/T)/b->sqxi/T) +
              fmin(fmin(1e-24, 5e-27*sqT), 1.5e-17/T));
  brem_cool = b->brem_c_pre*(3.3e-27*sqT);
  lambda = b->nH*(b->ne*comp_heat + b->nH*xray_heat
                  - b->ne*comp_cool - b->ne*line_cool - b->ne*brem_cool);
  return lambda;
}

static double zfunc(Bl *b, double temp)
{
  return (temp*b->n*C_kB/(2.0/3.0)) - b->E - b->dt_share*(b->hc_init + heatcool(b, temp))/2.0;
}

static double bfunc(Bl *b, int which, double x)
{
  return which ? heatcool(b, x) : zfunc(b, x);
}

static double zbrent(Bl *bl, int which, double x1, double x2, double tol)
{
  const int ITMAX = 100;
  const double EPS = 3.0e-8;
  int iter;
  double a = x1, b = x2, c = x2, d = 0.0, e = 0.0;
  double fa = bfunc(bl, which, a), fb = bfunc(bl, which, b), fc = fb;
  double p, q, r, s, tol1, xm;
  if (fb*fa > 0.0) return b;
  for (iter = 1; iter <= ITMAX; iter++) {
    if (fb*fc > 0.0) { c = a; fc = fa; e = d = b - a; }
    if (fabs(fc) < fabs(fb)) { a = b; b = c; c = a; fa = fb; fb = fc; fc = fa; }
    tol1 = 2.0*EPS*fabs(b) + 0.5*tol;
    xm = 0.5*(c - b);
    if (fabs(xm) <= tol1 || fb == 0.0) return b;
    if (fabs(e) >= tol1 && fabs(fa) > fabs(fb)) {
      s = fb/fa;
      if (a == c) {
        p = 2.0*xm*s;
        q = 1.0 - s;
      } else {
        q = fa/fc;
        r = fb/fc;
        p = s*(2.0*xm*q*(q - r) - (b - a)*(r - 1.0));
        q = (q - 1.0)*(r - 1.0)*(s - 1.0);
      }
      if (p > 0.0) q = -q;
      p = fabs(p);
      {
        double min1 = 3.0*xm*q - fabs(tol1*q), min2 = fabs(e*q);
        if (2.0*p < (min1 < min2 ? min1 : min2)) { e = d; d = p/q; }
        else { d = xm; e = d; }
      }
    } else {
      d = xm;
      e = d;
    }
    a = b;
    fa = fb;
    if (fabs(d) > tol1) b += d;
    else b += (xm > 0.0 ? fabs(tol1) : -fabs(tol1));
    fb = bfunc(bl, which, b);
  }
  return b;
}

static void blondin(W *w, double dt)
{
  const orc_wind_cfg *c = w->c;
  double lx = c->l_star*c->f_x, mu = c->mu, minT = 1.e4;
  int i, j;
  Bl b;
  b.dt_share = dt*w->UT;
  for (j = NG; j < NG + w->n2; j++)
  for (i = NG; i < NG + w->n1; i++) {
    double r, rho, p, T, T_f, t_l, t_u, test, E_f;
    double pre[5];
    prefactors(w, j, i, pre);                   /* cooling.c:74-79 */
    b.comp_h_pre = pre[0]; b.comp_c_pre = pre[1]; b.xray_h_pre = pre[2];
    b.brem_c_pre = pre[3]; b.line_c_pre = pre[4];
    r = w->x1[i]*w->UL;
    rho = w->V[VI(w, RHO, j, i)]*w->UD;
    p = w->V[VI(w, PRS, j, i)];
    T = p/w->V[VI(w, RHO, j, i)]*w->KELVIN*mu;
    b.E = (p*w->UP)/(c->gamma - 1);
    b.nH = rho/(1.43*C_mp);
    if (w->t <= 3.0) {
      b.xi = lx/b.nH/r/r;
      b.tx = c->t_x;
    } else {
      b.xi = w->xi_s[CI(w, j, i)];
      b.tx = w->tr_s[CI(w, j, i)];
    }
    b.n = rho/(mu*C_mp);
    T = b.E*(2.0/3.0)/(b.n*C_kB);
    if (T < minT) { cnt[CNT_BELOW_MINT]++; continue; }
    cnt[CNT_BL_SOLVE]++;
    b.sqxi = sqrt(b.xi);
    b.sqsqxi = pow(b.xi, 0.25);
    b.hc_init = heatcool(&b, T);
    t_l = T*0.9;
    t_u = T*1.1;
    test = zfunc(&b, t_l)*zfunc(&b, t_u);
    while (test > 0 && test == test) {
      cnt[CNT_BL_WIDEN]++;
      t_l *= 0.9;
      t_u *= 1.1;
      test = zfunc(&b, t_l)*zfunc(&b, t_u);
    }
    if (test != test) {
      T_f = T;
    } else {
      double hc_final;
      T_f = zbrent(&b, 0, t_l, t_u, 1.0);
      hc_final = heatcool(&b, T_f);
      if (hc_final*b.hc_init < 0.0) cnt[CNT_BL_SECOND]++;
      if (hc_final*b.hc_init < 0.0) T_f = zbrent(&b, 1, fmin(T_f, T), fmax(T_f, T), 1.0);
    }
    if (T_f < minT) cnt[CNT_FLOOR_T]++;
    T_f = MAXD(T_f, minT);
    E_f = T_f/(2.0/3.0)*(b.n*C_kB);
    w->V[VI(w, PRS, j, i)] = E_f*(c->gamma - 1)/w->UP;
  }
}

/* ------------------------------------------------------------------ */
/* sub-cycled plain integrator: CoolingSource + Radiat + RKF12 / CK45   */
/* ------------------------------------------------------------------ */
typedef struct { W *w; double x1; double xi_s, tr_s; double pre[5]; } Rad;


static void radiat(const Rad *R, double *v, double *rhs)
{
  const W *w = R->w;
  const orc_wind_cfg *c = w->c;
  double T, prs, mu = c->mu, lx = c->l_star*c->f_x, r = R->x1*w->UL, xi, tx;
  double rho = v[RHO]*w->UD, nH = rho/(1.43*C_mp), ne;
  double E_cost = w->UL/w->UD/pow(w->UV, 3.0);
  double sqT, sqxi, sqsqxi, comp_heat, comp_cool, xray_heat, line_cool, brem_cool, lambda;
  const double minT = 1.e4;
  prs = v[RHOE]*(c->gamma - 1.0);
  if (prs < 0.0) { prs = 1.e-12; v[RHOE] = prs/(c->gamma - 1.0); }
  T = prs/v[RHO]*w->KELVIN*mu;
  if (T < minT) { rhs[RHOE] = 0.0; return; }
  if (w->t < 3.0) { xi = lx/nH/r/r; tx = c->t_x; }
  else { xi = R->xi_s; tx = R->tr_s; }
  sqT = sqrt(T);
  sqxi = sqrt(xi);
  sqsqxi = pow(xi, 0.25);
  ne = 1.21*nH;
  /* cv_idl/radiat.c:118-123 */
  comp_heat = R->pre[0]*(8.9e-36*xi*tx);
  comp_cool = R->pre[1]*(8.9e-36*xi*(4.0*T));
  xray_heat = R->pre[2]*(1.5e-21*(sqsqxi/sqT));
  line_cool = R->pre[4]*((1e-16*exp(-1.3e5/T)/sqxi/T) + fmin(fmin(1e-24, 5e-27*sqT), 1.5e-17/T));
  brem_cool = R->pre[3]*(3.3e-27*sqT);
  lambda = nH*(ne*comp_heat + nH*xray_heat - ne*comp_cool - ne*line_cool - ne*brem_cool);
  if (v[TRC] > 0.01) rhs[RHOE] = lambda*E_cost;
  else rhs[RHOE] = 0.0;
  rhs[RHOE] *= 1.0 - 1.0/cosh(pow(T/minT, 12));
}

static double rkf12(const Rad *R, double *v0, double *k1, double *v2nd, double dt)
{
  double v1[NVC], v1st, vscal, k2[NVC], err, scrh;
  int nv;
  for (nv = 0; nv < NVC; nv++) { v1[nv] = v0[nv]; k2[nv] = 0.0; }
  v1[RHOE] = v0[RHOE] + 0.5*dt*k1[RHOE];
  radiat(R, v1, k2);
  v2nd[RHOE] = v0[RHOE] + dt*k2[RHOE];
  v1st = v0[RHOE] + dt*k1[RHOE];
  vscal = fabs(v0[RHOE]) + dt*fabs(k1[RHOE]);
  err = 0.0;
  scrh = fabs(v2nd[RHOE] - v1st)/fabs(vscal);
  err = MAXD(err, scrh);
  return err;
}

static double ck45(const Rad *R, double *v0, double *k1, double *v5th, double dt, double tol)
{
  const double c1 = 37.0/378.0, c2 = 0.0, c3 = 250.0/621.0, c4 = 125.0/594.0, c5 = 0.0,
               c6 = 512.0/1771.0;
  const double cs1 = 2825.0/27648.0, cs2 = 0.0, cs3 = 18575.0/48384.0,
               cs4 = 13525.0/55296.0, cs5 = 277.0/14336.0, cs6 = 0.25;
  const double b21 = 0.2, b31 = 3.0/40.0, b32 = 9.0/40.0, b41 = 0.3, b42 = -0.9, b43 = 1.2;
  const double b51 = -11.0/54.0, b52 = 2.5, b53 = -70.0/27.0, b54 = 35.0/27.0;
  const double b61 = 1631.0/55296.0, b62 = 175.0/512.0, b63 = 575.0/13824.0,
               b64 = 44275.0/110592.0, b65 = 253.0/4096.0;
  double v1[NVC], v4th, k2[NVC], k3[NVC], k4[NVC], k5[NVC], k6[NVC];
  double v0_copy[NVC], k1_copy[NVC], t, tend, err, scrh, vscal;
  int nv, ksub = 0;
  for (nv = NVC; nv--;) {
    v1[nv] = v0[nv]; v0_copy[nv] = v0[nv]; k1_copy[nv] = k1[nv];
    k2[nv] = k3[nv] = k4[nv] = k5[nv] = k6[nv] = 0.0;
  }
  t = 0.0;
  tend = dt;
  for (;;) {
    v1[RHOE] = v0[RHOE] + dt*b21*k1[RHOE];
    radiat(R, v1, k2);
    v1[RHOE] = v0[RHOE] + dt*(b31*k1[RHOE] + b32*k2[RHOE]);
    radiat(R, v1, k3);
    v1[RHOE] = v0[RHOE] + dt*(b41*k1[RHOE] + b42*k2[RHOE] + b43*k3[RHOE]);
    radiat(R, v1, k4);
    v1[RHOE] = v0[RHOE] + dt*(b51*k1[RHOE] + b52*k2[RHOE] + b53*k3[RHOE] + b54*k4[RHOE]);
    radiat(R, v1, k5);
    v1[RHOE] = v0[RHOE] + dt*(b61*k1[RHOE] + b62*k2[RHOE] + b63*k3[RHOE] + b64*k4[RHOE]
                              + b65*k5[RHOE]);
    radiat(R, v1, k6);
    v5th[RHOE] = v0[RHOE] + dt*(c1*k1[RHOE] + c2*k2[RHOE] + c3*k3[RHOE] + c4*k4[RHOE]
                                + c5*k5[RHOE] + c6*k6[RHOE]);
    v4th = v0[RHOE] + dt*(cs1*k1[RHOE] + cs2*k2[RHOE] + cs3*k3[RHOE] + cs4*k4[RHOE]
                          + cs5*k5[RHOE] + cs6*k6[RHOE]);
    vscal = fabs(v0[RHOE]) + dt*fabs(k1[RHOE]);
    err = 0.0;
    scrh = fabs(v5th[RHOE] - v4th)/fabs(vscal);
    err = MAXD(err, scrh);
    err /= tol;
    if (err < 1.0) {
      double dt_grow;
      ksub++;
      cnt[CNT_CK45_SUB]++;
      err = MAXD(0.0, 1.e-18);
      t += dt;
      dt_grow = 0.9*dt*pow(err, -0.2);
      dt_grow = MIND(dt_grow, 5.0*dt);
      dt = dt_grow;
      if (fabs(t/tend - 1.0) < 1.e-9) break;
      if (dt > (tend - t)) dt = tend - t;
      v0[RHOE] = v5th[RHOE];
      radiat(R, v0, k1);
      if (ksub > 1000) {                  /* cooling_ode_solver.c:190-192: QUIT_PLUTO */
        orc_set_quit(ORC_QUIT_COOL_SUBSTEPS);
        break;
      }
    } else {
      double dt_shrink = 0.9*dt*pow(err, -0.25);
      cnt[CNT_CK45_REJECT]++;
      dt = MAXD(dt_shrink, 0.05*dt);
    }
  }
  for (nv = NVC; nv--;) { v0[nv] = v0_copy[nv]; k1[nv] = k1_copy[nv]; }
  return dt;
}

static void cooling_plain(W *w, double dt)
{
  const orc_wind_cfg *c = w->c;
  const double min_tol = 2.e-5, minT = 1.e4, maxCoolingRate = 0.1;
  const double FHe = 0.2741/C_AHe*C_AH/0.7110;                 /* pluto.h:718-720 */
  const double FZ = (1.0 - 0.7110 - 0.2741)/C_AZ*C_AH/0.7110;
  const double mu_tab = (C_AH + FHe*C_AHe + FZ*C_AZ)/(2.0 + 3.0*FHe + FZ*(1.0 + C_AZ*0.5));
  int i, j, nv;
  for (j = NG; j < NG + w->n2; j++)
  for (i = NG; i < NG + w->n1; i++) {
    double v0[NVC], v1[NVC], k1[NVC], prs, T0, T1, err = 0.0, scrh;
    Rad R;
    int stiff;
    R.w = w; R.x1 = w->x1[i]; R.xi_s = w->xi_s[CI(w, j, i)]; R.tr_s = w->tr_s[CI(w, j, i)];
    prefactors(w, j, i, R.pre);
    for (nv = NVC; nv--;) v0[nv] = v1[nv] = k1[nv] = 0.0;
    for (nv = 0; nv < NVW; nv++) v0[nv] = v1[nv] = w->V[VI(w, nv, j, i)];
    prs = v0[PRS];
    T0 = v0[PRS]/v0[RHO]*w->KELVIN*mu_tab;
    v0[RHOE] = v1[RHOE] = prs/(c->gamma - 1.0);
    radiat(&R, v0, k1);
    stiff = (fabs(k1[RHOE]) > 0.5/dt ? 1 : 0);
    if (!stiff) {
      err = rkf12(&R, v0, k1, v1, dt);
      err /= min_tol;
    }
    if (stiff || err > 1.0) {
      double dt1, dts, t1 = 0.0;
      cnt[CNT_CK45_CALLS]++;
      int done = 0;
      dt1 = 0.5/(fabs(k1[RHOE]) + 1.0/dt);
      do {
        if ((t1 + dt1) >= dt) { dt1 = dt - t1; done = 1; }
        dts = ck45(&R, v0, k1, v1, dt1, min_tol);
        t1 += dt1;
        dt1 = dts;
        if (done) break;
        for (nv = NVC; nv--;) v0[nv] = v1[nv];
        radiat(&R, v0, k1);
      } while (!done);
    }
    else cnt[CNT_RKF12_OK]++;
    prs = v1[RHOE]*(c->gamma - 1.0);
    T1 = prs/v1[RHO]*w->KELVIN*mu_tab;
    if (T1 < minT && T0 > minT) cnt[CNT_FLOOR_T]++;
    if (prs < 0.0) prs = 1.e-12;
    if (T1 < minT && T0 > minT) prs = minT*v1[RHO]/(w->KELVIN*mu_tab);
    err = fabs(prs/w->V[VI(w, PRS, j, i)] - 1.0);
    scrh = dt*maxCoolingRate/err;
    w->dt_cool = MIND(w->dt_cool, scrh);
    w->V[VI(w, PRS, j, i)] = prs;
  }
}

static void split_source(W *w, double dt)
{
  if (w->c->cooling == ORC_COOL_BLONDIN) blondin(w, dt);
  else if (w->c->cooling == ORC_COOL_PLAIN) cooling_plain(w, dt);
}

/* ------------------------------------------------------------------ */
/* Set up the restatement's state from the configuration and the interior
 * primitive state Vin ([6][nx2][nx1]). */
static void w_init(W *pw, const orc_wind_cfg *c, const double *Vin)
{
  W w;
  int i, j, nv, a, d;
  size_t nc;
  memset(&w, 0, sizeof(w));
  w.c = c;
  w.n1 = c->nx1; w.n2 = c->nx2;
  w.t1 = c->nx1 + 2*NG; w.t2 = c->nx2 + 2*NG;
  nc = (size_t)w.t1*w.t2;
  w.x1 = c->x1; w.x1l = c->x1l; w.x1r = c->x1r; w.dx1 = c->dx1;
  w.x2 = c->x2; w.x2l = c->x2l; w.x2r = c->x2r; w.dx2 = c->dx2;
  w.UD = c->unit_density; w.UL = c->unit_length; w.UV = c->unit_velocity;
  w.UT = (w.UL/w.UV);
  w.UP = (w.UD*w.UV*w.UV);
  w.UA = (w.UV*w.UV/w.UL);
  w.KELVIN = (w.UV*w.UV*C_amu/C_kB);
  w.V = calloc(NVW*nc, sizeof(double));
  w.U = calloc(NVW*nc, sizeof(double));
  w.U0 = calloc(NVW*nc, sizeof(double));
  w.Cdt = calloc(nc, sizeof(double));
  w.xgc1 = calloc(w.t1, sizeof(double)); w.rt = calloc(w.t1, sizeof(double));
  w.inv_dx1 = calloc(w.t1, sizeof(double));
  w.xgc2 = calloc(w.t2, sizeof(double)); w.s = calloc(w.t2, sizeof(double));
  w.sp = calloc(w.t2, sizeof(double)); w.inv_dx2 = calloc(w.t2, sizeof(double));
  w.dV = calloc(nc, sizeof(double)); w.A1 = calloc(nc, sizeof(double));
  w.A2 = calloc(nc, sizeof(double));
  for (d = 0; d < 2; d++)
    for (a = 0; a < 6; a++) w.pc[d][a] = calloc(d ? w.t2 : w.t1, sizeof(double));
  w.na = c->nangles;
  w.fr = calloc(w.na*nc, sizeof(double)); w.ft = calloc(w.na*nc, sizeof(double));
  w.fp = calloc(w.na*nc, sizeof(double));
  w.roff = calloc(w.na*nc, sizeof(double)); w.toff = calloc(w.na*nc, sizeof(double));
  w.moff = calloc(w.na*nc, sizeof(double)); w.dvds = calloc(w.na*nc, sizeof(double));
  w.x11_0 = calloc(nc, sizeof(double)); w.x11_1 = calloc(nc, sizeof(double));
  w.x22_0 = calloc(nc, sizeof(double)); w.x22_1 = calloc(nc, sizeof(double));
  w.xi_s = calloc(nc, sizeof(double)); w.tr_s = calloc(nc, sizeof(double));
  w.flag = calloc(nc, 1);
  w.S = calloc(nc, sizeof(double)); w.Ue = calloc(nc, sizeof(double));
  w.U0e = calloc(nc, sizeof(double));

  geometry(&w);
  w.mid_row = midplane_row(&w);
  for (a = 0; a < w.na; a++)
    for (j = 0; j < w.n2; j++)
      for (i = 0; i < w.n1; i++) {
        size_t src = ((size_t)a*w.n2 + j)*w.n1 + i;
        w.fr[AI(&w, a, j + NG, i + NG)] = c->flux_r[src];
        w.ft[AI(&w, a, j + NG, i + NG)] = c->flux_t[src];
        w.fp[AI(&w, a, j + NG, i + NG)] = c->flux_p[src];
      }
  for (nv = 0; nv < NVW; nv++)
    for (j = 0; j < w.n2; j++)
      for (i = 0; i < w.n1; i++)
        w.V[VI(&w, nv, j + NG, i + NG)] = Vin[((size_t)nv*w.n2 + j)*w.n1 + i];
  /* read_sirocco_heatcool(flag = 0), line_connect.c:291-306 */
  for (j = NG; j < NG + w.n2; j++)
    for (i = NG; i < NG + w.n1; i++) {
      double rho = w.V[VI(&w, RHO, j, i)]*w.UD, r = w.x1[i]*w.UL;
      double nH = rho/(1.43*C_mp), lx = c->l_star*c->f_x;
      w.xi_s[CI(&w, j, i)] = lx/nH/(r*r);
      w.tr_s[CI(&w, j, i)] = c->t_x;
    }
  if (c->krad == 999 && c->alpharad == 999) {   /* line_connect.c:210-231 */
    size_t np = (size_t)c->fm_points, nint = (size_t)w.n1*w.n2, q;
    w.fm_lt = calloc(np ? np : 1, sizeof(double));
    w.fm_lm = calloc(np*nint + 1, sizeof(double));
    for (q = 0; q < np; q++) w.fm_lt[q] = log10(c->fm_t[q]);
    for (q = 0; q < np*nint; q++) w.fm_lm[q] = log10(c->fm_m[q]);
  }
  if (c->xi)                                    /* restart: the file's values */
    for (j = 0; j < w.n2; j++)
      for (i = 0; i < w.n1; i++) {
        w.xi_s[CI(&w, j + NG, i + NG)] = c->xi[(size_t)j*w.n1 + i];
        w.tr_s[CI(&w, j + NG, i + NG)] = c->t_r[(size_t)j*w.n1 + i];
      }

  *pw = w;
}

static void w_free(W *pw)
{
  W w = *pw;
  int d, a;
  free(w.V); free(w.U); free(w.U0); free(w.Cdt); free(w.xgc1); free(w.rt); free(w.inv_dx1);
  free(w.xgc2); free(w.s); free(w.sp); free(w.inv_dx2); free(w.dV); free(w.A1); free(w.A2);
  for (d = 0; d < 2; d++) for (a = 0; a < 6; a++) free(w.pc[d][a]);
  free(w.fr); free(w.ft); free(w.fp); free(w.roff); free(w.toff); free(w.moff); free(w.dvds);
  free(w.x11_0); free(w.x11_1); free(w.x22_0); free(w.x22_1); free(w.xi_s); free(w.tr_s);
  free(w.flag); free(w.S); free(w.Ue); free(w.U0e); free(w.fm_lt); free(w.fm_lm);
}

/* Copy the interior of the state back to [nv][nx2][nx1] arrays. */
static void w_out(const W *w, double *Vout, double *U_out)
{
  int i, j, nv;
  for (nv = 0; nv < NVW; nv++)
    for (j = 0; j < w->n2; j++)
      for (i = 0; i < w->n1; i++) {
        size_t o = ((size_t)nv*w->n2 + j)*w->n1 + i;
        if (Vout) Vout[o] = w->V[VI(w, nv, j + NG, i + NG)];
        if (U_out) U_out[o] = w->U[NVW*CI(w, j + NG, i + NG) + nv];
      }
}

/* ComputeUserVar, cv_idl/userdef_output.c:20-260, on the current state
 * with the dv/ds of the last VGradCalc (the stage-2 input of the last
 * step).  M_UV, t_UV, t_UV2 and dvds2 start at 0 where the reference
 * leaves them uninitialised (read only for cells whose first flux-carrying
 * bins all have dv/ds <= 0). */
static void user_vars(W *w, double *uv)
{
  const orc_wind_cfg *c = w->c;
  const size_t nint = (size_t)w->n1*w->n2;
  const double mu = c->mu, lx = c->l_star*c->f_x, tx = c->t_x, M_max = 4400.;
  const double sigma_e = C_sigmaT/(C_amu*1.18);          /* userdef_output.c:107 */
  const int tab = (c->krad == 999 && c->alpharad == 999);
  int i, j, a;
  enum { XI, T_, TR, CH, CC, LC, BC, XH, CHP, CCP, LCP, BCP, XHP, XIP, NE, NH, GR, GT, GP,
         DVDS, TT, M1, M2 };
  for (j = NG; j < NG + w->n2; j++)
  for (i = NG; i < NG + w->n1; i++) {
    size_t o = (size_t)(j - NG)*w->n1 + (i - NG);
    double v[NVW], grad[3], pre[5], xip;
    double r = w->x1[i]*w->UL, rho, T, nH, ne, ne_ratio, v_th, xi, sqrt_T, sqxi, sqsqxi;
    double M_UV = 0.0, t_UV = 0.0, t_UV2 = 0.0, dvds2 = 0.0, M_temp1 = -1.0, M_temp2 = -1.0;
    double fmax = 1e-99;
    int nv;
    for (nv = 0; nv < NVW; nv++) v[nv] = w->V[VI(w, nv, j, i)];
    rho = v[RHO]*w->UD;
    T = uv[T_*nint + o] = v[PRS]/v[RHO]*w->KELVIN*mu;
    nH = rho/(1.43*C_mp);
    uv[NH*nint + o] = nH;
    if (T < 1.5e4) ne_ratio = 1e-2 + pow(10, (-51.59417133 + 12.27740153*log10(T)));
    else if (T >= 1.5e4 && T < 3.3e4) ne_ratio = pow(10, (-3.80749689 + 0.86092628*log10(T)));
    else ne_ratio = 1.21;
    ne = nH*ne_ratio;
    uv[NE*nint + o] = ne;
    v_th = sqrt(2.0*C_kB*T/C_mp);
    if (w->t <= 1.0) {
      xi = lx/(nH*r*r);
      uv[XI*nint + o] = xi;
      uv[TR*nint + o] = tx;
    } else {
      xi = w->xi_s[CI(w, j, i)];
      uv[XI*nint + o] = xi;
      uv[TR*nint + o] = w->tr_s[CI(w, j, i)];
    }
    sqrt_T = sqrt(T);
    sqxi = sqrt(xi);
    sqsqxi = pow(xi, 0.25);
    prefactors(w, j, i, pre);
    xip = c->pre ? c->pre[(size_t)5*nint + o] : 1.0;
    uv[CHP*nint + o] = pre[0]; uv[CCP*nint + o] = pre[1]; uv[XHP*nint + o] = pre[2];
    uv[LCP*nint + o] = pre[4]; uv[BCP*nint + o] = pre[3]; uv[XIP*nint + o] = xip;
    if (w->t < 3.0) {
      uv[CH*nint + o] = (8.9e-36*xi*tx);
      uv[CC*nint + o] = (8.9e-36*xi*(4.0*T));
      uv[XH*nint + o] = (1.5e-21*(sqsqxi/sqrt_T));
      uv[LC*nint + o] = ((1e-16*exp(-1.3e5/T)/sqxi/T) + fmin(fmin(1e-24, 5e-27*sqrt_T), 1.5e-17/T));
      uv[BC*nint + o] = (3.3e-27*sqrt_T);
    } else {
      uv[CH*nint + o] = pre[0]*(8.9e-36*xi*tx);
      uv[CC*nint + o] = pre[1]*(8.9e-36*xi*(4.0*T));
      uv[XH*nint + o] = pre[2]*(1.5e-21*(sqsqxi/sqrt_T));
      uv[LC*nint + o] = pre[4]*((1e-16*exp(-1.3e5/T)/sqxi/T) +
                                fmin(fmin(1e-24, 5e-27*sqrt_T), 1.5e-17/T));
      uv[BC*nint + o] = pre[3]*(3.3e-27*sqrt_T);
    }
    line_force(w, v, grad, j, i);
    uv[GR*nint + o] = grad[0];
    uv[GT*nint + o] = grad[1];
    uv[GP*nint + o] = grad[2];
    T = v[PRS]/v[RHO]*w->KELVIN*mu;
    for (a = 0; a < w->na; a++) {
      size_t qa = AI(w, a, j, i);
      double dv = w->dvds[qa], fr = w->fr[qa], ft = w->ft[qa];
      if (dv > 0.0) {
        t_UV = sigma_e*rho*v_th/dv;
        if (tab) {
          const double *mcol = w->fm_lm + o;
          M_UV = linterp(log10(t_UV), w->fm_lt, mcol, nint, c->fm_points);
        } else {
          M_UV = c->krad*pow(t_UV, c->alpharad);
        }
        if (M_UV > M_max) M_UV = M_max;
        if (M_UV > M_temp1) M_temp1 = M_UV;
      }
      if (pow(fr, 2) + pow(ft, 2) > fmax) {
        fmax = pow(M_UV*fr, 2) + pow(M_UV*ft, 2);
        M_temp2 = M_UV;
        t_UV2 = t_UV;
        dvds2 = dv;
      }
    }
    uv[M1*nint + o] = M_temp1;
    uv[M2*nint + o] = M_temp2;
    uv[DVDS*nint + o] = dvds2;
    uv[TT*nint + o] = t_UV2;
  }
}

int orc_wind_run(const orc_wind_cfg *c, double *Vin, int nsteps, double *dt, double *t,
                 double *U_out, double *dt_hist)
{
  return orc_wind_run_uv(c, Vin, nsteps, dt, t, U_out, dt_hist, NULL);
}

int orc_wind_run_uv(const orc_wind_cfg *c, double *Vin, int nsteps, double *dt, double *t,
                    double *U_out, double *dt_hist, double *uv)
{
  W w;
  int step;
  memset(cnt, 0, sizeof(cnt));
  w_init(&w, c, Vin);
  w.t = *t;
  w.invDt_hyp = 0.0;
  w.dt_cool = 1.e38;
  orc_set_quit(ORC_QUIT_NONE);
  for (step = 0; step < nsteps; step++) {       /* main.c:215-337 */
    /* with cooling, invDt / dt_cool are reset and dt recomputed only every
       second step (main.c:406-415, :329-333); without, every step */
    int cooling = c->cooling != ORC_COOL_NONE;
    int last = 0;
    if ((*t + *dt) >= c->tstop*(1.0 - 1.e-8)) {  /* main.c:226-229 */
      *dt = (c->tstop - *t);
      last = 1;
    }
    if (dt_hist) dt_hist[step] = *dt;
    if (!cooling || (step - 1)%2 == 1) { w.invDt_hyp = 0.0; w.dt_cool = 1.e38; }
    if ((step%2) == 0) {
      advance_step(&w, *dt);
      split_source(&w, *dt);
    } else {
      split_source(&w, *dt);
      advance_step(&w, *dt);
    }
    w.t += *dt;
    *t = w.t;
    if (!cooling || step%2 == 1) {              /* NextTimeStep, main.c:521-697 */
      double dt_hyp = 1.0/w.invDt_hyp, dtn;
      dt_hyp *= c->cfl;
      dtn = dt_hyp;
      dtn = MIND(dtn, w.dt_cool);
      dtn = MIND(dtn, c->cfl_max_var*(*dt));
      if (dtn == c->cfl_max_var*(*dt)) cnt[CNT_DT_MAXVAR]++;
      else if (dtn == w.dt_cool) cnt[CNT_DT_COOL]++;
      else cnt[CNT_DT_CFL]++;
      if (c->cfl_max_var != 1.0) *dt = dtn;
      if (dtn < c->first_dt*1.e-9)              /* main.c:672-683 */
        orc_set_quit(ORC_QUIT_DT_TOO_SMALL);
    }
    /* a quit inside the step (the cooling's sub-step limit) or in
       NextTimeStep ends the run after it; so does the step that reached
       tstop */
    if (orc_last_quit() != ORC_QUIT_NONE) { step++; break; }
    if (last) { orc_set_quit(ORC_STOP_TSTOP); step++; break; }
  }

  if (uv) user_vars(&w, uv);
  w_out(&w, Vin, U_out);
  w_free(&w);
  return step;
}

int orc_wind_cool(const orc_wind_cfg *c, double *V, double dt, double t, const double *xi,
                  const double *t_r, double *dt_cool)
{
  W w;
  int i, j;
  memset(cnt, 0, sizeof(cnt));
  orc_set_quit(ORC_QUIT_NONE);
  w_init(&w, c, V);
  if (xi)
    for (j = 0; j < w.n2; j++)
      for (i = 0; i < w.n1; i++) {
        w.xi_s[CI(&w, j + NG, i + NG)] = xi[(size_t)j*w.n1 + i];
        w.tr_s[CI(&w, j + NG, i + NG)] = t_r[(size_t)j*w.n1 + i];
      }
  w.t = t;
  w.dt_cool = 1.e38;
  split_source(&w, dt);
  if (dt_cool) *dt_cool = w.dt_cool;
  w_out(&w, V, NULL);
  w_free(&w);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Exact time for one cell's energy to go from E0 = p0/(gamma-1) to      */
/* E1 = p1/(gamma-1) under dE/dt = Radiat(E) (rho, xi, T_r fixed,        */
/* tracer 1, g_time > 3): the integral of dE / f(E), by adaptive Simpson  */
/* quadrature in ln E.  Returns NaN if f vanishes or changes sign on the  */
/* path (the trajectory cannot get there).  Checks the surrogate's        */
/* refinement independently of the plain integrator.                      */
/* ------------------------------------------------------------------ */
static double ct_g(const Rad *R, double u, int *bad)
{
  double v[NVC], rhs[NVC], E = exp(u);
  memset(v, 0, sizeof(v)); memset(rhs, 0, sizeof(rhs));
  v[RHO] = R->w->V[0]; v[TRC] = 1.0; v[RHOE] = E;
  radiat(R, v, rhs);
  if (!(rhs[RHOE] != 0.0) || rhs[RHOE] != rhs[RHOE]) { *bad = 1; return 0.0; }
  return E/rhs[RHOE];
}

static double ct_simpson(const Rad *R, double a, double b, double fa, double fm, double fb,
                         double whole, double tol, int depth, int *bad, double s0)
{
  double m = 0.5*(a + b), lm = 0.5*(a + m), rm = 0.5*(m + b);
  double flm = ct_g(R, lm, bad), frm = ct_g(R, rm, bad);
  double left, right;
  if (*bad || flm*s0 <= 0.0 || frm*s0 <= 0.0) { *bad = 1; return 0.0; }
  left = (m - a)/6.0*(fa + 4.0*flm + fm);
  right = (b - m)/6.0*(fm + 4.0*frm + fb);
  if (depth > 48 || fabs(left + right - whole) <= 15.0*tol)
    return left + right + (left + right - whole)/15.0;
  return ct_simpson(R, a, m, fa, flm, fm, left, 0.5*tol, depth + 1, bad, s0) +
         ct_simpson(R, m, b, fm, frm, fb, right, 0.5*tol, depth + 1, bad, s0);
}

double orc_wind_cool_time(const orc_wind_cfg *c, double rho, double p0, double xi, double t_r,
                          double p1, double tol)
{
  static const double ones[5] = {1.0, 1.0, 1.0, 1.0, 1.0};
  return orc_wind_cool_time_pre(c, rho, p0, xi, t_r, p1, tol, ones);
}

double orc_wind_cool_time_pre(const orc_wind_cfg *c, double rho, double p0, double xi,
                              double t_r, double p1, double tol, const double *pre)
{
  W w;
  Rad R;
  double rv[1];
  double a, b, fa, fb, fm, whole, s0;
  int bad = 0;
  memset(&w, 0, sizeof(w));
  w.c = c;
  w.UD = c->unit_density; w.UL = c->unit_length; w.UV = c->unit_velocity;
  w.KELVIN = (w.UV*w.UV*C_amu/C_kB);
  w.t = 5.0;
  rv[0] = rho;
  w.V = rv;
  R.w = &w; R.x1 = 1.0; R.xi_s = xi; R.tr_s = t_r;
  R.pre[0] = pre[0]; R.pre[1] = pre[1]; R.pre[2] = pre[2]; R.pre[3] = pre[3]; R.pre[4] = pre[4];
  a = log(p0/(c->gamma - 1.0));
  b = log(p1/(c->gamma - 1.0));
  if (a == b) return 0.0;
  fa = ct_g(&R, a, &bad); fb = ct_g(&R, b, &bad); fm = ct_g(&R, 0.5*(a + b), &bad);
  s0 = fa > 0.0 ? 1.0 : -1.0;
  /* f keeps its sign on the path and points from p0 towards p1 */
  if (bad || fb*fa <= 0.0 || fm*fa <= 0.0 || (b - a)*fa <= 0.0) return NAN;
  whole = (b - a)/6.0*(fa + 4.0*fm + fb);
  {
    double t = ct_simpson(&R, a, b, fa, fm, fb, whole, tol, 0, &bad, s0);
    return bad ? NAN : t;
  }
}
