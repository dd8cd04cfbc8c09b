/* PLUTO-side drop-in of libplutomx: AdvanceStep and SplitSource over
 * PLUTO's own Data / timeStep / Grid, for a PLUTO build that links this
 * object and libplutomx.so in place of the reference's rk_step.o,
 * update_stage.o and split_source.o (the object selection the reference's
 * makefiles already make between integrators).
 *
 *   int  AdvanceStep (Data *d, timeStep *Dts, Grid *grid)
 *        Src/prototypes.h:5, body Src/Time_Stepping/rk_step.c:29-384,
 *        called from Integrate (Src/main.c:476, :480, :484)
 *   void SplitSource (Data *d, double dt, timeStep *Dts, Grid *grid)
 *        Src/prototypes.h:204, body Src/split_source.c:28-79,
 *        called from Integrate (Src/main.c:475-483)
 *
 * Compiled against the problem's definitions.h and PLUTO's headers (it
 * includes pluto.h), like any other PLUTO source.  Two kinds of problem:
 *
 *   GEOMETRY CARTESIAN (HD, IDEAL, no tracers, no body force, no cooling):
 *     the Cartesian hydro step (pmx_hydro_*).  RECONSTRUCTION LINEAR with
 *     LIMITER VANLEER_LIM / MC_LIM / MINMOD_LIM / DEFAULT, or PARABOLIC
 *     (PPM_ORDER 4); TIME_STEPPING RK2 / RK3; [Solver] hllc or roe;
 *     periodic, outflow or reflective sides.
 *   GEOMETRY SPHERICAL with LINE_DRIVEN_WIND: the line-driven disk wind
 *     (pmx_wind_*), 2-D, one tracer, COOLING NO / BLONDIN / TABULATED (the
 *     cv_idl CoolingSource); userdef sides are the cv_idl problem's
 *     (Test_Problems/LineDrivenWind/cv_idl/init.c:155-362) with its
 *     INTERNAL_BOUNDARY floors (parameter DFLOOR).
 *
 * Every call hands the host state over (Vc in, Vc and Uc out), so that
 * everything else PLUTO does between steps (output, analysis, restart)
 * sees the arrays it expects.  A library error prints the library's
 * message and quits, as the reference does with QUIT_PLUTO; so do the
 * library's run-stopping conditions that the reference quits on (a
 * Cash-Karp cooling integration over its sub-step limit).
 */
#include "pluto.h"
#include "plutomx.h"

#include <stdlib.h>
#include <string.h>

static void pmx_glue_fail(const char *what)
{
  printLog("! libplutomx: %s: %s\n", what, pmx_last_error());
  QUIT_PLUTO(1);
}

#define PMX_CALL(x) do { if ((x) != 0) pmx_glue_fail(#x); } while (0)

/* FailSafe (Src/failsafe.c:5-158) retries a step whose cons->prim failed
 * from the saved state, with HLL / flat / minmod flags around the failed
 * zones; it needs AdvanceStep to return at the first failing stage
 * (rk_step.c:186-188, :262-264, :323-325) and honour per-zone flags the
 * device step does not read.  Refused at compile time rather than run
 * without the retry. */
#if FAILSAFE == YES
 #error "libplutomx: FAILSAFE YES is not supported (the device step has no FailSafe retry)"
#endif

/* ------------------------------------------------------------------ */
#if GEOMETRY == CARTESIAN && !(defined LINE_DRIVEN_WIND && LINE_DRIVEN_WIND != NO)
/* ------------------------------------------------------------------ */

#if PHYSICS != HD || EOS != IDEAL || NTRACER != 0 || COOLING != NO
 #error "libplutomx: the Cartesian path is HD, IDEAL, without tracers or cooling"
#endif
#if defined BODY_FORCE && BODY_FORCE != NO
 #error "libplutomx: the Cartesian path has no body force"
#endif

static int pmx_bc(int b)
{
  if (b == PERIODIC) return PMX_BC_PERIODIC;
  if (b == OUTFLOW) return PMX_BC_OUTFLOW;
  if (b == REFLECTIVE) return PMX_BC_REFLECTIVE;
  printLog("! libplutomx: boundary type %d is not supported\n", b);
  QUIT_PLUTO(1);
  return -1;
}

/* The context of this run, made at the first step from the grid and the
 * run-time set-up (Src/initialize.c, Src/set_grid.c). */
static pmx_hydro *hydro_context(Grid *grid)
{
  static pmx_hydro *h = NULL;
  pmx_hydro_desc desc;
  Runtime *rt = RuntimeGet();
  int dir;

  if (h != NULL) return h;
  memset(&desc, 0, sizeof(desc));
  desc.dims = DIMENSIONS;
  desc.nx1 = NX1; desc.nx2 = NX2; desc.nx3 = NX3;
#if RECONSTRUCTION == LINEAR
  #if LIMITER == VANLEER_LIM
  desc.recon = PMX_REC_PLM_VANLEER;
  #elif LIMITER == MC_LIM
  desc.recon = PMX_REC_PLM_MC;
  #elif LIMITER == MINMOD_LIM
  desc.recon = PMX_REC_PLM_MINMOD;
  #elif LIMITER == DEFAULT
  desc.recon = PMX_REC_PLM_DEFAULT;
  #else
   #error "libplutomx: LIMITER must be VANLEER_LIM, MC_LIM, MINMOD_LIM or DEFAULT"
  #endif
  if (grid->nghost[IDIR] != 2) pmx_glue_fail("PLM runs with 2 ghost zones");
#elif RECONSTRUCTION == PARABOLIC
  #if PPM_ORDER != 4
   #error "libplutomx: PPM_ORDER 4 only"
  #endif
  desc.recon = PMX_REC_PPM;
  if (grid->nghost[IDIR] != 3) pmx_glue_fail("PPM runs with 3 ghost zones");
#else
 #error "libplutomx: RECONSTRUCTION LINEAR or PARABOLIC"
#endif
#if TIME_STEPPING == RK2
  desc.rk = 2;
#elif TIME_STEPPING == RK3
  desc.rk = 3;
#else
 #error "libplutomx: TIME_STEPPING RK2 or RK3"
#endif
  if (!strcmp(rt->solv_type, "hllc")) desc.solver = PMX_SOLVER_HLLC;
  else if (!strcmp(rt->solv_type, "roe")) desc.solver = PMX_SOLVER_ROE;
  else {
    printLog("! libplutomx: [Solver] %s is not supported (hllc, roe)\n", rt->solv_type);
    QUIT_PLUTO(1);
  }
  for (dir = 0; dir < 3; dir++) {
    desc.bc[2*dir]     = dir < DIMENSIONS ? pmx_bc(rt->left_bound[dir])  : PMX_BC_PERIODIC;
    desc.bc[2*dir + 1] = dir < DIMENSIONS ? pmx_bc(rt->right_bound[dir]) : PMX_BC_PERIODIC;
  }
  desc.gamma = g_gamma;
  desc.cfl = rt->cfl;
  desc.cfl_max_var = rt->cfl_max_var;
  desc.first_dt = rt->first_dt;
  /* interior cell widths grid->dx[d][IBEG..IEND] (Src/structs.h:133-187) */
  desc.dx1 = grid->dx[IDIR] + IBEG;
  desc.dx2 = grid->dx[JDIR] + JBEG;
  desc.dx3 = grid->dx[KDIR] + KBEG;
  PMX_CALL(pmx_hydro_create(&desc, &h));
  return h;
}

int AdvanceStep (Data *d, timeStep *Dts, Grid *grid)
{
  pmx_hydro *h = hydro_context(grid);
  static double *ucons = NULL;
  double inv = Dts->invDt_hyp;   /* the stage-1 maximum folds into it */
  int i, j, k, nv;
  long q = 0;

  if (ucons == NULL) ucons = (double *)malloc(sizeof(double)*NVAR*NX1*NX2*NX3);
  if (ucons == NULL) pmx_glue_fail("host buffer");
  PMX_CALL(pmx_hydro_set_prim(h, &d->Vc[0][0][0][0], 1));
  PMX_CALL(pmx_hydro_advance_step(h, g_dt, &inv));
  PMX_CALL(pmx_hydro_get_prim(h, &d->Vc[0][0][0][0], 1));
  PMX_CALL(pmx_hydro_get_cons(h, ucons, 1));
  DOM_LOOP(k,j,i) {
    NVAR_LOOP(nv) d->Uc[k][j][i][nv] = ucons[q*NVAR + nv];
    q++;
  }
  Dts->invDt_hyp = inv;
  g_maxMach = 0.0;   /* not tracked on the device (log output only) */
  return 0;
}

void SplitSource (Data *d, double dt, timeStep *Dts, Grid *grid)
{
  /* no cooling and no parabolic terms on this path: nothing to do
     (Src/split_source.c:28-79 with COOLING NO) */
}

/* ------------------------------------------------------------------ */
#elif GEOMETRY == SPHERICAL && defined LINE_DRIVEN_WIND && LINE_DRIVEN_WIND != NO
/* ------------------------------------------------------------------ */

#if DIMENSIONS != 2 || NTRACER != 1 || TIME_STEPPING != RK2 || RECONSTRUCTION != LINEAR
 #error "libplutomx: the wind path is 2-D, one tracer, RK2, PLM"
#endif

/* interior [nv][nx2][nx1] buffers of the hand-over */
static double *gbuf_v = NULL, *gbuf_u = NULL, *gbuf_a = NULL, *gbuf_b = NULL;

static void gather3(double ***a, double *out)
{
  int i, j, k = KBEG;
  long q = 0;
  JDOM_LOOP(j) IDOM_LOOP(i) out[q++] = a[k][j][i];
}

static int pmx_side(int b, int dir, int beg)
{
  if (b == OUTFLOW) return PMX_BC_OUTFLOW;
  if (b == REFLECTIVE) return PMX_BC_REFLECTIVE;
  if (b == PERIODIC) return PMX_BC_PERIODIC;
  if (b == USERDEF && beg && dir == IDIR) return PMX_BC_WIND_INNER;
  if (b == USERDEF && beg && dir == JDIR) return PMX_BC_WIND_POLE;
  printLog("! libplutomx: boundary type %d on side %d of direction %d is not supported\n",
           b, beg ? 0 : 1, dir);
  QUIT_PLUTO(1);
  return -1;
}

static pmx_wind *wind_context(Data *d, Grid *grid)
{
  static pmx_wind *w = NULL;
  static double *fm_t = NULL, *fm_m = NULL;
  pmx_wind_desc desc;
  Runtime *rt = RuntimeGet();
  int a, i, j, k = KBEG;
  int g1 = grid->nghost[IDIR] - 2, g2 = grid->nghost[JDIR] - 2;   /* the descriptor takes 2 */
  long q, n = (long)NX1*NX2;
  double *fr, *ft, *fp;

  if (w != NULL) return w;
  memset(&desc, 0, sizeof(desc));
  desc.nx1 = NX1; desc.nx2 = NX2;
#if COOLING == NO
  desc.cooling = PMX_COOL_NONE;
#elif COOLING == BLONDIN
  desc.cooling = PMX_COOL_BLONDIN;
#elif COOLING == TABULATED
  desc.cooling = PMX_COOL_PLAIN;   /* the cv_idl CoolingSource / Radiat */
#else
 #error "libplutomx: COOLING NO, BLONDIN or TABULATED (cv_idl)"
#endif
  desc.bc[0] = pmx_side(rt->left_bound[IDIR], IDIR, 1);
  desc.bc[1] = pmx_side(rt->right_bound[IDIR], IDIR, 0);
  desc.bc[2] = pmx_side(rt->left_bound[JDIR], JDIR, 1);
  desc.bc[3] = pmx_side(rt->right_bound[JDIR], JDIR, 0);
  desc.gamma = g_gamma;
  desc.cfl = rt->cfl;
  desc.cfl_max_var = rt->cfl_max_var;
  desc.first_dt = rt->first_dt;
  desc.mu = g_inputParam[MU];
  desc.l_star = g_inputParam[L_star];
  desc.f_xray = g_inputParam[f_x];
  desc.t_x = g_inputParam[T_x];
  desc.krad = g_inputParam[KRAD];
  desc.alpharad = g_inputParam[ALPHARAD];
  desc.cent_mass = g_inputParam[CENT_MASS];
  desc.unit_density = UNIT_DENSITY;
  desc.unit_length = UNIT_LENGTH;
  desc.unit_velocity = UNIT_VELOCITY;
  desc.x1 = grid->x[IDIR] + g1;  desc.x1l = grid->xl[IDIR] + g1;
  desc.x1r = grid->xr[IDIR] + g1; desc.dx1 = grid->dx[IDIR] + g1;
  desc.x2 = grid->x[JDIR] + g2;  desc.x2l = grid->xl[JDIR] + g2;
  desc.x2r = grid->xr[JDIR] + g2; desc.dx2 = grid->dx[JDIR] + g2;
#if INTERNAL_BOUNDARY == YES
  desc.dfloor = g_inputParam[DFLOOR];
  desc.rho_0 = g_inputParam[RHO_0];
  desc.rho_alpha = g_inputParam[RHO_ALPHA];
  desc.disk_mdot = g_inputParam[DISK_MDOT];
#endif
#if SHOCK_FLATTENING == MULTID
  desc.shock_flattening = 1;
#endif
#if CHAR_LIMITING == YES
  desc.char_limiting = 1;
#endif
#if ENTROPY_SWITCH == ALWAYS
  desc.entropy_switch = 1;
#elif ENTROPY_SWITCH != NO
 #error "libplutomx: ENTROPY_SWITCH NO or ALWAYS"
#endif
  if (!strcmp(rt->solv_type, "hllc")) desc.solver = PMX_SOLVER_HLLC;
  else if (!strcmp(rt->solv_type, "hll")) desc.solver = PMX_SOLVER_HLL;
  else if (!strcmp(rt->solv_type, "roe")) desc.solver = PMX_SOLVER_ROE;
  else {
    printLog("! libplutomx: [Solver] %s is not supported (hllc, hll, roe)\n", rt->solv_type);
    QUIT_PLUTO(1);
  }
  desc.force_power = PMX_FORCE_POWER_REF;
  /* the SIROCCO fluxes read_sirocco_fluxes() left in flux_{r,t,p}_UV
     [angle][k][j][i] (Src/LineDriven/line_connect.c:43-165), interior */
  desc.nangles = NFLUX_ANGLES;
  fr = (double *)malloc(sizeof(double)*3*NFLUX_ANGLES*n);
  if (fr == NULL) pmx_glue_fail("host buffer");
  ft = fr + NFLUX_ANGLES*n;
  fp = ft + NFLUX_ANGLES*n;
  for (a = 0; a < NFLUX_ANGLES; a++) {
    q = a*n;
    JDOM_LOOP(j) IDOM_LOOP(i) {
      fr[q] = flux_r_UV[a][k][j][i];
      ft[q] = flux_t_UV[a][k][j][i];
      fp[q] = flux_p_UV[a][k][j][i];
      q++;
    }
  }
  desc.flux_r = fr; desc.flux_t = ft; desc.flux_p = fp;
  if (desc.krad == 999 && desc.alpharad == 999) {
    /* the tabulated force multiplier: the fits of M_UV_data.dat as the
       file holds them (the reference keeps their log10), read again */
    int found = 0, pts = 0;
    long matched = 0;
    PMX_CALL(pmx_sirocco_fm_points(NULL, &pts));
    if (pts <= 0) pmx_glue_fail("M_UV_data.dat");
    fm_t = (double *)malloc(sizeof(double)*pts);
    fm_m = (double *)calloc((size_t)pts*n, sizeof(double));
    if (fm_t == NULL || fm_m == NULL) pmx_glue_fail("host buffer");
    PMX_CALL(pmx_sirocco_read_force_multiplier(NULL, &desc, pts, fm_t, fm_m, &found,
                                               &matched));
    desc.fm_points = pts;
    desc.fm_t = fm_t; desc.fm_m = fm_m;
  }
  PMX_CALL(pmx_wind_create(&desc, &w));
  free(fr);
  free(fm_t); free(fm_m);
  gbuf_v = (double *)malloc(sizeof(double)*6*n);
  gbuf_u = (double *)malloc(sizeof(double)*6*n);
  gbuf_a = (double *)malloc(sizeof(double)*6*n);
  gbuf_b = (double *)malloc(sizeof(double)*n);
  if (!gbuf_v || !gbuf_u || !gbuf_a || !gbuf_b) pmx_glue_fail("host buffer");
  return w;
}

/* Host Vc (interior, the 6 HD + tracer variables) and the time state to
 * the device, before a step or a cooling call. */
static void wind_upload(pmx_wind *w, Data *d, timeStep *Dts, double dt)
{
  int i, j, k = KBEG, nv;
  long q = 0, n = (long)NX1*NX2;
  for (nv = 0; nv < 6; nv++) JDOM_LOOP(j) IDOM_LOOP(i) gbuf_v[q++] = d->Vc[nv][k][j][i];
  PMX_CALL(pmx_wind_set_prim(w, gbuf_v));
  PMX_CALL(pmx_wind_set_time(w, g_time, dt, g_stepNumber));
  PMX_CALL(pmx_wind_set_dts(w, Dts->invDt_hyp, Dts->dt_cool));
#if COOLING != NO
  /* the SIROCCO heating / cooling inputs (read_sirocco_heatcool,
     line_connect.c:267-497), as the cooling reads them */
  gather3(d->sirocco_xi, gbuf_a);
  gather3(d->sirocco_t_r, gbuf_b);
  PMX_CALL(pmx_wind_set_heatcool(w, gbuf_a, gbuf_b));
  gather3(d->comp_h_pre, gbuf_a);
  gather3(d->comp_c_pre, gbuf_a + n);
  gather3(d->xray_h_pre, gbuf_a + 2*n);
  gather3(d->brem_c_pre, gbuf_a + 3*n);
  gather3(d->line_c_pre, gbuf_a + 4*n);
  gather3(d->xi_ion_pre, gbuf_a + 5*n);
  PMX_CALL(pmx_wind_set_prefactors(w, gbuf_a));
#endif
}

static void wind_download(pmx_wind *w, Data *d, timeStep *Dts, int cons)
{
  int i, j, k = KBEG, nv, flags = 0;
  long q = 0;
  double inv, dtc;
  PMX_CALL(pmx_wind_sync(w));
  PMX_CALL(pmx_wind_get_prim(w, gbuf_v));
  for (nv = 0; nv < 6; nv++) JDOM_LOOP(j) IDOM_LOOP(i) d->Vc[nv][k][j][i] = gbuf_v[q++];
  if (cons) {
    long n = (long)NX1*NX2, c = 0;
    PMX_CALL(pmx_wind_get_cons(w, gbuf_u));
    JDOM_LOOP(j) IDOM_LOOP(i) {
      for (nv = 0; nv < 6; nv++) d->Uc[k][j][i][nv] = gbuf_u[nv*n + c];
      c++;
    }
  }
  PMX_CALL(pmx_wind_get_time(w, NULL, NULL, NULL, &inv, &dtc, &flags));
  Dts->invDt_hyp = inv;
  Dts->dt_cool = dtc;
  if (flags & PMX_FLAG_COOL_SUBSTEPS) {
    /* Src/Cooling/cooling_ode_solver.c:190-192 */
    printLog("! SolveODE_CK45: Number of substeps too large\n");
    QUIT_PLUTO(1);
  }
}

int AdvanceStep (Data *d, timeStep *Dts, Grid *grid)
{
  pmx_wind *w = wind_context(d, grid);
  wind_upload(w, d, Dts, g_dt);
  PMX_CALL(pmx_wind_advance_step(w));
  wind_download(w, d, Dts, 1);
  g_maxMach = 0.0;   /* not tracked on the device (log output only) */
  return 0;
}

void SplitSource (Data *d, double dt, timeStep *Dts, Grid *grid)
{
#if COOLING != NO
  pmx_wind *w = wind_context(d, grid);
  wind_upload(w, d, Dts, dt);
  PMX_CALL(pmx_wind_split_source(w));
  wind_download(w, d, Dts, 0);
#endif
}

#else
 #error "libplutomx: a Cartesian HD problem or the spherical line-driven wind"
#endif
