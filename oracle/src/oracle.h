/* Oracle: a plain-C CPU restatement of PLUTO-sirocco's finite-volume hydro
 * step, used ONLY as the checker of the GPU library in tests/, in the
 * package smoke test and in bench.py's cpu_baseline leg.  It is never
 * linked into, loaded by or called from the product path.
 *
 * Each routine cites the reference file:line whose arithmetic it follows;
 * the operation order matches the reference so that results agree bit for
 * bit with the compiled reference (oracle/_ref) on the same inputs.
 *
 * Array layout (all interior-only, float64):
 *     V[nv][k][j][i]   primitive variables  (rho, vx1, vx2, vx3, p)
 *     U[nv][k][j][i]   conservative variables (rho, m1, m2, m3, E)
 */
#ifndef PMX_ORACLE_H
#define PMX_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_PLM = 0, ORC_PPM = 1 };
enum { ORC_LIM_DEFAULT = 0, ORC_LIM_VANLEER = 1, ORC_LIM_MC = 2,
       ORC_LIM_MINMOD = 3, ORC_LIM_FLAT = 4 };
enum { ORC_HLLC = 0, ORC_ROE = 1 };
enum { ORC_BC_PERIODIC = 0, ORC_BC_OUTFLOW = 1, ORC_BC_REFLECTIVE = 2,
       ORC_BC_WIND_INNER = 3,   /* cv_idl X1-beg userdef (cv_idl/init.c:320-330) */
       ORC_BC_WIND_POLE = 4 };  /* cv_idl X2-beg userdef (cv_idl/init.c:346-359) */

typedef struct {
  int    dims;          /* 2 or 3                                        */
  int    nx[3];         /* interior cells per direction (nx[2]=1 in 2-D) */
  int    recon;         /* ORC_PLM / ORC_PPM                              */
  int    limiter;       /* PLM limiter (ignored by PPM)                   */
  int    solver;        /* ORC_HLLC / ORC_ROE                             */
  int    rk;            /* 2 or 3                                         */
  int    bc[6];         /* x1beg, x1end, x2beg, x2end, x3beg, x3end       */
  double gamma, cfl, cfl_max_var;
  double xbeg[3], xend[3];
  double tstop;         /* runtime.tstop (1e30 in the test configurations)  */
  double first_dt;      /* runtime.first_dt (NextTimeStep's "too small")    */
} orc_cfg;

/* Why the last run stopped before nsteps (orc_last_quit): the reference's
 * QUIT_PLUTO points on this path.  ORC_STOP_TSTOP: the step that reached
 * tstop was taken (the main loop ends there). */
enum { ORC_QUIT_NONE = 0, ORC_QUIT_DT_TOO_SMALL = 1, ORC_QUIT_COOL_SUBSTEPS = 2,
       ORC_STOP_TSTOP = 3 };
int orc_last_quit(void);

/* Advance V (interior, in/out) by nsteps time steps exactly as PLUTO's
 * main loop does (Integrate + NextTimeStep, Src/main.c:215-337), or until
 * the loop ends earlier: the step that reaches tstop has its dt cut to
 * tstop - t and is the last (main.c:226-229); NextTimeStep's "dt is too
 * small" quits after the step (main.c:672-683).
 *   *dt   in: dt of the first step; out: dt of the next step
 *   *t    in/out: simulation time
 *   U_out optional: Uc after the last step (interior, [nv][k][j][i])
 *   dt_hist optional: nsteps entries, dt used at each step
 * Returns the number of steps taken, or a negative value on bad
 * configuration. */
int orc_run(const orc_cfg *cfg, double *V, int nsteps, double *dt, double *t,
            double *U_out, double *dt_hist);

/* Pencil-level entry points (used by tests to pin single functions). */
int  orc_nghost(const orc_cfg *cfg);
void orc_prim_to_cons(double gamma, const double *v, double *u);
int  orc_cons_to_prim(double gamma, double *u, double *v);
/* One HLLC / Roe interface: vL,vR primitive states (5), dir 0..2.
 * Outputs flux[5], *press, *cmax. */
void orc_riemann(int solver, double gamma, int dir, const double *vL,
                 const double *vR, double *flux, double *press, double *cmax);

/* ------------------------------------------------------------------ */
/* Line-driven disk wind: 2-D spherical (r, theta), PLM (van Leer,       */
/* general-grid weights) + HLLC + RK2, one passive tracer, point-mass    */
/* gravity, SIROCCO-mode line force, split cooling.                      */
/* V has 6 variables: rho, v_r, v_theta, v_phi, p, tracer.                */
/* ------------------------------------------------------------------ */
enum { ORC_COOL_NONE = 0, ORC_COOL_BLONDIN = 1, ORC_COOL_PLAIN = 2 };

typedef struct {
  int    nx1, nx2;            /* interior cells (r, theta)                    */
  int    cooling;             /* ORC_COOL_*                                    */
  int    bc[4];               /* r-beg, r-end, theta-beg, theta-end            */
  double gamma, cfl, cfl_max_var;
  double mu, t_iso, l_star, f_x, f_uv, t_x, krad, alpharad, cent_mass;
  double unit_density, unit_length, unit_velocity;
  /* grid with 2 ghost cells per side: nx + 4 entries each */
  const double *x1, *x1l, *x1r, *dx1;
  const double *x2, *x2l, *x2r, *dx2;
  int    nangles;             /* 36 */
  const double *flux_r, *flux_t, *flux_p;   /* [nangles][nx2][nx1]            */
  double dfloor;              /* > 0: cv_idl internal-boundary floors (g/cm^3) */
  int    shock_flattening;    /* 1: SHOCK_FLATTENING MULTID (flag_shock.c)     */
  int    char_limiting;       /* 1: CHAR_LIMITING YES (plm_states.c:479-746)   */
  int    entropy_switch;      /* 1: ENTROPY_SWITCH ALWAYS (pressure from the
                                 advected entropy, HD/mappers.c:186-204)      */
  int    solver;              /* 0: HLLC, 1: Roe, 2: HLL (Src/HD/hll.c)       */
  /* What read_sirocco_heatcool(flag = 1) leaves on a restart
   * (line_connect.c:308-497); NULL = its start-up values:                   */
  const double *xi, *t_r;     /* [nx2][nx1] sirocco_xi, sirocco_t_r (read by
                                 the cooling once g_time > 3)                */
  const double *pre;          /* [6][nx2][nx1] comp_h, comp_c, xray_h, brem_c,
                                 line_c, xi_ion prefactors (NULL: all 1.0)   */
  /* krad = alpharad = 999: the tabulated force multiplier
   * (line_connect.c:183-257, :858-887)                                      */
  int    fm_points;
  const double *fm_t;         /* [fm_points] t of the fits, as in the file   */
  const double *fm_m;         /* [fm_points][nx2][nx1] M per cell, as in the
                                 file (the reader keeps log10 of both)       */
  double tstop;               /* runtime.tstop                                */
  double first_dt;            /* runtime.first_dt                             */
  /* the cv_idl disk of the midplane row (parameters RHO_0 [g/cm^3],
   * RHO_ALPHA, DISK_MDOT [g/s]; cv_idl/init.c:276-311) */
  double rho_0, rho_alpha, disk_mdot;
} orc_wind_cfg;

/* Advance V ([6][nx2][nx1], in/out) by nsteps steps of the reference's
 * main loop with cooling (Src/main.c:215-337, :383-485, :521-697) starting
 * at step number 0, or until it ends earlier as in orc_run; a Cash-Karp
 * integration over 1000 sub-steps (cooling_ode_solver.c:190-192) also
 * quits, after the step.  dt/t in/out as in orc_run.  U_out optional.
 * Returns the number of steps taken. */
int orc_wind_run(const orc_wind_cfg *cfg, double *V, int nsteps, double *dt,
                 double *t, double *U_out, double *dt_hist);

/* orc_wind_run, and then the wind problem's ComputeUserVar
 * (cv_idl/userdef_output.c:20-260) on the final state: uv[23][nx2][nx1] in
 * the order of ORC_UV_NAMES (the production problem's uservar list). */
#define ORC_NUV 23
#define ORC_UV_NAMES "XI T T_r ch cc lc bc xh ch_pre cc_pre lc_pre bc_pre xh_pre XI_pre " \
                     "ne nh gr gt gp dv_ds t M_max1 M_max2"
int orc_wind_run_uv(const orc_wind_cfg *cfg, double *V, int nsteps, double *dt, double *t,
                    double *U_out, double *dt_hist, double *uv);

/* One SplitSource (cooling) call on V ([6][nx2][nx1], in/out) at time t
 * with step dt.  xi / t_r ([nx2][nx1]) are the SIROCCO arrays, or NULL for
 * the start-up values derived from V.  *dt_cool receives the plain
 * integrator's limit (1e38 if none). */
int orc_wind_cool(const orc_wind_cfg *cfg, double *V, double dt, double t, const double *xi,
                  const double *t_r, double *dt_cool);

/* Exact time for a cell's energy to go from p0 to p1 (code units) under
 * the cooling ODE dE/dt = Radiat(E) with density rho, ionisation parameter
 * xi and T_r fixed (tracer 1, g_time > 3); adaptive Simpson in ln E to the
 * absolute tolerance tol.  NaN if the rate vanishes or changes sign on the
 * way (p1 cannot be reached). */
double orc_wind_cool_time(const orc_wind_cfg *cfg, double rho, double p0, double xi, double t_r,
                          double p1, double tol);
/* The same with the cell's five cooling prefactors pre[5] (comp_h, comp_c,
 * xray_h, brem_c, line_c: cv_idl/radiat.c:118-123) in place of 1. */
double orc_wind_cool_time_pre(const orc_wind_cfg *cfg, double rho, double p0, double xi,
                              double t_r, double p1, double tol, const double *pre);

/* Branch counters of the last orc_wind_run / orc_wind_cool: rkf12 accepted, ck45 calls,
 * ck45 substeps, ck45 rejections, BLONDIN solves, bracket widenings,
 * second (heat-cool root) solves, cells below the 1e4 K cut, dt limited by
 * CFL / cooling / CFL_max_var, cons->prim failures, temperature floors,
 * cells changed by the internal-boundary floors, zones flagged as shocked,
 * zones reset by the cv_idl midplane row.
 * Copies min(n, count) entries; returns the count. */
int orc_wind_counters(long *out, int n);

#ifdef __cplusplus
}
#endif
#endif
