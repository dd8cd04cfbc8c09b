/* libplutomx: MI355X (gfx950) implementation of PLUTO-sirocco's
 * finite-volume hydro step behind a plain C ABI.
 *
 * All arrays are float64.  Primitive variables are (rho, v1, v2, v3, p) and
 * conservative variables (rho, m1, m2, m3, E), as in the reference HD module
 * (Src/HD/mod_defs.h:19-41).  Unless stated otherwise host arrays are
 * interior-only, variable-major: A[nv][k][j][i] with i fastest.
 *
 * Every function returns 0 on success and -1 on error; pmx_last_error()
 * then describes the failure.  A context owns its device buffers and one
 * HIP stream; calls on one context must not overlap.
 */
#ifndef PLUTOMX_H
#define PLUTOMX_H

#ifdef __cplusplus
extern "C" {
#endif

/* reconstruction (definitions.h RECONSTRUCTION / LIMITER) */
#define PMX_REC_PLM_VANLEER 0   /* LINEAR + VANLEER_LIM   */
#define PMX_REC_PLM_MC      1   /* LINEAR + MC_LIM        */
#define PMX_REC_PLM_MINMOD  2   /* LINEAR + MINMOD_LIM    */
#define PMX_REC_PLM_DEFAULT 3   /* LINEAR + DEFAULT       */
#define PMX_REC_PPM         5   /* PARABOLIC, PPM_ORDER 4 */
/* Riemann solver (pluto.ini [Solver]) */
#define PMX_SOLVER_HLLC 0
#define PMX_SOLVER_ROE  1
#define PMX_SOLVER_HLL  2   /* the wind only: Src/HD/hll.c, the production problem's [Solver] */
/* boundary types (pluto.ini [Boundary]) */
#define PMX_BC_PERIODIC   0
#define PMX_BC_OUTFLOW    1
#define PMX_BC_REFLECTIVE 2
/* The wind's userdef sides of the production problem
 * (Test_Problems/LineDrivenWind/cv_idl/init.c): */
#define PMX_BC_WIND_INNER 3   /* X1-beg: copy of the first cell, v_r = MIN(v_r, 0) (:318-329) */
#define PMX_BC_WIND_POLE  4   /* X2-beg: reflective for the velocities and the tracer, rho and
                                 p copied from the first cell (:343-359) */

typedef struct pmx_hydro pmx_hydro;

/* Configuration of one grid.  Mirrors what the reference reads from
 * definitions.h (DIMENSIONS, RECONSTRUCTION, LIMITER, TIME_STEPPING) and
 * pluto.ini ([Grid], [Time] CFL / CFL_max_var, [Solver], [Boundary]);
 * dx1..dx3 are the interior cell widths grid->dx[d][IBEG..IEND]
 * (Src/structs.h:133-187).  Geometry is Cartesian. */
typedef struct {
  int dims;              /* 2 or 3 */
  int nx1, nx2, nx3;     /* interior cells (nx3 = 1 in 2-D) */
  int recon;             /* PMX_REC_*    */
  int solver;            /* PMX_SOLVER_* */
  int rk;                /* 2 or 3 (TIME_STEPPING RK2 / RK3) */
  int bc[6];             /* x1-beg, x1-end, x2-beg, x2-end, x3-beg, x3-end */
  double gamma;          /* g_gamma */
  double cfl;            /* runtime->cfl */
  double cfl_max_var;    /* runtime->cfl_max_var */
  const double *dx1, *dx2, *dx3;   /* host arrays, nx1 / nx2 / nx3 entries */
  double first_dt;       /* runtime->first_dt: > 0 enables NextTimeStep's "dt is too
                            small" stop (dt < 1e-9 first_dt, Src/main.c:672-683),
                            reported as PMX_FLAG_DT_TOO_SMALL */
} pmx_hydro_desc;

/* Flags that *_get_time return.  The run-stopping ones (DT_TOO_SMALL,
 * COOL_SUBSTEPS, TSTOP) are sticky and halt the main loop: once one is set,
 * further steps of *_run / *_run_graph / pmx_wind_batch_run do nothing (t,
 * dt and the state stay where the run stopped) until *_set_time is called.
 * The single-step entries (*_advance_step, pmx_wind_split_source), which a
 * driver calls with its own dt, are not halted. */
/* NextTimeStep produced a dt below 1e-9 first_dt: the point where the
 * reference prints "dt is too small" and quits (Src/main.c:672-683).  The
 * state after it is not meaningful. */
#define PMX_FLAG_DT_TOO_SMALL 4096
/* The step that reached tstop has been taken: the reference's main loop
 * cuts that step's dt to tstop - t (Src/main.c:226-229) and ends. */
#define PMX_FLAG_TSTOP 8192
/* A cell's Cash-Karp cooling integration exceeded 1000 sub-steps: the
 * reference prints "! SolveODE_CK45: too many steps" and quits
 * (Src/Cooling/cooling_ode_solver.c:190-192), also raised when a cell's
 * integration does not end (the device caps every loop).  The run halts
 * after the step in which it happened; that step's state is not meaningful. */
#define PMX_FLAG_COOL_SUBSTEPS 256

const char *pmx_last_error(void);

/* Create a context: allocates V (3 buffers), U and the grid tables on the
 * current HIP device.  Replaces the reference's one-time set-up in
 * Initialize() (Src/initialize.c:381-443) for the hot path. */
int pmx_hydro_create(const pmx_hydro_desc *desc, pmx_hydro **out);
int pmx_hydro_destroy(pmx_hydro *h);

/* Host <-> device transfer of the primitive state Vc.  ghosts != 0 means
 * the host array is PLUTO's Vc[nv][NX3_TOT][NX2_TOT][NX1_TOT] with
 * nghost ghost cells in every active direction (Src/initialize.c:442);
 * on download the ghost cells are filled by the boundary rule. */
int pmx_hydro_set_prim(pmx_hydro *h, const double *V, int ghosts);
int pmx_hydro_get_prim(pmx_hydro *h, double *V, int ghosts);
/* Download Uc after the last stage.  pluto_layout != 0 writes the
 * reference's Uc[k][j][i][nv] interior layout, else U[nv][k][j][i]. */
int pmx_hydro_get_cons(pmx_hydro *h, double *U, int pluto_layout);

/* Time state: simulation time g_time and the dt of the next step g_dt. */
int pmx_hydro_set_time(pmx_hydro *h, double t, double dt);
int pmx_hydro_get_time(pmx_hydro *h, double *t, double *dt, double *inv_dt_hyp,
                       int *c2p_flags);
/* runtime.tstop (pluto.ini [Time] tstop; default: none).  The main loop
 * (pmx_hydro_run / _run_graph) cuts the dt of the step that would reach
 * tstop to tstop - t, as the reference does at the top of each step
 * (Src/main.c:226-229), and halts after it with PMX_FLAG_TSTOP.  Also
 * clears PMX_FLAG_TSTOP, so that a halted run continues to a later tstop.
 * set_time applies the cut to its (t, dt) at once; get_time then reports
 * the cut dt. */
int pmx_hydro_set_tstop(pmx_hydro *h, double tstop);

/* One AdvanceStep(d, Dts, grid) of the reference (Src/Time_Stepping/
 * rk_step.c:29, prototype Src/prototypes.h:5) with time step dt:
 * boundaries, all RK stages, cons->prim.  *inv_dt_hyp is Dts->invDt_hyp:
 * on entry the value the stage-1 maximum of C_dt folds into, on return
 * MAX(entry, max C_dt) / DIMENSIONS (update_stage.c:389-394); NULL: 0 on
 * entry, nothing returned.  Synchronises the stream. */
int pmx_hydro_advance_step(pmx_hydro *h, double dt, double *inv_dt_hyp);

/* No counterpart in the reference: pmx_hydro_advance_step for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  No context: desc as for
 * pmx_hydro_create; V (in: V^n, out: V^{n+1}) and U (out: Uc after the last
 * stage) are the caller's DEVICE arrays [nv][k][j][i] of the interior; dt
 * and *inv_dt_hyp as in pmx_hydro_advance_step.  The intermediate stage
 * states, the geometry tables and the time scalars live in one fixed
 * device array of the library (40 MiB on each GPU it runs on, held until
 * the process ends), so the grid may hold at most 524288 cells with
 * nx1 + nx2 + nx3 <= 8192 (larger grids are refused with an error).  The
 * call allocates nothing, runs on the null stream and waits for it.  Every
 * call on a GPU shares that GPU's array: calls that overlap on one GPU
 * (threads, or streams other than the null stream) race. */
int pmx_hydro_advance_step_noalloc(const pmx_hydro_desc *desc, double *V, double *U, double dt,
                                   double *inv_dt_hyp);

/* nsteps iterations of the reference's main loop body for this path
 * (Integrate + time update + NextTimeStep, Src/main.c:215-337,
 * :383-485, :521-697) entirely on the device: dt and t stay in device
 * memory and no host synchronisation happens between steps.
 * Asynchronous; pmx_hydro_sync() waits. */
int pmx_hydro_run(pmx_hydro *h, int nsteps);
int pmx_hydro_sync(pmx_hydro *h);

/* Same as pmx_hydro_run, but replays a captured HIP graph of 8 steps
 * (built on first use); a remainder of fewer steps runs eagerly. */
int pmx_hydro_run_graph(pmx_hydro *h, int nsteps);

/* Per-kernel timing with HIP events recorded on the context's stream
 * around every stage-kernel launch while enabled.  Reports the summed
 * duration (ms) and the number of stage launches since the last reset. */
int pmx_hydro_stage_timing(pmx_hydro *h, int enable, int reset, double *total_ms,
                           int *launches);

/* Device pointers of the current primitive state V[nv][k][j][i] and of
 * the conservative U (for zero-copy consumers such as torch). */
int pmx_hydro_device_state(pmx_hydro *h, double **V, double **U);

/* ------------------------------------------------------------------ */
/* Line-driven disk wind on a 2-D spherical (r, theta) grid.             */
/*                                                                      */
/* HD + IDEAL with one passive tracer: V = (rho, v_r, v_theta, v_phi, p, */
/* tracer), U = (rho, m_r, m_theta, m_phi, E, rho*tracer), host arrays   */
/* A[nv][j][i] (i = radius fastest), interior only.  PLM (van Leer,      */
/* general-grid weights) + HLLC + RK2, point-mass gravity, the SIROCCO-  */
/* mode line force and one split cooling step per time step.             */
/* ------------------------------------------------------------------ */
#define PMX_COOL_NONE    0
#define PMX_COOL_BLONDIN 1   /* COOLING BLONDIN: Brent root solve per cell */
#define PMX_COOL_PLAIN   2   /* cv_idl CoolingSource: RKF12 / Cash-Karp 45 */

typedef struct pmx_wind pmx_wind;
typedef struct pmx_mlp pmx_mlp;           /* the cooling surrogate, see below */

/* Mirrors the reference's set-up of the wind problem: definitions.h
 * (UNIT_*), pluto.ini ([Time] CFL / CFL_max_var, [Boundary], [Parameters]
 * MU, L_star, f_x, T_x, KRAD, ALPHARAD, CENT_MASS, GAMMA, DFLOOR) and the
 * Grid arrays x, xl, xr, dx of both directions WITH 2 ghost cells per side
 * (nx + 4 entries, Src/structs.h:133-187).  flux_r / flux_t / flux_p are
 * the SIROCCO directional fluxes (cgs) [nangles][nx2][nx1] that
 * read_sirocco_fluxes() loads (Src/LineDriven/line_connect.c:43-165). */
typedef struct {
  int nx1, nx2;                 /* interior cells (r, theta)            */
  int cooling;                  /* PMX_COOL_*                           */
  int bc[4];                    /* r-beg, r-end, theta-beg, theta-end   */
  double gamma, cfl, cfl_max_var;
  double mu, l_star, f_xray, t_x, krad, alpharad, cent_mass;   /* f_xray: f_x (a field
                                   name apart from the problem's f_x parameter macro) */
  double unit_density, unit_length, unit_velocity;
  const double *x1, *x1l, *x1r, *dx1;
  const double *x2, *x2l, *x2r, *dx2;
  int nangles;                  /* 1 .. 64 (the reference's files hold 36) */
  const double *flux_r, *flux_t, *flux_p;
  /* > 0: INTERNAL_BOUNDARY YES with the production problem's
   * UserDefBoundary(side 0) (cv_idl/init.c:196-273) at the start of every
   * stage: density floor DFLOOR (g/cm^3, velocities scaled to keep the
   * momentum, pressure to keep the sound speed), temperature floor 500 K,
   * pressure floor DFLOOR*500/(KELVIN mu), Uc of changed cells from
   * PrimToCons.  0: no internal boundary. */
  double dfloor;
  /* 1: SHOCK_FLATTENING MULTID (Src/flag_shock.c:81-253): at the first
   * stage of each step, zones with div v < 0 and a pressure jump above 5x
   * the local minimum take the HLL flux (Src/HD/hllc.c:93-104) and, with
   * their neighbours, the minmod slope (Src/States/plm_states.c:177-190).
   * The reference then runs with 3 ghost zones; the test covers the first
   * two ghost layers. */
  int shock_flattening;
  /* 1: CHAR_LIMITING YES: PLM limits the differences of the characteristic
   * variables of the primitive HD system (Src/States/plm_states.c:479-676,
   * Src/HD/eigenv.c:92-189, :575-605), then bounds the primitive slopes by
   * the primitive differences; the tracer is limited in primitive form. */
  int char_limiting;
  /* 1: ENTROPY_SWITCH ALWAYS: an entropy s = p / rho^gamma, set from the
   * bounded state before each stage (Src/entropy_switch.c:14-37), is
   * reconstructed and advected with the mass flux (Src/adv_flux.c:131-134);
   * cons->prim takes the pressure from it (Src/HD/mappers.c:186-204). */
  int entropy_switch;
  int solver;                   /* PMX_SOLVER_HLLC (0), PMX_SOLVER_HLL or PMX_SOLVER_ROE
                                   (Src/HD/roe.c, with its 2-D strong-shock switch and,
                                   next to flagged zones, its HLL branch :105-131; a
                                   batch takes HLLC and HLL grids only) */
  double first_dt;              /* as in pmx_hydro_desc (PMX_FLAG_DT_TOO_SMALL) */
  /* krad == alpharad == 999: the tabulated force multiplier of
   * M_UV_data.dat (line_connect.c:183-257, :858-887).  Per angle bin,
   * M = linterp(log10 t) over the cell's fit, 10^ of the linear
   * interpolation in log10 M between the two fit points around log10 t
   * (the end value outside the fit range), capped at 4400.  fm_t [fm_points]
   * and fm_m [fm_points][nx2][nx1] are t and M as the file holds them
   * (pmx_sirocco_read_force_multiplier); their log10 is taken once at
   * create.  Ignored for other krad / alpharad. */
  int fm_points;                /* 1 .. 4096 */
  const double *fm_t, *fm_m;
  /* How the line force evaluates M = krad t^alpha per angle bin (ignored
   * with the tabulated multiplier):
   *   PMX_FORCE_POWER_REF (0): as the reference (line_connect.c:879-889),
   *     t = sigma_e rho v_th / dvds and pow(t, alpha) for each of the two
   *     LineForce calls of a cell; the device pow, like glibc's, is within
   *     an ulp of the exact power.
   *   PMX_FORCE_POWER_FACTORED (1): both calls share one power per bin,
   *     (sigma_e rho v_th)^alpha once per cell and call times
   *     exp2(-alpha log2 dvds) per bin: a few 1e-15 relative from the
   *     reference's M, faster. */
  int force_power;
  /* The cv_idl disk (parameters RHO_0 [g/cm^3], RHO_ALPHA, DISK_MDOT [g/s]):
   * with dfloor > 0 and shock_flattening (the production set-up, whose
   * MULTID runs the reference with 3 ghost zones), the internal-boundary
   * pass of cv_idl/init.c:276-311 resets the last interior polar row at
   * every stage, after the floors, when that row's centre and volume
   * centroid agree to 1e-6 (from about 48 polar zones on): the disk
   * density rho_0 (r / r_in)^-alpha with the polar momentum kept, v_r = 0,
   * the Keplerian v_phi, the pressure of the disk temperature, tracer 1. */
  double rho_0, rho_alpha, disk_mdot;
} pmx_wind_desc;
#define PMX_FORCE_POWER_REF      0
#define PMX_FORCE_POWER_FACTORED 1

/* Create a context: uploads geometry, PLM weights and the line-force
 * tables (the one-time part of VGradCalc, line_connect.c:534-625). */
int pmx_wind_create(const pmx_wind_desc *desc, pmx_wind **out);
int pmx_wind_destroy(pmx_wind *w);

/* Primitive state V[6][nx2][nx1] host <-> device. */
int pmx_wind_set_prim(pmx_wind *w, const double *V);
int pmx_wind_get_prim(pmx_wind *w, double *V);
/* Uc after the last RK stage, U[6][nx2][nx1]. */
int pmx_wind_get_cons(pmx_wind *w, double *U);

/* SIROCCO ionisation parameter and radiation temperature per cell
 * ([nx2][nx1]), used by the cooling once g_time > 3.  With xi == NULL they
 * are set from the current state exactly as read_sirocco_heatcool(flag 0)
 * does at start-up (line_connect.c:291-306): xi = L_x / n_H / r^2,
 * t_r = T_x. */
int pmx_wind_set_heatcool(pmx_wind *w, const double *xi, const double *t_r);

/* The heating / cooling prefactors per cell, pre[6][nx2][nx1] in the
 * column order of prefactors.dat: comp_h, comp_c, xray_h, brem_c, line_c,
 * xi_ion (pmx_sirocco_read_prefactors).  CoolingSource passes the first
 * five to Radiat (cv_idl/cooling_source.c:101-105, cv_idl/radiat.c:118-123)
 * and BLONDIN cooling multiplies its rates by them
 * (Src/Cooling/BLONDIN/cooling.c:74-79, :190-195); xi_ion is read but the
 * reference uses it nowhere.  pre == NULL: all 1.0, what
 * read_sirocco_heatcool sets at start-up (line_connect.c:366-380).  Cells
 * with a prefactor other than 1 are cooled by the plain integrator also
 * when a surrogate is attached (the network is trained on unit prefactors). */
int pmx_wind_set_prefactors(pmx_wind *w, const double *pre);

/* g_time, g_dt and the step counter g_stepNumber of the next step.  Also
 * clears the run-stopping flags and cuts dt to tstop (pmx_wind_set_tstop). */
int pmx_wind_set_time(pmx_wind *w, double t, double dt, long step);
/* runtime.tstop (default: none): as pmx_hydro_set_tstop, for the wind's
 * main loop (pmx_wind_run / _run_graph, and each grid of a batch). */
int pmx_wind_set_tstop(pmx_wind *w, double tstop);
/* The timeStep fields the step reduces into: Dts->invDt_hyp (max with the
 * stage-1 C_dt, then / DIMENSIONS) and Dts->dt_cool (min with the plain
 * integrator's limits).  The reference's Integrate() resets them
 * (Src/main.c:406-415); a driver that calls pmx_wind_advance_step /
 * pmx_wind_split_source itself passes its values here first. */
int pmx_wind_set_dts(pmx_wind *w, double inv_dt_hyp, double dt_cool);
/* Also returns Dts->invDt_hyp, Dts->dt_cool and the OR of the cons->prim
 * flags and the PMX_FLAG_* bits (PMX_FLAG_COOL_SUBSTEPS: a cooling
 * integration hit its sub-step limit).  step: the steps the main loop has
 * taken since set_time, added to its step; halted steps do not count. */
int pmx_wind_get_time(pmx_wind *w, double *t, double *dt, long *step, double *inv_dt_hyp,
                      double *dt_cool, int *flags);

/* AdvanceStep (Src/Time_Stepping/rk_step.c:29, Src/prototypes.h:5) for
 * the wind: both RK2 stages with boundaries, line force, sources and the
 * stage-1 CFL reduction, using the context's dt.  Asynchronous. */
int pmx_wind_advance_step(pmx_wind *w);
/* SplitSource (Src/split_source.c:28, Src/prototypes.h:204): the cooling
 * step over the context's dt; the plain integrator also lowers dt_cool.
 * Asynchronous. */
int pmx_wind_split_source(pmx_wind *w);
/* nsteps iterations of the reference's main loop with cooling
 * (Src/main.c:215-337, :383-485, :521-697): Strang order by step parity,
 * invDt / dt_cool reset and NextTimeStep every second step, all on the
 * device.  Asynchronous; pmx_wind_sync() waits. */
int pmx_wind_run(pmx_wind *w, int nsteps);
/* Same as pmx_wind_run, but replays a captured HIP graph of 8 steps
 * (from an even step; built on first use, rebuilt after
 * pmx_wind_set_surrogate); an odd first step and a remainder of fewer
 * steps run eagerly.  Bit-identical to pmx_wind_run; with timing enabled
 * it runs eagerly. */
int pmx_wind_run_graph(pmx_wind *w, int nsteps);
int pmx_wind_sync(pmx_wind *w);
/* Timing (HIP events on the context's stream) while enabled: the RK
 * stages (each the line-force and the stage kernel) and the cooling calls,
 * summed milliseconds and counts (stage_launches counts RK stages). */
int pmx_wind_timing(pmx_wind *w, int enable, int reset, double *stage_ms, int *stage_launches,
                    double *cool_ms, int *cool_launches);

/* ------------------------------------------------------------------ */
/* Output: snapshots of the device state for the .dbl files a SIROCCO    */
/* cycle reads back, with the production problem's user variables.       */
/* ------------------------------------------------------------------ */
/* ComputeUserVar of Test_Problems/LineDrivenWind/cv_idl/userdef_output.c
 * (:20-260) for the production uservar list (cv_idl/pluto_sirocco_sub.py:86),
 * in this order: */
#define PMX_NUV 23
#define PMX_UV_NAMES "XI T T_r ch cc lc bc xh ch_pre cc_pre lc_pre bc_pre xh_pre XI_pre " \
                     "ne nh gr gt gp dv_ds t M_max1 M_max2"

/* Take an output snapshot, asynchronously: on the context's stream (after
 * the work queued before it) V is copied to a staging buffer in device
 * memory and, with user_vars, the 23 user variables are computed next to
 * it (as WriteData calls ComputeUserVar before writing, Src/write_data.c:
 * 96-101: the line-force quantities use the dv/ds of the last stage's
 * input, so user_vars needs a step first); a copy stream then moves the
 * snapshot to pinned host memory while later steps run.  A second snapshot
 * waits on the device for the first one's transfer.  For a grid of a
 * batch, call pmx_wind_batch_sync first. */
int pmx_wind_snapshot(pmx_wind *w, int user_vars);
/* Wait for the last snapshot and copy it to out[6 (+ 23)][nx2][nx1] (V,
 * then the user variables); t, dt: g_time and g_dt at the snapshot; step:
 * the context's step counter. */
int pmx_wind_snapshot_wait(pmx_wind *w, double *out, double *t, double *dt, long *step);
/* The last snapshot in place, after pmx_wind_snapshot_wait: *data points at
 * the pinned host copy, [6 (+ 23)][nx2][nx1] as above, valid until the next
 * snapshot or pmx_wind_destroy (no copy; out may then be NULL above). */
int pmx_wind_snapshot_host(pmx_wind *w, const double **data);

/* WriteData for DBL_OUTPUT in single_file mode (Src/write_data.c:29-200,
 * :420-450, Src/bin_io.c:216-258): dir/data.NNNN.dbl holds the nvar arrays
 * data[nvar][nx3][nx2][nx1] (interior cells, host byte order) one after
 * another; dir/dbl.out gets the line "nfile t dt step single_file little
 * names..." after its first nfile lines (the file is rewritten from there,
 * "w" for nfile 0).  names: nvar names, e.g. "rho vx1 vx2 vx3 prs tr1" and
 * PMX_UV_NAMES split at spaces. */
int pmx_write_dbl(const char *dir, int nfile, const double *data, int nvar,
                  const char *const *names, int nx1, int nx2, int nx3, double t, double dt,
                  long step);
/* RestartDump (Src/restart.c:276-315): record `counter` (0, 1, ...) of
 * dir/restart.out, the reference's Restart structure (Src/structs.h:256-
 * 262: int nstep, int nfile[16], double t, double dt, 40 fill bytes; 128
 * bytes), nfile[16] being each output type's file counter.  The fill bytes
 * are written as zeros (the reference writes whatever its stack held). */
int pmx_write_restart(const char *dir, int counter, long nstep, const int *nfile, double t,
                      double dt);
/* RestartFromFile / RestartGet for .dbl (Src/restart.c:17-92, :195-274):
 * the byte order of output nrestart from its dbl.out line (*swap = 1 when
 * it is not the host's), then the first restart.out record whose nfile[0]
 * (the .dbl slot of the output table, Src/runtime_setup.c:244-249) equals
 * nrestart: its step number, the 16 file counters, t and dt. */
int pmx_read_restart(const char *dir, int nrestart, long *nstep, int *nfile, double *t,
                     double *dt, int *swap);
/* The first nvar arrays of dir/data.NNNN.dbl (nfile NNNN), data[nvar][nx3]
 * [nx2][nx1], byte-swapped when swap (Src/restart.c:104-128). */
int pmx_read_dbl(const char *dir, int nfile, int nvar, int nx1, int nx2, int nx3, int swap,
                 double *data);

/* ------------------------------------------------------------------ */
/* SIROCCO file interface (Src/LineDriven/line_connect.c:43-497): host-  */
/* side readers of the files a SIROCCO cycle writes.  Records are matched */
/* to cells of the grid of `grid` (nx1, nx2, x1, x2 with ghosts,          */
/* unit_length) by the reference's relative tolerances; every matching   */
/* cell takes the record.  Outputs are interior arrays [..][nx2][nx1].    */
/* A missing file is not an error (the reference prints a note and goes  */
/* on): *found is 0 and the outputs hold what the reference leaves.  A    */
/* malformed record is an error (the reference exits).  `dir` is the      */
/* directory of the files (NULL: the working directory).  *matched counts */
/* the records that matched a cell.  No device work.                     */
/* ------------------------------------------------------------------ */
/* Angle bins of directional_flux_r.dat: the third word of its second
 * line (line_connect.c:84-94).  *nangles = 0 when the file is missing. */
int pmx_sirocco_flux_angles(const char *dir, int *nangles);
/* read_sirocco_fluxes(), flux part (line_connect.c:68-165):
 * directional_flux_{r,theta,phi}.dat, records "i j k r[cm] theta f_0 ..
 * f_{nangles-1}", r matched within 1e-6 of x1 * unit_length and theta
 * within 1e-6 of x2 (relative).  Cells no record names keep the caller's
 * values (the reference's arrays are uninitialised there).  *found: bit a
 * set when the file of axis a (r, theta, phi) was read; matched[3]. */
int pmx_sirocco_read_fluxes(const char *dir, const pmx_wind_desc *grid, int nangles,
                            double *flux_r, double *flux_t, double *flux_p, int *found,
                            long *matched);
/* Fit points of M_UV_data.dat: the second word of its first line
 * (line_connect.c:193-198); 0 when the file is missing. */
int pmx_sirocco_fm_points(const char *dir, int *points);
/* The force multiplier fits (line_connect.c:183-257): t[points] from the
 * second line (after one word), records "i j r[cm] theta M_0 .. M_{points-1}"
 * matched as the fluxes.  Values as in the file (pmx_wind_create takes the
 * log10).  Unmatched cells keep the caller's values. */
int pmx_sirocco_read_force_multiplier(const char *dir, const pmx_wind_desc *grid, int points,
                                      double *t, double *m, int *found, long *matched);
/* read_sirocco_heatcool(flag = 1), py_heatcool.dat part
 * (line_connect.c:308-363): 18-word records "i j r[cm] theta vol t_e t_r
 * xi ne ..."; r and theta matched within 1e-5 (relative to the record's
 * values); xi floored at 1 and t_r at 1e3 K.  Without the file, every cell
 * gets the start-up values from rho (code units, [nx2][nx1]): xi = L_x /
 * n_H / r^2, t_r = T_x.  Unmatched cells keep the caller's values. */
int pmx_sirocco_read_heatcool(const char *dir, const pmx_wind_desc *grid, const double *rho,
                              double *xi, double *t_r, int *found, long *matched);
/* read_sirocco_heatcool(flag = 1), prefactors.dat part
 * (line_connect.c:365-497): 11-word records "i r[cm] j theta rho comp_h
 * comp_c xray_h brem_c line_c xi_ion" matched as py_heatcool; pre[6][nx2][nx1]
 * in that column order, 1.0 for every cell first (and everywhere without
 * the file).  *negative counts cell prefactors below 0 (the reference
 * prints each). */
int pmx_sirocco_read_prefactors(const char *dir, const pmx_wind_desc *grid, double *pre,
                                int *found, long *matched, long *negative);

/* ------------------------------------------------------------------ */
/* Batched sweep (no counterpart in the reference, which advances one    */
/* grid per process): many independent wind grids of uneven sizes, each  */
/* with its own descriptor (sizes, L_star, KRAD, ...), advanced together: */
/* one persistent kernel per RK stage over the (grid, tile) work list,   */
/* one batched cooling step, per-grid dt on the device, optional HIP      */
/* graph replay of the step loop.  All grids share the cooling mode       */
/* (PMX_COOL_NONE or PMX_COOL_PLAIN) and the step parity.                 */
/* ------------------------------------------------------------------ */
typedef struct pmx_wind_batch pmx_wind_batch;

int pmx_wind_batch_create(const pmx_wind_desc *descs, int ngrids, pmx_wind_batch **out);
int pmx_wind_batch_destroy(pmx_wind_batch *b);
/* The batch's context of grid g, for pmx_wind_set_prim / get_prim /
 * get_cons / set_heatcool / set_time / get_time (owned by the batch; do
 * not destroy it, do not run it on its own). */
int pmx_wind_batch_grid(pmx_wind_batch *b, int g, pmx_wind **w);
/* g_stepNumber of the next step (Strang order and the dt cadence). */
int pmx_wind_batch_set_step(pmx_wind_batch *b, long step);
/* Surrogate-cooled SplitSource for every grid (see pmx_wind_set_surrogate). */
int pmx_wind_batch_set_surrogate(pmx_wind_batch *b, pmx_mlp *m, int guess);
/* Times the stage kernel for each tile shape (64x2, 32x4, 16x8, 128x1,
 * 64x4, 32x8 cells) and persistent workgroup count (one per tile, or 1, 2,
 * 4, 8 per compute unit) and keeps the fastest; the state is unchanged. */
int pmx_wind_batch_autotune(pmx_wind_batch *b, int *shape, int *blocks, double *ms_per_stage);
/* Sets the tile shape (index as above) and workgroup count (<= 0: one per tile). */
int pmx_wind_batch_set_shape(pmx_wind_batch *b, int shape, int blocks);
/* nsteps main-loop iterations of every grid; use_graph != 0 replays a
 * captured (even, odd) pair of steps.  Asynchronous. */
int pmx_wind_batch_run(pmx_wind_batch *b, int nsteps, int use_graph);
int pmx_wind_batch_sync(pmx_wind_batch *b);
/* Per-kernel timing (stage and cooling launches) while enabled; graph
 * replay is not used while timing is enabled. */
int pmx_wind_batch_timing(pmx_wind_batch *b, int enable, int reset, double *stage_ms,
                          int *stage_launches, double *cool_ms, int *cool_launches);
/* Interior cells of all grids. */
int pmx_wind_batch_cells(pmx_wind_batch *b, long long *cells);

/* ------------------------------------------------------------------ */
/* MLP cooling surrogate (no counterpart in the reference).              */
/*                                                                      */
/* 4 -> 64 -> 64 -> 64 -> 1, SiLU, BF16 weights and activations on the   */
/* matrix cores, FP32 accumulation, FP32 biases.  Inputs: log10 rho      */
/* [g/cm^3], log10 T [K], log10 xi, log10(dt / t_cool) with t_cool =     */
/* E / |dE/dt| of the cell's state; output: log10(T_final / T_0).        */
/* Parameters are a float32 vector of 8705 entries: W1[64][4] b1[64]     */
/* W2[64][64] b2[64] W3[64][64] b3[64] w4[64] b4.                        */
/* ------------------------------------------------------------------ */
int pmx_mlp_create(unsigned long long seed, pmx_mlp **out);
int pmx_mlp_destroy(pmx_mlp *m);
int pmx_mlp_set_params(pmx_mlp *m, const float *P);   /* also resets Adam's moments */
int pmx_mlp_get_params(pmx_mlp *m, float *P);
/* n training samples drawn log-uniformly over the hull rho 1e-16..1e-10
 * g/cm^3, T 1e4..1e8 K, xi 1..1e6, dt 1e-2..1e2 t_cool, labelled by the
 * plain cooling integrator (the device code of pmx_wind_split_source with
 * PMX_COOL_PLAIN) with the physics scalars of `desc` (its arrays are not
 * read).  The dataset stays on the device. */
int pmx_mlp_sample(pmx_mlp *m, const pmx_wind_desc *desc, long long n, unsigned long long seed);
int pmx_mlp_set_dataset(pmx_mlp *m, const float *X, const float *Y, long long n);
/* X [n][4] raw features, Y [n] labels, nev [n] Radiat() calls the plain
 * integrator made per sample, aux [n][4] the sample's exact inputs (rho and
 * p in code units, xi, dt in code units); any pointer may be NULL. */
int pmx_mlp_get_dataset(pmx_mlp *m, float *X, float *Y, int *nev, double *aux, long long n);
/* Mean loss and gradient (8705 floats) of L = sum (y - Y/2)^2 / count over
 * the first `count` samples (y: network output, Y: label). */
int pmx_mlp_loss_grad(pmx_mlp *m, long long count, float *loss, float *grad);
/* `steps` Adam steps (batch a multiple of 64, cosine learning rate lr0 ->
 * lr1); loss_hist (optional) receives each step's mean loss. */
int pmx_mlp_train(pmx_mlp *m, int steps, int batch, float lr0, float lr1, float *loss_hist);
/* One cooling step for n independent cells (no hydro), each with its own
 * density rho and pressure p (code units), ionisation parameter xi and dt
 * (code units), tracer 1, T_r = T_x: mode 0 the plain integrator, 1 the
 * surrogate (network + refinement, plain fallback), 2 the surrogate's
 * refinement without the network.  p_out receives the new pressures;
 * stats (optional, 4 entries): cells given to the network, accepted,
 * sent to the plain integrator, Radiat() calls; *ms (optional) the
 * device time of the cooling kernels. */
int pmx_mlp_cool_cells(pmx_mlp *m, const pmx_wind_desc *desc, long long n, const double *rho,
                       const double *p, const double *xi, const double *dt, double *p_out,
                       int mode, long long *stats, double *ms);
/* Y[n] = predicted log10(T_final / T_0) for raw features X[n][4]. */
int pmx_mlp_predict(pmx_mlp *m, const float *X, long long n, float *Y);

/* Surrogate-cooled SplitSource (no counterpart in the reference) for a
 * wind context with PMX_COOL_PLAIN: cells the reference integrator treats
 * as stiff (|dE/dt| > 0.5 E / dt) and that lie inside the training hull get
 * the network's guess, refined by Newton's method on the cooling ODE's time
 * integral; cells outside the hull or whose refinement does not converge
 * take the plain integrator, as do all non-stiff cells.  m == NULL switches
 * back to the plain integrator.  The context keeps a pointer to m. */
int pmx_wind_set_surrogate(pmx_wind *w, pmx_mlp *m, int guess);
/* guess: 0 = the network's prediction starts the refinement; 1 = ablation
 * without the network (start from the explicit step limited to one
 * e-fold), to measure what the network contributes. */
/* Running totals since the last reset: stiff cells handed to the network,
 * accepted refinements, cells sent to the plain integrator, and Radiat()
 * calls of all cooling steps (with or without the surrogate). */
int pmx_wind_surrogate_stats(pmx_wind *w, long long *listed, long long *accepted,
                             long long *fallback, long long *radiat_calls, int reset);

/* Measurement helper (no counterpart in the reference): times `reps`
 * launches of a plain streaming copy of nbytes (multiple of 16) between two
 * fresh device buffers; reports the mean milliseconds per copy.  Used to
 * calibrate HBM byte counters and as the achievable-bandwidth reference. */
int pmx_copy_probe(long long nbytes, int reps, double *ms_per_copy);
/* Measurement helper (no counterpart in the reference): the same copy by
 * the HIP runtime's own device-to-device hipMemcpyAsync, as a second
 * reference for the achievable bandwidth. */
int pmx_copy_probe_runtime(long long nbytes, int reps, double *ms_per_copy);

/* Arithmetic check (no counterpart in the reference): the line force's
 * divisions by per-cell and per-run constants (Src/LineDriven/
 * line_connect.c:753 and :899-900) are evaluated from a reciprocal with
 * FMA corrections where the run's inputs keep the operands in range; this
 * runs that evaluation and the plain division of each of the n values x
 * (0, or |x| and |x / c| within 2^-900..2^900) by c > 0 on the device and
 * reports how many of the quotients differ in any bit (0 expected). */
int pmx_check_div_rc(const double *x, long long n, double c, long long *mismatches);

/* Arithmetic check (no counterpart in the reference): the line force's
 * force-multiplier power M = krad pow(t, alpha), t = num / dvds
 * (Src/LineDriven/line_connect.c:880-881), as the device evaluates it:
 * log2 num per LineForce call and log2 dvds per bin, shared, in
 * double-double, and 2^(alpha log2 t) with the quotient's rounding error
 * folded in.  out[i] = krad pow(num[i] / dvds[i], alpha) for dvds[i] > 0,
 * else 0 (host arrays of n values). */
int pmx_check_force_power(const double *num, const double *dvds, long long n, double krad,
                          double alpha, double *out);

#ifdef __cplusplus
}
#endif
#endif
