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
/* boundary types (pluto.ini [Boundary]) */
#define PMX_BC_PERIODIC   0
#define PMX_BC_OUTFLOW    1
#define PMX_BC_REFLECTIVE 2

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
} pmx_hydro_desc;

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

/* One AdvanceStep(d, Dts, grid) of the reference (Src/Time_Stepping/
 * rk_step.c:29, prototype Src/prototypes.h:5) with time step dt:
 * boundaries, all RK stages, cons->prim.  Returns the stage-1 inverse
 * hyperbolic time step Dts->invDt_hyp (update_stage.c:389-394).
 * Synchronises the stream. */
int pmx_hydro_advance_step(pmx_hydro *h, double dt, double *inv_dt_hyp);

/* nsteps iterations of the reference's main loop body for this path
 * (Integrate + time update + NextTimeStep, Src/main.c:215-337,
 * :383-485, :521-697) entirely on the device: dt and t stay in device
 * memory and no host synchronisation happens between steps.
 * Asynchronous; pmx_hydro_sync() waits. */
int pmx_hydro_run(pmx_hydro *h, int nsteps);
int pmx_hydro_sync(pmx_hydro *h);

/* Same as pmx_hydro_run, but replays a captured HIP graph of one step
 * (built on first use). */
int pmx_hydro_run_graph(pmx_hydro *h, int nsteps);

/* Per-kernel timing with HIP events recorded on the context's stream
 * around every stage-kernel launch while enabled.  Reports the summed
 * duration (ms) and the number of stage launches since the last reset. */
int pmx_hydro_stage_timing(pmx_hydro *h, int enable, int reset, double *total_ms,
                           int *launches);

/* Device pointers of the current primitive state V[nv][k][j][i] and of
 * the conservative U (for zero-copy consumers such as torch). */
int pmx_hydro_device_state(pmx_hydro *h, double **V, double **U);

/* Measurement helper (no counterpart in the reference): times `reps`
 * launches of a plain streaming copy of nbytes (multiple of 16) between two
 * fresh device buffers; reports the mean milliseconds per copy.  Used to
 * calibrate HBM byte counters and as the achievable-bandwidth reference. */
int pmx_copy_probe(long long nbytes, int reps, double *ms_per_copy);

#ifdef __cplusplus
}
#endif
#endif
