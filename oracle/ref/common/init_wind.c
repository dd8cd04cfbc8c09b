/* Problem file for the line-driven disk-wind oracle builds of the
 * unmodified PLUTO reference (test infrastructure only, never shipped in
 * the product).  Same I/O conventions as init_file_ic.c, plus the point-
 * mass gravity g_r = -G M / r^2 of the wind problem.
 *
 * - Init() sets a placeholder state; InitDomain() then overwrites the
 *   interior of Vc with the binary initial condition named by the
 *   environment variable PMX_IC (float64, [nv][k][j][i], interior only),
 *   so the reference and the GPU library start from identical bits.
 * - Analysis() (called every step with "analysis -1 1" in pluto.ini)
 *   overwrites the file named by PMX_DUMP with the current state:
 *       int32  magic(0x504D5844), step, nx1, nx2, nx3, nvar
 *       double time, dt, wall_seconds_since_first_analysis_call
 *       double V[nv][k][j][i]   (interior primitive variables)
 *       double U[nv][k][j][i]   (interior conservative variables, Uc)
 *   and appends (step, time, dt) to PMX_DUMP".dt".
 *   With PMX_TIME_ONLY set, nothing is written and only the wall time
 *   between the first and the last Analysis() call is printed to stdout.
 * - A SIROCCO coupling cycle restarts PLUTO with the files SIROCCO wrote.
 *   Two variables give a fresh start the state such a restart has: PMX_T0
 *   sets g_time, as RestartFromFile does (Src/restart.c:259), and with
 *   PMX_SIROCCO_RESTART set read_sirocco_heatcool() is called with
 *   restart_flag = 1 (Src/main.c:197), so the reference's own reader loads
 *   py_heatcool.dat and prefactors.dat.  The call reaches it through the
 *   linker's --wrap=read_sirocco_heatcool; the reader is compiled as it lies.
 */
#define _POSIX_C_SOURCE 199309L
#include <time.h>
#include "pluto.h"

static double wall_now(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1.e-9*(double)ts.tv_nsec;
}

void Init (double *v, double x1, double x2, double x3)
{
  int nv;
#ifdef GAMMA_EOS
  g_gamma = g_inputParam[GAMMA_EOS];
#else
  g_gamma = g_inputParam[GAMMA];      /* the cv_idl parameter set */
#endif
  for (nv = 0; nv < NVAR; nv++) v[nv] = 0.0;
  v[RHO] = 1.0;
  v[PRS] = 1.0;
  v[TRC] = 1.0;
  dvds_setup_flag = 0;
}

void InitDomain (Data *d, Grid *grid)
{
  int i, j, k, nv;
  const char *path = getenv("PMX_IC");
  FILE *fp;
  size_t n;
  double *buf;
  long nint = (long)NX1*(long)NX2*(long)NX3;

  if (path == NULL) {
    print ("! InitDomain(): PMX_IC not set\n");
    QUIT_PLUTO(1);
  }
  fp = fopen(path, "rb");
  if (fp == NULL) {
    print ("! InitDomain(): cannot open %s\n", path);
    QUIT_PLUTO(1);
  }
  buf = (double *)malloc(sizeof(double)*nint);
  for (nv = 0; nv < NVAR; nv++) {
    n = fread(buf, sizeof(double), nint, fp);
    if ((long)n != nint) {
      print ("! InitDomain(): short read in %s (var %d)\n", path, nv);
      QUIT_PLUTO(1);
    }
    DOM_LOOP(k,j,i) {
      d->Vc[nv][k][j][i] = buf[((long)(k-KBEG)*NX2 + (j-JBEG))*NX1 + (i-IBEG)];
    }
  }
  free(buf);
  fclose(fp);
  if (getenv("PMX_T0") != NULL) g_time = atof(getenv("PMX_T0"));
}

void __real_read_sirocco_heatcool(Data *d, Grid *grid, int flag);

void __wrap_read_sirocco_heatcool(Data *d, Grid *grid, int flag)
{
  __real_read_sirocco_heatcool(d, grid, getenv("PMX_SIROCCO_RESTART") != NULL ? 1 : flag);
}

void Analysis (const Data *d, Grid *grid)
{
  static double t0 = -1.0;
  static int ncall = 0;
  const char *path = getenv("PMX_DUMP");
  double now = wall_now();
  int i, j, k, nv;
  FILE *fp;
  char dtpath[1024];

  if (t0 < 0.0) t0 = now;
  ncall++;
  if (getenv("PMX_TIME_ONLY") != NULL) {
    printf ("PMX_TIMING step=%d wall=%.9f\n", g_stepNumber, now - t0);
    return;
  }
  if (path == NULL) return;

  fp = fopen(path, "wb");
  {
    int hdr[6] = {0x504D5844, g_stepNumber, NX1, NX2, NX3, NVAR};
    double dh[3];
    dh[0] = g_time; dh[1] = g_dt; dh[2] = now - t0;
    fwrite(hdr, sizeof(int), 6, fp);
    fwrite(dh, sizeof(double), 3, fp);
  }
  for (nv = 0; nv < NVAR; nv++) {
    DOM_LOOP(k,j,i) fwrite(&d->Vc[nv][k][j][i], sizeof(double), 1, fp);
  }
  for (nv = 0; nv < NVAR; nv++) {
    DOM_LOOP(k,j,i) fwrite(&d->Uc[k][j][i][nv], sizeof(double), 1, fp);
  }
  fclose(fp);

  snprintf(dtpath, sizeof(dtpath), "%s.dt", path);
  fp = fopen(dtpath, ncall == 1 ? "w" : "a");
  fprintf(fp, "%d %.17e %.17e\n", g_stepNumber, g_time, g_dt);
  fclose(fp);
}

#ifndef PMX_CVIDL_BOUNDARY
void UserDefBoundary (const Data *d, RBox *box, int side, Grid *grid)
{
}
#endif

void BodyForceVector(double *v, double *g, double x1, double x2, double x3)
{
  double gm_code = CONST_G*g_inputParam[CENT_MASS]
                   /(UNIT_LENGTH*UNIT_VELOCITY*UNIT_VELOCITY);
  g[IDIR] = -1.0*gm_code/(x1*x1);
  g[JDIR] = 0.0;
  g[KDIR] = 0.0;
}

double BodyForcePotential(double x1, double x2, double x3)
{
  return 0.0;
}
