/* Oracle internals shared between the Cartesian hydro restatement
 * (hydro.c) and the line-driven wind restatement (wind.c).  Test
 * infrastructure only. */
#ifndef PMX_ORACLE_HD_INTERNAL_H
#define PMX_ORACLE_HD_INTERNAL_H

/* physical flux of the HD equations, normal momentum flux without pressure
 * (Src/HD/fluxes.c:23-51) */
void orc_phys_flux(int n, const double *v, const double *u, double *f);
/* HLLC interface solver (Src/HD/hllc.c:28-180) */
void orc_hllc(double gamma, int n, const double *vL, const double *uL,
              const double *vR, const double *uR, double *flux,
              double *press, double *cmax, double *mach);
/* Roe interface solver (Src/HD/roe.c:42-346) */
void orc_roe(double gamma, int ndim, int n, const double *ql,
             const double *uL, const double *qr, const double *uR,
             double *flux, double *press, double *cmax, double *mach);

/* records why a run ended before its step count (orc_last_quit) */
void orc_set_quit(int why);

#endif
