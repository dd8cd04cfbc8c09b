// libplutomx: line-driven disk wind on a 2-D spherical grid (see the wind
// section of include/plutomx.h).
//
// Kernels
//   wind_stage_kernel<S>  one RK2 stage: boundaries (by index mapping, no
//                         ghost storage), VGradCalc, PLM, HLLC + tracer,
//                         spherical RHS, geometric / gravity / line-force
//                         sources, RK combination, cons->prim and the
//                         stage-1 CFL reduction.  One thread per cell.
//   wind_cool_kernel<M>   SplitSource cooling (BLONDIN or the plain
//                         integrator) and the dt_cool reduction.
//   wind_end_kernel       time update and NextTimeStep (every second step
//                         with cooling, as the reference).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/plutomx.h"
#include "pmx_common.h"
#include "pmx_wind_math.h"
#include "pmx_wind_host.h"

namespace pmx {
namespace wind {

constexpr int NG = 2;          // ghost cells of the reference grid (PLM)
constexpr int BX = 64, BY = 2; // stage-kernel workgroup
constexpr int NLF = 5;         // line-force tables per (angle, cell)
constexpr int FM_POINTS_MAX = 4096;   // fit points of a tabulated force multiplier

// Waves per SIMD that the register allocation of the batch stage kernels
// and the one-grid feature kernels targets.  The plain
// stage kernel fits 2 waves by itself (200-226 VGPRs); kernels built with
// characteristic limiting, the entropy switch, the tabulated multiplier or
// Roe would take more than 256 registers and fall to 1 wave, so they are
// capped and spill.  (With the line force still inside it, the batch
// kernel took 276 registers and 1 wave; capped at 2 it ran 1.5x faster.)
#ifndef PMX_WIND_WAVES
#define PMX_WIND_WAVES 2
#endif
// Per (angle, cell): the offset point's polar bilinear weight (-999 for a
// bin without flux, which the reference skips), sin / cos of its polar
// angle, and the two flux components.  The radial weight and ds are
// recomputed per bin from the cell's geometry with the same IEEE
// operations as the host set-up (bit-identical, 2 fewer tables to read).
enum { LF_F2, LF_SIN, LF_COS, LF_FR, LF_FT };

// Geometry and line-force tables on the device.  Radial and polar 1-D
// arrays include the 2 ghost cells (index = interior index + 2).
struct Tables {
  const double *x1, *dx1, *x1r, *rt, *inv_dx1, *dVr, *pr;   // pr: [6][nr+4]
  const double *x2, *dx2, *s, *sp, *dmu, *inv_dx2, *ct, *sin2, *cos2, *pt;
  const double *lf;            // [NLF][na][nt][nr]
  const double *f1c, *f2c;     // [nr], [nt]: centre bilinear weights
  const double *sta, *cta;     // [na]: sin / cos of the bin angle
  double x1l0, sin_x2l0;       // x1 and |sin x2| of the lower edge of the first ghost zone
  // the cv_idl midplane row (k.midplane): per interior column, the disk's
  // density, Keplerian v_phi and pressure, [3][nr]
  const double* mid;
  // tabulated force multiplier (k.fm_points > 0): log10 t [fm_points] and
  // log10 M [fm_points][nt][nr] of the fits, as read_sirocco_fluxes keeps them
  const double *fm_lt, *fm_lm;
};

struct StageArgs {
  const double* Vin;   // stage input (stencil)
  const double* Vn;    // V^n (stage 2: U^n = P2C(V^n))
  double* U;
  double* Vout;
  int nr, nt, na;
  int bc[4];
  long long sv;
  Const k;
  Tables T;
  Scalars* s;
  // SHOCK_FLATTENING MULTID: 1 per shocked zone over the interior and two
  // ghost layers, [nt+4][nr+4]; null without flattening
  const unsigned char* shock;
  double* Ue;          // ENTROPY_SWITCH: Uc[ENTR] = rho s per zone; null without it
  // the stage's line-force accelerations per zone, g_r [sv] then g_theta
  // [sv]: written by the line-force kernel, read by the stage kernel
  double* glf;
  int loop;            // a main-loop step: nothing to do once s->halt is set
};

// The wind's userdef sides: r-beg maps like outflow (plus v_r <= 0), theta-
// beg like reflective (plus rho, p from the first row).  Kept out of the
// shared map_bc: one more compare there costs the Cartesian stage kernel,
// which sits at its register limit, twice its spills.
enum { BC_WIND_INNER = PMX_BC_WIND_INNER, BC_WIND_POLE = PMX_BC_WIND_POLE };

__device__ __forceinline__ int map_bc_wind(int g, int n, int bcl, int bcr, bool& flip) {
  return map_bc(g, n, bcl == BC_WIND_INNER ? BC_OUTFLOW : bcl, bcr, flip);
}

// The midplane row of the cv_idl side-0 pass (cv_idl/init.c:276-311) on
// zone (last polar row, column im), after its floors: the disk density
// with the polar momentum kept, v_r = 0, the Keplerian v_phi, the disk
// pressure, tracer 1 (the per-column values come from the host set-up).
__device__ __forceinline__ void midplane_cell(const StageArgs& a, int im, double* v) {
  const double rho_mid = a.T.mid[im];
  v[VX2] = (v[RHO] * v[VX2]) / rho_mid;
  v[RHO] = rho_mid;
  v[VX1] = 0.0;
  v[VX3] = a.T.mid[a.nr + im];
  v[PRS] = a.T.mid[2 * a.nr + im];
  v[TRC] = 1.0;
}

// The state of cell (jj, ii) as the reference's Boundary() leaves it at
// the start of a stage (Src/boundary.c:56-480): ghost cells are never
// stored; an out-of-domain index maps to the interior cell the boundary
// rule copies from.  With the internal boundary on, that cell's floors
// (floor_cell) apply first, as UserDefBoundary(side 0) runs before the
// sides are filled; then the side rules in the reference's order (x1
// sides, then x2 sides over the full x1 range, so corners see both):
// reflective flips the normal velocity; WIND_INNER caps v_r at 0; WIND_POLE
// reflects the velocities and takes rho and p from the first interior row.
// Returns 1 when the floors changed an interior cell (used for the cell's
// own Uc).
__device__ __forceinline__ int ld6(const StageArgs& a, const double* __restrict__ V, int jj,
                                   int ii, double* v) {
  bool fr = false, ft = false;
  int im = map_bc_wind(ii, a.nr, a.bc[0], a.bc[1], fr);
  int jm = map_bc_wind(jj, a.nt, a.bc[2], a.bc[3], ft);
  const double* p = V + (long long)jm * a.nr + im;
#pragma unroll
  for (int nv = 0; nv < NVW; nv++) v[nv] = p[nv * a.sv];
  int changed = 0;
  if (a.k.dfloor > 0.0) changed = floor_cell(a.k, v);
  if (a.k.midplane && jm == a.nt - 1) midplane_cell(a, im, v);
  if (fr) v[VX1] = -v[VX1];
  if (ii < 0 && a.bc[0] == BC_WIND_INNER) v[VX1] = dmin(v[VX1], 0.0);
  if (ft) v[VX2] = -v[VX2];
  if (jj < 0 && a.bc[2] == BC_WIND_POLE) {
    const double* q = V + im;          // row 0 of the same column
    double w[NVW];
#pragma unroll
    for (int nv = 0; nv < NVW; nv++) w[nv] = q[nv * a.sv];
    if (a.k.dfloor > 0.0) floor_cell(a.k, w);
    v[RHO] = w[RHO];
    v[PRS] = w[PRS];
  }
  return changed;
}

__device__ __forceinline__ void ld6(const StageArgs& a, int jj, int ii, double* v) {
  ld6(a, a.Vin, jj, ii, v);
}

__device__ __forceinline__ void ld_vel(const StageArgs& a, int jj, int ii, double& v1,
                                       double& v2) {
  bool fr = false, ft = false;
  int im = map_bc_wind(ii, a.nr, a.bc[0], a.bc[1], fr);
  int jm = map_bc_wind(jj, a.nt, a.bc[2], a.bc[3], ft);
  const double* p = a.Vin + (long long)jm * a.nr + im;
  v1 = p[VX1 * a.sv];
  v2 = p[VX2 * a.sv];
  if (a.k.dfloor > 0.0) {
    // floor_cell's velocity scaling, which needs only rho
    double rho = p[RHO * a.sv];
    if (rho < a.k.dfloor) {
      if (rho < 0.0) rho = a.k.dfloor;
      const double dfact = rho / a.k.dfloor;
      v1 = dfact * v1;
      v2 = dfact * v2;
      rho = a.k.dfloor;
    }
    if (a.k.midplane && jm == a.nt - 1) {   // midplane_cell's velocities
      v2 = (rho * v2) / a.T.mid[im];
      v1 = 0.0;
    }
  }
  if (fr) v1 = -v1;
  if (ii < 0 && a.bc[0] == BC_WIND_INNER) v1 = dmin(v1, 0.0);
  if (ft) v2 = -v2;
}

// VGradCalc for one cell (line_connect.c:628-738) from the stage input:
// the velocity averages at the four cell corners (:638-660) and the
// centre velocity by bilinear interpolation (:746-777), in the (x, z)
// frame; vgrad_dvds then gives one bin's |dv/ds| between the centre and
// that bin's offset point, or -999 for a bin without flux.
struct VGrad {
  double v11_0, v11_1, v12_0, v12_1, v22_0, v22_1, v21_0, v21_1;
  double vx1, vz1;
  // the cell's offset geometry (line_connect.c:562-607): centre (x, z),
  // the offset length maxds and the radial corner positions
  double x, z, maxds, x11_0, dx0, rdx0;   // rdx0 = 1 / dx0 for div_rc
};

__device__ __forceinline__ void vgrad_cell(const StageArgs& a, int j, int i, VGrad& g) {
  const Const& k = a.k;
  const Tables& T = a.T;
  double P1[3][3], P2[3][3];
#pragma unroll
  for (int dj = 0; dj < 3; dj++)
#pragma unroll
    for (int di = 0; di < 3; di++) ld_vel(a, j + dj - 1, i + di - 1, P1[dj][di], P2[dj][di]);
  // P[dj][di] = V(j+dj-1, i+di-1)
  g.v11_0 = (P1[0][0] + P1[0][1] + P1[1][0] + P1[1][1]) / 4.0;
  g.v11_1 = (P2[0][0] + P2[0][1] + P2[1][0] + P2[1][1]) / 4.0;
  g.v12_0 = (P1[2][0] + P1[1][0] + P1[2][1] + P1[1][1]) / 4.0;
  g.v12_1 = (P2[2][0] + P2[1][0] + P2[2][1] + P2[1][1]) / 4.0;
  g.v22_0 = (P1[2][1] + P1[2][2] + P1[1][2] + P1[1][1]) / 4.0;
  g.v22_1 = (P2[2][1] + P2[2][2] + P2[1][2] + P2[1][1]) / 4.0;
  g.v21_0 = (P1[1][2] + P1[0][2] + P1[0][1] + P1[1][1]) / 4.0;
  g.v21_1 = (P2[1][2] + P2[0][2] + P2[0][1] + P2[1][1]) / 4.0;
  double f1 = T.f1c[i], f2 = T.f2c[j], t1, t2;
  t1 = (1.0 - f1) * g.v11_0 + f1 * g.v21_0;
  t2 = (1.0 - f1) * g.v12_0 + f1 * g.v22_0;
  double ans0 = (1.0 - f2) * t1 + f2 * t2;
  t1 = (1.0 - f1) * g.v11_1 + f1 * g.v21_1;
  t2 = (1.0 - f1) * g.v12_1 + f1 * g.v22_1;
  double ans1 = (1.0 - f2) * t1 + f2 * t2;
  const double sj = T.sin2[j + NG], cj = T.cos2[j + NG];
  g.vx1 = (ans0 * k.UV * sj + ans1 * k.UV * cj);
  g.vz1 = (ans0 * k.UV * cj - ans1 * k.UV * sj);
  // the set-up's expressions (as in pmx_wind_create, line_connect.c:562-607)
  const int I = i + NG, J = j + NG;
  const double UL = k.UL;
  const double x11_0 = (T.x1[I - 1] + T.x1[I]) / 2.0 * UL, x11_1 = (T.x2[J - 1] + T.x2[J]) / 2.0;
  const double x22_0 = (T.x1[I + 1] + T.x1[I]) / 2.0 * UL, x22_1 = (T.x2[J + 1] + T.x2[J]) / 2.0;
  double maxds = fabs(x22_0 - x11_0);
  if (maxds > (T.x1[I] * UL * fabs(x22_1 - x11_1))) maxds = T.x1[I] * UL * fabs(x22_1 - x11_1);
  maxds /= 2.0;
  g.maxds = maxds;
  g.x11_0 = x11_0;
  g.dx0 = x22_0 - x11_0;
  g.rdx0 = 1.0 / g.dx0;
  g.x = T.x1[I] * sj * UL;
  g.z = T.x1[I] * cj * UL;
}

// g2, so, co: the cell's LF_F2, LF_SIN and LF_COS entries of bin `ang`.
template <bool FAST = false>
__device__ __forceinline__ double vgrad_dvds(const StageArgs& a, const VGrad& g, double g2,
                                             double so, double co, int ang) {
  const Const& k = a.k;
  if (g2 == -999) return -999;
  const double ddx1 = g.maxds * a.T.sta[ang], ddx2 = g.maxds * a.T.cta[ang];
  const double ds = sqrt(ddx1 * ddx1 + ddx2 * ddx2);
  const double roff = sqrt((g.x + ddx1) * (g.x + ddx1) + (g.z + ddx2) * (g.z + ddx2));
  double g1 = qdiv<FAST>(roff - g.x11_0, g.dx0, g.rdx0), t1, t2;
  t1 = (1.0 - g1) * g.v11_0 + g1 * g.v21_0;
  t2 = (1.0 - g1) * g.v12_0 + g1 * g.v22_0;
  double b0 = (1.0 - g2) * t1 + g2 * t2;
  t1 = (1.0 - g1) * g.v11_1 + g1 * g.v21_1;
  t2 = (1.0 - g1) * g.v12_1 + g1 * g.v22_1;
  double b1 = (1.0 - g2) * t1 + g2 * t2;
  double vx2 = (b0 * k.UV * so + b1 * k.UV * co);
  double vz2 = (b0 * k.UV * co - b1 * k.UV * so);
  double st = a.T.sta[ang], ctt = a.T.cta[ang];
  double u1 = st * g.vx1 + ctt * g.vz1;
  double u2 = st * vx2 + ctt * vz2;
  return fabs((u2 - u1) / ds);
}

// q: the cell's line-force table entries of bin `ang`; sa: table stride.
template <bool FAST = false>
__device__ __forceinline__ double vgrad_dvds(const StageArgs& a, const VGrad& g,
                                             const double* q, long long sa, int ang) {
  return vgrad_dvds<FAST>(a, g, q[LF_F2 * sa], q[LF_SIN * sa], q[LF_COS * sa], ang);
}

__device__ __forceinline__ void ldc(const double* tab, int n, int idx, double* c) {
#pragma unroll
  for (int q = 0; q < 6; q++) c[q] = tab[q * (n + 2 * NG) + idx];
}

// HLL at an interface next to a shocked zone: HLLC's speeds and supersonic
// cases, then the HLL flux (Src/HD/hllc.c:57-104 with FLAG_HLL).
template <int DIR>
__device__ __forceinline__ void hll_face(const double* vL, const double* vR, double gamma,
                                         double gmm1, double* flux, double& press,
                                         double& cmax) {
  const int n = VX1 + DIR;
  double uL[NV], uR[NV], fL[NV], fR[NV];
  prim_to_cons(vL, uL, gmm1);
  prim_to_cons(vR, uR, gmm1);
  double a2L = gamma * vL[PRS] / vL[RHO];
  double a2R = gamma * vR[PRS] / vR[RHO];
  double aL = sqrt(a2L), aR = sqrt(a2R);
  double SL = dmin(vL[n] - aL, vR[n] - aR);
  double SR = dmax(vL[n] + aL, vR[n] + aR);
  cmax = dmax(fabs(SL), fabs(SR));
  phys_flux<DIR>(vL, uL, fL);
  phys_flux<DIR>(vR, uR, fR);
  if (SL > 0.0) {
#pragma unroll
    for (int q = 0; q < NV; q++) flux[q] = fL[q];
    press = vL[PRS];
  } else if (SR < 0.0) {
#pragma unroll
    for (int q = 0; q < NV; q++) flux[q] = fR[q];
    press = vR[PRS];
  } else {
    const double scrh = 1.0 / (SR - SL);
#pragma unroll
    for (int q = 0; q < NV; q++) {
      flux[q] = SL * SR * (uR[q] - uL[q]) + SR * fL[q] - SL * fR[q];
      flux[q] *= scrh;
    }
    press = (SR * vL[PRS] - SL * vR[PRS]) * scrh;
  }
}

// Roe's HLL switch next to a shocked zone (Src/HD/roe.c:105-131): Davis
// speeds, bmin = min(0, SL), bmax = max(0, SR) and the HLL formula for every
// interface, without HLLC's supersonic shortcut.
template <int DIR>
__device__ __forceinline__ void roe_hll_face(const double* vL, const double* vR, double gamma,
                                             double gmm1, double* flux, double& press,
                                             double& cmax) {
  const int n = VX1 + DIR;
  double uL[NV], uR[NV], fL[NV], fR[NV];
  prim_to_cons(vL, uL, gmm1);
  prim_to_cons(vR, uR, gmm1);
  phys_flux<DIR>(vL, uL, fL);
  phys_flux<DIR>(vR, uR, fR);
  double aL = sqrt(gamma * vL[PRS] / vL[RHO]), aR = sqrt(gamma * vR[PRS] / vR[RHO]);
  double bmin = dmin(vL[n] - aL, vR[n] - aR);
  double bmax = dmax(vL[n] + aL, vR[n] + aR);
  cmax = dmax(fabs(bmin), fabs(bmax));
  bmin = dmin(0.0, bmin);
  bmax = dmax(0.0, bmax);
  const double scrh = 1.0 / (bmax - bmin);
#pragma unroll
  for (int q = 0; q < NV; q++) {
    flux[q] = bmin * bmax * (uR[q] - uL[q]) + bmax * fL[q] - bmin * fR[q];
    flux[q] *= scrh;
  }
  press = (bmax * vL[PRS] - bmin * vR[PRS]) * scrh;
}

// HLLC (or HLL next to a shocked zone), or with ROE the Roe solver, at one
// interface plus the tracer's upwind flux (Src/adv_flux.c:42-52)
template <int DIR, bool ROE = false>
__device__ __forceinline__ void face(const double* vL, const double* vR, const Const& k,
                                     double* F, double& p, double& cmax, bool hll = false) {
  if (ROE && k.roe) {   // a batch kernel built with Roe also runs HLLC / HLL grids
    if (hll) roe_hll_face<DIR>(vL, vR, k.gamma, k.gmm1, F, p, cmax);
    else roe<DIR, 2>(vL, vR, k.gamma, k.gmm1, F, p, cmax);
  } else if (hll || k.hll) {
    // HLL_Solver (Src/HD/hll.c:29-103) is HLLC's FLAG_HLL branch
    hll_face<DIR>(vL, vR, k.gamma, k.gmm1, F, p, cmax);
  } else {
    hllc<DIR>(vL, vR, k.gamma, k.gmm1, F, p, cmax);
  }
  F[TRC] = F[RHO] * (F[RHO] > 0.0 ? vL[TRC] : vR[TRC]);
}

// Shock flags of zone (j, i), j in [-2, nt + 1], i in [-2, nr + 1].
__device__ __forceinline__ bool shocked(const StageArgs& a, int j, int i) {
  return a.shock[(long long)(j + 2) * (a.nr + 4) + (i + 2)] != 0;
}
// FLAG_MINMOD of zone (j, i), j in [-1, nt], i in [-1, nr]: the zone or a
// face neighbour is shocked (flag_shock.c:214-225).
__device__ __forceinline__ bool minmod_zone(const StageArgs& a, int j, int i) {
  if (!a.shock) return false;
  return shocked(a, j, i) || shocked(a, j, i - 1) || shocked(a, j, i + 1) ||
         shocked(a, j - 1, i) || shocked(a, j + 1, i);
}
// FLAG_HLL at the interface between two face-neighbour zones.
__device__ __forceinline__ bool hll_interface(const StageArgs& a, int j0, int i0, int j1,
                                              int i1) {
  if (!a.shock) return false;
  return shocked(a, j0, i0) || shocked(a, j1, i1);
}

// CHAR: the kernel is built with the characteristic-limiting path (kernels
// for grids without it leave it out: it costs registers).
template <int DIR, int CHAR>
__device__ __forceinline__ void plm_zone(const StageArgs& a, int j, int i, const double* vm1,
                                         const double* v0, const double* vp1, const double* c,
                                         double* vp, double* vm) {
  const bool mm = minmod_zone(a, j, i);
  if (CHAR && a.k.char_lim) plm_char_general<DIR>(vm1, v0, vp1, c, a.k.gamma, mm, vp, vm);
  else if (mm) plm_mm_general(vm1, v0, vp1, c, vp, vm);
  else plm_vl_general(vm1, v0, vp1, c, vp, vm);
}

// One RK2 stage (Src/Time_Stepping/rk_step.c:117-330 with UpdateStage,
// Src/Time_Stepping/update_stage.c:35-410, for GEOMETRY SPHERICAL in 2-D).
// FEAT: bit 0 builds the characteristic-limiting path, bit 1 the entropy
// switch, bit 2 the tabulated force multiplier (kernels for grids without
// them leave them out: they cost registers); grids of a kernel built with
// a feature still check it.
enum { FEAT_CHAR = 1, FEAT_ENT = 2, FEAT_TAB = 4, FEAT_ROE = 8 };

// VGradCalc + LineForce of both sweeps for zone (i, j) of the stage input
// (line_connect.c:628-738, :815-903): the x1 call takes the zone's PLM
// face average (vp + vm)/2, the x2 call the zone centre; g_r and g_theta
// go to a.glf.  A kernel of its own: the 36-bin loop with its two powers
// per bin needs few registers beside the accumulators, while inside the
// stage it would sit on top of the whole stencil (and drop the stage to
// one wave per SIMD).
template <int FEAT>
__device__ __forceinline__ void lforce_cell(const StageArgs& a, int i, int j) {
  const Const& k = a.k;
  const Tables& T = a.T;
  const int I = i + NG;
  const long long cell = (long long)j * a.nr + i;
  double vc[NVW], vc1[NVW];
  {
    double w1[NVW], e1[NVW], c[6], vpi[NVW], vmi[NVW];
    ld6(a, a.Vin, j, i, vc);
    ld6(a, j, i - 1, w1);
    ld6(a, j, i + 1, e1);
    ldc(T.pr, a.nr, I, c);
    plm_zone<0, FEAT & FEAT_CHAR>(a, j, i, w1, vc, e1, c, vpi, vmi);
#pragma unroll
    for (int nv = 0; nv < NVW; nv++) vc1[nv] = 0.5 * (vpi[nv] + vmi[nv]);
  }
  VGrad vg;
  vgrad_cell(a, j, i, vg);
  // thermal speeds of the two LineForce calls (line_connect.c:848-866)
  const double rho1 = vc1[RHO] * k.UD, rho2 = vc[RHO] * k.UD;
  const double T1 = vc1[PRS] / vc1[RHO] * k.KELVIN * k.mu;
  const double T2 = vc[PRS] / vc[RHO] * k.KELVIN * k.mu;
  const double vth1 = sqrt((2.0 * C_kB * T1) / C_mp);
  const double vth2 = sqrt((2.0 * C_kB * T2) / C_mp);
  const bool tab = (FEAT & FEAT_TAB) && k.fm_points > 0;
  // factored power: A = (sigma_e rho v_th)^alpha per call; the reference's
  // form: log2 of num = sigma_e rho v_th per call, once per zone
  double A1 = 0.0, A2 = 0.0;
  PowNum P1{}, P2{};
  if (!tab && k.fm_fast) {
    A1 = force_factor(k, rho1, vth1);
    A2 = force_factor(k, rho2, vth2);
  } else if (!tab) {
    P1 = pow_num(k.sigma_e * rho1 * vth1);
    P2 = pow_num(k.sigma_e * rho2 * vth2);
  }
  const long long sa = a.sv * a.na;
  const double* lf = T.lf + cell;
  const double rc_c = 1.0 / C_c, r_ua = 1.0 / k.UA;
  double g1r = 0.0, g2t = 0.0;
  // the bin loop in two copies, chosen once per run (k.fast_div): its
  // quotients through div_rc or through the division
  auto bins = [&](auto fast) {
    constexpr bool FAST = decltype(fast)::value;
// two bins at a time give the scheduler two independent chains
#ifndef PMX_ANG_UNROLL
#define PMX_ANG_UNROLL 2
#endif
    // a bin's five table entries are loaded one bin ahead (the loop is
    // bound by their latency otherwise: two waves per SIMD hide little)
    double nx[NLF];
#pragma unroll
    for (int f = 0; f < NLF; f++) nx[f] = lf[f * sa];
#pragma unroll PMX_ANG_UNROLL
    for (int ang = 0; ang < a.na; ang++) {
      double cur[NLF];
#pragma unroll
      for (int f = 0; f < NLF; f++) cur[f] = nx[f];
      {
        const double* q = lf + (long long)min(ang + 1, a.na - 1) * a.sv;
#pragma unroll
        for (int f = 0; f < NLF; f++) nx[f] = q[f * sa];
      }
      const double dvds = vgrad_dvds<FAST>(a, vg, cur[LF_F2], cur[LF_SIN], cur[LF_COS], ang);
      const double fr = cur[LF_FR], ft = cur[LF_FT];
      double fm1, fm2;
      if (tab) {   // both calls read the fit of this cell (line_connect.c:858-862)
        fm1 = force_mult_tab(k, rho1, vth1, dvds, T.fm_lt, T.fm_lm + cell, a.sv);
        fm2 = force_mult_tab(k, rho2, vth2, dvds, T.fm_lt, T.fm_lm + cell, a.sv);
      } else if (k.fm_fast) {
        force_mult_pair(k, A1, A2, dvds, fm1, fm2);
      } else {
        force_mult_shared(k, P1, P2, dvds, fm1, fm2);
      }
      // ((fm sigma_e f) / c) / UA, each quotient bit-identical
      g1r += qdiv<FAST>(qdiv<FAST>(fm1 * k.sigma_e * fr, C_c, rc_c), k.UA, r_ua);
      g2t += qdiv<FAST>(qdiv<FAST>(fm2 * k.sigma_e * ft, C_c, rc_c), k.UA, r_ua);
    }
  };
  if (k.fast_div) bins(std::true_type{});
  else bins(std::false_type{});
  a.glf[cell] = g1r;
  a.glf[a.sv + cell] = g2t;
}

template <int STAGE, int FEAT>
__device__ __forceinline__ void cell_stage(const StageArgs& a, int i, int j, double& cdt,
                                           int& flags) {
  {
    const Const& k = a.k;
    const Tables& T = a.T;
    const double dt = a.s->dt;
    const int I = i + NG, J = j + NG;
    const long long cell = (long long)j * a.nr + i;

    // ---- x1 stencil and reconstruction -------------------------------
    double vc[NVW], w1[NVW], e1[NVW];
    const int floored = ld6(a, a.Vin, j, i, vc);
    ld6(a, j, i - 1, w1);
    ld6(a, j, i + 1, e1);
    const bool ent = (FEAT & FEAT_ENT) && k.ent;
    const bool smm = !k.char_lim;      // scalars follow FLAG_MINMOD without characteristics
    double sc = 0.0, sw1 = 0.0, se1 = 0.0;
    if (ent) {
      sc = entropy_of(vc, k.gamma);
      sw1 = entropy_of(w1, k.gamma);
      se1 = entropy_of(e1, k.gamma);
    }
    double c[6];
    double vpi[NVW], vmi[NVW];
    ldc(T.pr, a.nr, I, c);
    plm_zone<0, FEAT & FEAT_CHAR>(a, j, i, w1, vc, e1, c, vpi, vmi);

    // ---- the line force of both sweeps, from the line-force kernel
    const double g1r = a.glf[cell], g2t = a.glf[a.sv + cell];

    // ---- x1 faces i-1/2 and i+1/2 --------------------------------------
    double rhs1[NVW], Fm[NVW], Fp[NVW], pm, pp, cm1, cp1;
    double fem = 0.0, fep = 0.0;       // entropy fluxes (adv_flux.c:131-134)
    {
      double w2[NVW], e2[NVW], vpl[NVW], vml[NVW], vpr[NVW], vmr[NVW];
      ld6(a, j, i - 2, w2);
      ld6(a, j, i + 2, e2);
      double cl[6], cr[6];
      ldc(T.pr, a.nr, I - 1, cl);
      plm_zone<0, FEAT & FEAT_CHAR>(a, j, i - 1, w2, w1, vc, cl, vpl, vml);
      ldc(T.pr, a.nr, I + 1, cr);
      plm_zone<0, FEAT & FEAT_CHAR>(a, j, i + 1, vc, e1, e2, cr, vpr, vmr);
      face<0, (FEAT & FEAT_ROE) != 0>(vpl, vmi, k, Fm, pm, cm1, hll_interface(a, j, i - 1, j, i));
      face<0, (FEAT & FEAT_ROE) != 0>(vpi, vmr, k, Fp, pp, cp1, hll_interface(a, j, i, j, i + 1));
      if (ent) {
        double spl, sml, spi, smi, spr, smr;
        plm_scalar(entropy_of(w2, k.gamma), sw1, sc, cl, smm && minmod_zone(a, j, i - 1), spl, sml);
        plm_scalar(sw1, sc, se1, c, smm && minmod_zone(a, j, i), spi, smi);
        plm_scalar(sc, se1, entropy_of(e2, k.gamma), cr, smm && minmod_zone(a, j, i + 1), spr, smr);
        fem = Fm[RHO] >= 0.0 ? spl * Fm[RHO] : smi * Fm[RHO];
        fep = Fp[RHO] >= 0.0 ? spi * Fp[RHO] : smr * Fp[RHO];
      }
    }
    double rhs1e = 0.0;
    {
      const double dmu = T.dmu[J];
      double xrm = T.x1r[I - 1], xrp = T.x1r[I];
      double Am = xrm * xrm * dmu, Ap = xrp * xrp * dmu;
      double fAm[NVW], fAp[NVW];
#pragma unroll
      for (int nv = 0; nv < NVW; nv++) { fAm[nv] = Fm[nv] * Am; fAp[nv] = Fp[nv] * Ap; }
      fAm[VX3] *= fabs(xrm);
      fAp[VX3] *= fabs(xrp);
      double dtdV = dt / (T.dVr[I] * dmu);
      double dtdl = dt / T.dx1[I];
#pragma unroll
      for (int nv = 0; nv < NVW; nv++) rhs1[nv] = -dtdV * (fAp[nv] - fAm[nv]);
      if (ent) rhs1e = -dtdV * (fep * Ap - fem * Am);
      rhs1[VX1] -= dtdl * (pp - pm);
      rhs1[VX3] /= fabs(T.x1[I]);
      // rhs_source.c:227-297
      double vs[NVW];
#pragma unroll
      for (int nv = 0; nv < NVW; nv++) vs[nv] = 0.5 * (vpi[nv] + vmi[nv]);
      double r_1 = 1.0 / T.x1[I];
      double vphi = vs[VX3];
      double Sm = vs[RHO] * (vs[VX2] * vs[VX2] + vphi * vphi);
      Sm += -0.0 - 0.0;
      rhs1[VX1] += dt * Sm * r_1;
      double g0 = -1.0 * k.gm_code / (T.x1[I] * T.x1[I]);
      rhs1[VX1] += dt * vs[RHO] * g0;
      rhs1[ENG] += dt * 0.5 * (Fp[RHO] + Fm[RHO]) * g0;
      rhs1[VX1] += dt * vs[RHO] * g1r;
      rhs1[ENG] += dt * 0.5 * (Fp[RHO] + Fm[RHO]) * g1r;
    }

    // ---- x2 faces j-1/2 and j+1/2 --------------------------------------
    double rhs2[NVW], Gm[NVW], Gp[NVW], qm, qp, cm2, cp2;
    double gem = 0.0, gep = 0.0;
    {
      double s1[NVW], n1[NVW], s2[NVW], n2[NVW];
      ld6(a, j - 1, i, s1);
      ld6(a, j + 1, i, n1);
      ld6(a, j - 2, i, s2);
      ld6(a, j + 2, i, n2);
      double vpc[NVW], vmc[NVW], vpl[NVW], vml[NVW], vpr[NVW], vmr[NVW];
      double cl[6], cr[6];
      ldc(T.pt, a.nt, J, c);
      plm_zone<1, FEAT & FEAT_CHAR>(a, j, i, s1, vc, n1, c, vpc, vmc);
      ldc(T.pt, a.nt, J - 1, cl);
      plm_zone<1, FEAT & FEAT_CHAR>(a, j - 1, i, s2, s1, vc, cl, vpl, vml);
      ldc(T.pt, a.nt, J + 1, cr);
      plm_zone<1, FEAT & FEAT_CHAR>(a, j + 1, i, vc, n1, n2, cr, vpr, vmr);
      face<1, (FEAT & FEAT_ROE) != 0>(vpl, vmc, k, Gm, qm, cm2, hll_interface(a, j - 1, i, j, i));
      face<1, (FEAT & FEAT_ROE) != 0>(vpc, vmr, k, Gp, qp, cp2, hll_interface(a, j, i, j + 1, i));
      if (ent) {
        const double ss1 = entropy_of(s1, k.gamma), sn1 = entropy_of(n1, k.gamma);
        double spl, sml, spc, smc, spr, smr;
        plm_scalar(entropy_of(s2, k.gamma), ss1, sc, cl, smm && minmod_zone(a, j - 1, i), spl, sml);
        plm_scalar(ss1, sc, sn1, c, smm && minmod_zone(a, j, i), spc, smc);
        plm_scalar(sc, sn1, entropy_of(n2, k.gamma), cr, smm && minmod_zone(a, j + 1, i), spr, smr);
        gem = Gm[RHO] >= 0.0 ? spl * Gm[RHO] : smc * Gm[RHO];
        gep = Gp[RHO] >= 0.0 ? spc * Gp[RHO] : smr * Gp[RHO];
      }
    }
    double rhs2e = 0.0;
    {
      double base = fabs(T.x1[I]) * T.dx1[I];
      double Am = base * T.sp[J - 1], Ap = base * T.sp[J];
      double fAm[NVW], fAp[NVW];
#pragma unroll
      for (int nv = 0; nv < NVW; nv++) { fAm[nv] = Gm[nv] * Am; fAp[nv] = Gp[nv] * Ap; }
      fAm[VX3] *= fabs(T.sp[J - 1]);
      fAp[VX3] *= fabs(T.sp[J]);
      double dtdV = dt / (T.dVr[I] * T.dmu[J]);
      double dtdl = dt / T.dx2[J] * (1.0 / T.rt[I]);
#pragma unroll
      for (int nv = 0; nv < NVW; nv++) rhs2[nv] = -dtdV * (fAp[nv] - fAm[nv]);
      if (ent) rhs2e = -dtdV * (gep * Ap - gem * Am);
      rhs2[VX2] -= dtdl * (qp - qm);
      rhs2[VX3] /= fabs(T.s[J]);
      // rhs_source.c:313-396
      double r_1 = 1.0 / T.rt[I], ct = T.ct[J];
      double vphi = vc[VX3];
      double Sm = vc[RHO] * (-vc[VX2] * vc[VX1] + ct * vphi * vphi);
      Sm += 0.0 - ct * 0.0;
      rhs2[VX2] += dt * Sm * r_1;
      const double g1 = 0.0, g2 = 0.0;
      rhs2[VX2] += dt * vc[RHO] * g1;
      rhs2[ENG] += dt * 0.5 * (Gp[RHO] + Gm[RHO]) * g1;
      rhs2[VX3] += dt * vc[RHO] * g2;
      rhs2[ENG] += dt * vc[RHO] * vc[VX3] * g2;
      const double l2 = 0.0;
      rhs2[VX2] += dt * vc[RHO] * g2t;
      rhs2[ENG] += dt * 0.5 * (Gp[RHO] + Gm[RHO]) * g2t;
      rhs2[VX3] += dt * vc[RHO] * l2;
      rhs2[ENG] += dt * vc[RHO] * vc[VX3] * l2;
    }

    // ---- update, RK2 combination, cons->prim ---------------------------
    double u[NVW];
    // a zone of the midplane row that the floors changed: its Uc is
    // PrimToCons of the floored state, before the midplane reset
    double vfl[NVW];
    const bool mid_row = k.midplane && j == a.nt - 1;
    if (STAGE == 2 && floored && mid_row) {
#pragma unroll
      for (int nv = 0; nv < NVW; nv++) vfl[nv] = a.Vin[nv * a.sv + cell];
      floor_cell(k, vfl);
    }
    if (STAGE == 1 || floored) {
      // stage 1: U0 = PrimToCons of the bounded state (rk_step.c:124-125);
      // stage 2: a cell the floors changed gets PrimToCons of its new state
      p2c6((STAGE == 2 && mid_row) ? vfl : vc, u, k.gmm1);
    } else {
#pragma unroll
      for (int nv = 0; nv < NVW; nv++) u[nv] = a.U[nv * a.sv + cell];
    }
#pragma unroll
    for (int nv = 0; nv < NVW; nv++) u[nv] += rhs1[nv];
#pragma unroll
    for (int nv = 0; nv < NVW; nv++) u[nv] += rhs2[nv];
    if (STAGE == 2) {
      double vn[NVW], u0[NVW];
#pragma unroll
      for (int nv = 0; nv < NVW; nv++) vn[nv] = a.Vn[nv * a.sv + cell];
      if (k.dfloor > 0.0) floor_cell(k, vn);   // U0 was formed after the floors
      if (mid_row) midplane_cell(a, i, vn);    // and the midplane row
      p2c6(vn, u0, k.gmm1);
#pragma unroll
      for (int nv = 0; nv < NVW; nv++) u[nv] = 0.5 * u0[nv] + 0.5 * u[nv];
    }
    double vo[NVW];
    if (ent) {
      // Uc[ENTR]: stage 1 PrimToCons of the bounded state; stage 2 the
      // stage-1 value, or for a zone the floors changed rho times the
      // entropy the stage-1 cons->prim left (Ue / rho of the stage-1 U)
      double ue;
      if (STAGE == 1) ue = vc[RHO] * sc;
      else if (floored) ue = (mid_row ? vfl[RHO] : vc[RHO]) *
                             (a.Ue[cell] * (1.0 / a.U[RHO * a.sv + cell]));
      else ue = a.Ue[cell];
      ue += rhs1e;
      ue += rhs2e;
      if (STAGE == 2) {
        double vn[NVW];
#pragma unroll
        for (int nv = 0; nv < NVW; nv++) vn[nv] = a.Vn[nv * a.sv + cell];
        if (k.dfloor > 0.0) floor_cell(k, vn);
        if (mid_row) midplane_cell(a, i, vn);
        const double u0e = vn[RHO] * entropy_of(vn, k.gamma);
        ue = 0.5 * u0e + 0.5 * ue;
      }
      flags = c2p6_entropy(u, ue, vo, k.gmm1);
      a.Ue[cell] = ue;
    } else {
      flags = c2p6(u, vo, k.gmm1);
    }
#pragma unroll
    for (int nv = 0; nv < NVW; nv++) {
      a.U[nv * a.sv + cell] = u[nv];
      a.Vout[nv * a.sv + cell] = vo[nv];
    }
    if (STAGE == 1) {   // update_stage.c:303-316
      cdt = 0.0;
      cdt += 0.5 * (cm1 + cp1) * T.inv_dx1[I] * 1.0;
      cdt += 0.5 * (cm2 + cp2) * (T.inv_dx2[J] * (1.0 / T.x1[I])) * 1.0;
    }
  }
}

// Block reduction of a stage: max C_dt (stage 1) and OR of the cons->prim
// flags into the grid's scalars.  Every thread of the block calls it.
template <int STAGE, int NT>
__device__ __forceinline__ bool block_reduce_stage(Scalars* s, double cdt, int flags) {
  __shared__ double smax[NT / 64];
  __shared__ int sflag[NT / 64];
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  for (int off = 32; off > 0; off >>= 1) {
    cdt = dmax(cdt, __shfl_xor(cdt, off));
    flags |= __shfl_xor(flags, off);
  }
  if ((tid & 63) == 0) { smax[tid >> 6] = cdt; sflag[tid >> 6] = flags; }
  __syncthreads();
  if (tid == 0) {
    double m = smax[0];
    int f = sflag[0];
    for (int q = 1; q < NT / 64; q++) { m = dmax(m, smax[q]); f |= sflag[q]; }
    if (STAGE == 1 && m > 0.0)
      atomicMax(&s->cdt_red[red_index(blockIdx.x + blockIdx.y * gridDim.x)],
                (unsigned long long)__double_as_longlong(m));
    if (f) atomicOr(&s->err, f);
  }
  __syncthreads();   // the shared slots are reused by the next call
  return tid == 0;
}

// FlagShock with SHOCK_FLATTENING MULTID (Src/flag_shock.c:81-253) for the
// interior and two ghost layers (the reference runs MULTID with 3 ghost
// zones and tests all but the outermost), from the stage-1 input as the
// boundary leaves it (floors and side rules through ld6).  Spherical 2-D:
// div v from the face areas r^2 dmu and r dr |sin theta| and the volume
// dVr dmu (Src/set_geometry.c:118-186); shocked if div v < 0 and
// |dp_1| + |dp_2| > 5 min(p of the zone and its 4 neighbours).
__global__ __launch_bounds__(BX* BY) void wind_flag_kernel(StageArgs a, unsigned char* shock) {
  const int I = blockIdx.x * BX + threadIdx.x, J = blockIdx.y * BY + threadIdx.y;
  if (I >= a.nr + 4 || J >= a.nt + 4) return;
  if (a.loop && a.s->halt) return;
  const int i = I - NG, j = J - NG;
  const Tables& T = a.T;
  double c[NVW], e[NVW], w[NVW], n[NVW], s[NVW];
  ld6(a, a.Vin, j, i, c);
  ld6(a, a.Vin, j, i + 1, e);
  ld6(a, a.Vin, j, i - 1, w);
  ld6(a, a.Vin, j + 1, i, n);
  ld6(a, a.Vin, j - 1, i, s);
  const double dmu = T.dmu[J];
  const double x1p = T.x1r[I], x1m = I > 0 ? T.x1r[I - 1] : T.x1l0;
  const double A1p = x1p * x1p * dmu, A1m = x1m * x1m * dmu;
  const double base = fabs(T.x1[I]) * T.dx1[I];
  const double A2p = base * T.sp[J], A2m = base * (J > 0 ? T.sp[J - 1] : T.sin_x2l0);
  const double dvx1 = A1p * (e[VX1] + c[VX1]) - A1m * (w[VX1] + c[VX1]);
  const double dvx2 = A2p * (n[VX2] + c[VX2]) - A2m * (s[VX2] + c[VX2]);
  const double divv = (dvx1 + dvx2) / (T.dVr[I] * dmu);
  unsigned char f = 0;
  if (divv < 0.0) {
    double pt_min = c[PRS];
    const double pt_min1 = dmin(e[PRS], w[PRS]), pt_min2 = dmin(n[PRS], s[PRS]);
    pt_min = dmin(pt_min, pt_min1);
    pt_min = dmin(pt_min, pt_min2);
    const double gradp = fabs(e[PRS] - w[PRS]) + fabs(n[PRS] - s[PRS]);
    f = gradp > 5.0 * pt_min;
  }
  shock[(long long)J * (a.nr + 4) + I] = f;
}

template <int STAGE, int FEAT>
__device__ __forceinline__ void stage_grid(const StageArgs& a) {
  if (a.loop && a.s->halt) return;     // the run has stopped (uniform over the launch)
  const int i = blockIdx.x * BX + threadIdx.x;
  const int j = blockIdx.y * BY + threadIdx.y;
  double cdt = 0.0;
  int flags = 0;
  if (i < a.nr && j < a.nt) cell_stage<STAGE, FEAT>(a, i, j, cdt, flags);
  if (block_reduce_stage<STAGE, BX * BY>(a.s, cdt, flags)) {
    if (STAGE == 2 && blockIdx.x == 0 && blockIdx.y == 0) {
      // update_stage.c:389-394: invDt = MAX(invDt, C_dt) over the domain,
      // then / DIMENSIONS (stage-1 values, complete when stage 2 runs)
      double inv = a.s->inv_dt_hyp;
      inv = dmax(inv, take_cdt(a.s));
      inv /= (double)2;
      a.s->inv_dt_hyp = inv;
    }
  }
}

// The plain stage fits 2 waves per SIMD by itself (forcing the attribute
// on it cost 13% while the line force was still part of it); with
// characteristic limiting, the entropy switch or the
// tabulated force multiplier the allocation grows past 256 registers
// (VGPRs + AGPRs) and falls to 1 wave unless capped.
template <int STAGE, int FEAT>
__global__ __launch_bounds__(BX* BY) void wind_stage_kernel(StageArgs a) {
  stage_grid<STAGE, FEAT>(a);
}

template <int STAGE, int FEAT>
__global__ __launch_bounds__(BX* BY) __attribute__((amdgpu_waves_per_eu(PMX_WIND_WAVES, PMX_WIND_WAVES)))
void wind_stage_kernel_w(StageArgs a) {
  stage_grid<STAGE, FEAT>(a);
}

// The stage's line force (lforce_cell) over the interior, before the stage
// kernel.  FEAT: FEAT_CHAR and FEAT_TAB as for the stage.
template <int FEAT>
__global__ __launch_bounds__(BX* BY) void wind_lforce_kernel(StageArgs a) {
  if (a.loop && a.s->halt) return;     // uniform over the launch
  const int i = blockIdx.x * BX + threadIdx.x;
  const int j = blockIdx.y * BY + threadIdx.y;
  if (i < a.nr && j < a.nt) lforce_cell<FEAT>(a, i, j);
}

template <int MODE>
__global__ __launch_bounds__(128) void wind_cool_kernel(CoolArgs a, unsigned long long* calls) {
  if (a.loop && a.s->halt) return;     // uniform over the launch
  const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double cand = 1.0e300;
  int bad = 0;
  long long ncalls = 0;
  if (cell < a.sv) {
    const int i = (int)(cell % a.nr);
    const double t = a.s->t, dt = a.s->dt;
    double rho = a.V[RHO * a.sv + cell], p = a.V[PRS * a.sv + cell];
    double pn;
    if (MODE == PMX_COOL_BLONDIN) {
      double pre[5];
      load_pre(a.pre, a.sv, cell, pre);
      pn = blondin_cell(a.k, rho, p, a.x1[i + NG], a.xi_s[cell], a.tr_s[cell], t, dt, pre);
    } else {
      RadCell R;
      R.rho = rho;
      R.trc = a.V[TRC * a.sv + cell];
      R.x1 = a.x1[i + NG];
      R.xi_s = a.xi_s[cell];
      R.tr_s = a.tr_s[cell];
      load_pre(a.pre, a.sv, cell, R.pre);
      R.t = t;
      R.nev = 0;
      pn = plain_cell(a.k, R, p, dt, cand, bad);
      ncalls = R.nev;
    }
    a.V[PRS * a.sv + cell] = pn;
  }
  if (MODE == PMX_COOL_PLAIN) {
    // one device atomic per workgroup and reduction, each into the
    // workgroup's spread slot
    __shared__ double smin[2];
    __shared__ int sbad[2];
    __shared__ long long scalls[2];
    for (int off = 32; off > 0; off >>= 1) {
      cand = dmin(cand, __shfl_xor(cand, off));
      bad |= __shfl_xor(bad, off);
      ncalls += __shfl_xor(ncalls, off);
    }
    if ((threadIdx.x & 63) == 0) {
      smin[threadIdx.x >> 6] = cand;
      sbad[threadIdx.x >> 6] = bad;
      scalls[threadIdx.x >> 6] = ncalls;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const long long nc = scalls[0] + scalls[1];
      if (nc) atomicAdd(calls + red_index(blockIdx.x), (unsigned long long)nc);
      double m = dmin(smin[0], smin[1]);
      if (m >= 0.0)
        atomicMin(&a.s->dtcool_red[red_index(blockIdx.x)],
                  (unsigned long long)__double_as_longlong(m));
      if (sbad[0] | sbad[1]) atomicOr(&a.s->err, PMX_FLAG_COOL_SUBSTEPS);
    }
  }
}

// Src/main.c:296 (t += dt) and NextTimeStep (main.c:521-697) followed by
// the invDt / dt_cool reset that the next step performs first
// (main.c:406-415).  With cooling both happen every second step
// (main.c:329-333), without every step: `next` says whether this is one.
// Then the run-stopping conditions and the cut of the next dt to tstop
// (end_of_step); a halted run does nothing.
__global__ void wind_end_kernel(Scalars* s, int next, double cfl, double cfl_max_var,
                                double first_dt) {
  if (s->halt) return;
  double dt = s->dt;
  s->t += dt;
  fold_dtcool(s);
  double dt_cool = __longlong_as_double((long long)s->dtcool_bits);
  s->dt_cool = dt_cool;
  if (next) {
    double dt_hyp = 1.0 / s->inv_dt_hyp;
    dt_hyp *= cfl;
    double dtn = dt_hyp;
    dtn = dmin(dtn, dt_cool);
    dtn = dmin(dtn, cfl_max_var * dt);
    if (dtn < first_dt * 1.e-9) s->halt |= PMX_FLAG_DT_TOO_SMALL;   // main.c:672-683
    if (cfl_max_var != 1.0) s->dt = dtn;
    s->inv_dt_hyp = 0.0;
    s->dtcool_bits = (unsigned long long)__double_as_longlong(1.e38);
  }
  end_of_step(s);
}

// ComputeUserVar of the production problem (cv_idl/userdef_output.c:20-260)
// for one cell: uv[q * sv + cell] in PMX_UV_* order, from the state Vf, the
// SIROCCO arrays, the prefactors pre[6][sv] and, for the line-force
// quantities, the dv/ds of the last stage's input a.Vin (the reference's
// dvds_array, set by the last VGradCalc).  The reference's LineForce and
// its M loop use two differently rounded sigma_e (sigmaT/amu/1.18 and
// sigmaT/(amu*1.18)); M_UV, t_UV, t_UV2 and dvds2 start at 0 where the
// reference leaves them uninitialised.
enum { UV_XI, UV_T, UV_TR, UV_CH, UV_CC, UV_LC, UV_BC, UV_XH, UV_CHP, UV_CCP, UV_LCP, UV_BCP,
       UV_XHP, UV_XIP, UV_NE, UV_NH, UV_GR, UV_GT, UV_GP, UV_DVDS, UV_TT, UV_M1, UV_M2,
       UV_COUNT };
static_assert(UV_COUNT == PMX_NUV, "user variable count");

__global__ __launch_bounds__(128) void wind_uservar_kernel(StageArgs a, const double* Vf,
                                                           const double* xi_s,
                                                           const double* tr_s,
                                                           const double* pre, double* uv) {
  const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= a.sv) return;
  const int i = (int)(cell % a.nr), j = (int)(cell / a.nr);
  const Const& k = a.k;
  const Tables& T = a.T;
  const long long sv = a.sv;
  double* o = uv + cell;
  double v[NVW];
#pragma unroll
  for (int nv = 0; nv < NVW; nv++) v[nv] = Vf[nv * sv + cell];
  const double t = a.s->t;
  const double r = T.x1[i + NG] * k.UL;
  const double rho = v[RHO] * k.UD;
  const double Tk = v[PRS] / v[RHO] * k.KELVIN * k.mu;
  o[UV_T * sv] = Tk;
  const double nH = rho / (1.43 * C_mp);
  o[UV_NH * sv] = nH;
  double ne_ratio;
  if (Tk < 1.5e4) ne_ratio = 1e-2 + pow(10.0, (-51.59417133 + 12.27740153 * log10(Tk)));
  else if (Tk >= 1.5e4 && Tk < 3.3e4) ne_ratio = pow(10.0, (-3.80749689 + 0.86092628 * log10(Tk)));
  else ne_ratio = 1.21;
  o[UV_NE * sv] = nH * ne_ratio;
  const double v_th = sqrt(2.0 * C_kB * Tk / C_mp);
  double xi;
  if (t <= 1.0) {
    xi = k.lx / (nH * r * r);
    o[UV_TR * sv] = k.t_x;
  } else {
    xi = xi_s[cell];
    o[UV_TR * sv] = tr_s[cell];
  }
  o[UV_XI * sv] = xi;
  const double sqrt_T = sqrt(Tk), sqxi = sqrt(xi), sqsqxi = pow(xi, 0.25);
  double p6[6];
#pragma unroll
  for (int q = 0; q < 6; q++) p6[q] = pre[q * sv + cell];
  o[UV_CHP * sv] = p6[0]; o[UV_CCP * sv] = p6[1]; o[UV_XHP * sv] = p6[2];
  o[UV_BCP * sv] = p6[3]; o[UV_LCP * sv] = p6[4]; o[UV_XIP * sv] = p6[5];
  const double lc = ((1e-16 * exp(-1.3e5 / Tk) / sqxi / Tk) +
                     fmin(fmin(1e-24, 5e-27 * sqrt_T), 1.5e-17 / Tk));
  if (t < 3.0) {
    o[UV_CH * sv] = (8.9e-36 * xi * k.t_x);
    o[UV_CC * sv] = (8.9e-36 * xi * (4.0 * Tk));
    o[UV_XH * sv] = (1.5e-21 * (sqsqxi / sqrt_T));
    o[UV_LC * sv] = lc;
    o[UV_BC * sv] = (3.3e-27 * sqrt_T);
  } else {
    o[UV_CH * sv] = p6[0] * (8.9e-36 * xi * k.t_x);
    o[UV_CC * sv] = p6[1] * (8.9e-36 * xi * (4.0 * Tk));
    o[UV_XH * sv] = p6[2] * (1.5e-21 * (sqsqxi / sqrt_T));
    o[UV_LC * sv] = p6[4] * lc;
    o[UV_BC * sv] = p6[3] * (3.3e-27 * sqrt_T);
  }
  // LineForce(Vc, grad) and the M loop over the bins
  VGrad vg;
  vgrad_cell(a, j, i, vg);
  const bool tab = k.fm_points > 0;
  const double sig_u = C_sigmaT / (C_amu * 1.18);
  const long long sa = a.sv * a.na;
  const double* lf = T.lf + cell;
  double gr = 0.0, gt = 0.0;
  double M_UV = 0.0, t_UV = 0.0, t_UV2 = 0.0, dvds2 = 0.0, M1 = -1.0, M2 = -1.0, fmax = 1e-99;
  for (int ang = 0; ang < a.na; ang++) {
    const double* q = lf + (long long)ang * a.sv;
    const double dvds = vgrad_dvds(a, vg, q, sa, ang);
    const double fr = q[LF_FR * sa], ft = q[LF_FT * sa];
    double M = 0.0;
    if (dvds > 0.0) {
      const double tl = k.sigma_e * rho * v_th / dvds;
      M = tab ? linterp_fm(log10(tl), T.fm_lt, T.fm_lm + cell, sv, k.fm_points)
              : k.krad * pow(tl, k.alpharad);
      if (M > 4400.) M = 4400.;
    }
    gr += ((1.0 + M) * k.sigma_e * fr / C_c) / k.UA;
    gt += ((1.0 + M) * k.sigma_e * ft / C_c) / k.UA;
    if (dvds > 0.0) {
      t_UV = sig_u * rho * v_th / dvds;
      M_UV = tab ? linterp_fm(log10(t_UV), T.fm_lt, T.fm_lm + cell, sv, k.fm_points)
                 : k.krad * pow(t_UV, k.alpharad);
      if (M_UV > 4400.) M_UV = 4400.;
      if (M_UV > M1) M1 = M_UV;
    }
    if (pow(fr, 2.0) + pow(ft, 2.0) > fmax) {
      fmax = pow(M_UV * fr, 2.0) + pow(M_UV * ft, 2.0);
      M2 = M_UV;
      t_UV2 = t_UV;
      dvds2 = dvds;
    }
  }
  o[UV_GR * sv] = gr;
  o[UV_GT * sv] = gt;
  o[UV_GP * sv] = 0.0;
  o[UV_M1 * sv] = M1;
  o[UV_M2 * sv] = M2;
  o[UV_DVDS * sv] = dvds2;
  o[UV_TT * sv] = t_UV2;
}

__global__ void heatcool_from_state_kernel(const double* V, int nr, long long sv,
                                           const double* x1, Const k, double* xi_s,
                                           double* tr_s) {
  const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= sv) return;
  const int i = (int)(cell % nr);
  double rho = V[RHO * sv + cell] * k.UD, r = x1[i + NG] * k.UL;
  double nH = rho / (1.43 * C_mp);
  xi_s[cell] = k.lx / nH / (r * r);
  tr_s[cell] = k.t_x;
}

}  // namespace wind
}  // namespace pmx

using namespace pmx;
using namespace pmx::wind;

struct pmx_wind {
  int nr, nt, na, cooling;
  int bc[4];
  double cfl, cfl_max_var;
  long long sv;
  Const k;
  double* V[2];       // V[0] = V^n (state), V[1] = stage-1 result
  double* U;
  double* xi_s;
  double* tr_s;
  double* pre;        // [6][sv] prefactors (1.0 until set); the cooling reads the first 5
  double* fm;         // tabulated force multiplier: log10 t, then log10 M; or null
  double* tab;        // one allocation for all geometry tables
  double* lf;
  unsigned char* shock;  // SHOCK_FLATTENING MULTID zone flags, or null
  double* Ue;            // ENTROPY_SWITCH: Uc[ENTR], or null
  double* glf;           // the stage's line force, [2][sv]
  Tables T;
  Scalars* s;
  long step;
  hipStream_t stream;
  bool timing;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pend_stage, pend_cool;
  double stage_ms, cool_ms;
  int stage_n, cool_n;
  pmx_mlp* mlp;        // surrogate for the stiff cells (PMX_COOL_PLAIN), or null
  SurrWork sw;
  CoolArgs* cargs;     // device copy of this grid's cooling arguments (1 entry)
  long long* coff;     // {0, sv}
  bool stepped;        // a step has run: V[1] holds the last stage-2 input
  // output snapshot: staging copy in HBM, pinned host copy, copy stream
  hipStream_t copy_stream;
  double* snap_dev;
  double* snap_host;
  hipEvent_t snap_ready, snap_done;
  bool snap_pending;
  int snap_nvar;
  long snap_step;
  // pmx_wind_run_graph: GRAPH_STEPS captured steps from an even one
  hipGraphExec_t graph;
  bool have_graph;
};

namespace {

constexpr double C_AH = 1.008, C_AHe = 4.004, C_AZ = 30.0;   // Src/pluto.h:296-301
constexpr double C_G = 6.6726e-8, C_PI = 3.14159265358979;
constexpr double C_sigma = 5.67051e-5;                          // Src/pluto.h:320

int timed_begin(pmx_wind* w, hipEvent_t* e) {
  if (!w->timing) return 0;
  if (w->ev_pool.size() < 2) {
    for (int q = 0; q < 64; q++) {
      hipEvent_t ev;
      PMX_HIP_CHECK(hipEventCreate(&ev));
      w->ev_pool.push_back(ev);
    }
  }
  e[0] = w->ev_pool.back(); w->ev_pool.pop_back();
  e[1] = w->ev_pool.back(); w->ev_pool.pop_back();
  PMX_HIP_CHECK(hipEventRecord(e[0], w->stream));
  return 0;
}

int timed_end(pmx_wind* w, hipEvent_t* e, bool cool) {
  if (!w->timing) return 0;
  PMX_HIP_CHECK(hipEventRecord(e[1], w->stream));
  (cool ? w->pend_cool : w->pend_stage).push_back({e[0], e[1]});
  return 0;
}

int collect(pmx_wind* w) {
  for (int c = 0; c < 2; c++) {
    auto& v = c ? w->pend_cool : w->pend_stage;
    for (auto& p : v) {
      float ms = 0.f;
      PMX_HIP_CHECK(hipEventSynchronize(p.second));
      PMX_HIP_CHECK(hipEventElapsedTime(&ms, p.first, p.second));
      if (c) { w->cool_ms += ms; w->cool_n++; } else { w->stage_ms += ms; w->stage_n++; }
      w->ev_pool.push_back(p.first);
      w->ev_pool.push_back(p.second);
    }
    v.clear();
  }
  return 0;
}

StageArgs stage_args(pmx_wind* w, const double* Vin, double* Vout, bool loop = false) {
  StageArgs a;
  a.Vin = Vin;
  a.Vn = w->V[0];
  a.U = w->U;
  a.Vout = Vout;
  a.nr = w->nr; a.nt = w->nt; a.na = w->na;
  for (int q = 0; q < 4; q++) a.bc[q] = w->bc[q];
  a.sv = w->sv;
  a.k = w->k;
  a.T = w->T;
  a.s = w->s;
  a.shock = w->shock;
  a.Ue = w->Ue;
  a.glf = w->glf;
  a.loop = loop ? 1 : 0;
  return a;
}

// FlagShock of the step (SHOCK_FLATTENING MULTID) from V^n, before stage 1.
int flag_shock(pmx_wind* w, hipStream_t st, bool loop) {
  if (!w->shock) return 0;
  dim3 grid((w->nr + 4 + BX - 1) / BX, (w->nt + 4 + BY - 1) / BY), block(BX, BY);
  hipLaunchKernelGGL(wind_flag_kernel, grid, block, 0, st,
                     stage_args(w, w->V[0], w->V[1], loop), w->shock);
  PMX_HIP_CHECK(hipGetLastError());
  return 0;
}

int features(const pmx_wind* w) {
  return (w->k.char_lim ? FEAT_CHAR : 0) | (w->k.ent ? FEAT_ENT : 0) |
         (w->k.fm_points > 0 ? FEAT_TAB : 0) | (w->k.roe ? FEAT_ROE : 0);
}

// Stage 1 of the feature kernels needs the 2-wave cap (it also carries the
// CFL reduction), and so do the stage-2 kernels that would otherwise take
// accumulation registers and drop to one wave per SIMD (characteristic
// limiting with the entropy switch, and the tabulated force multiplier);
// stage 2 of the others fits 2 waves by itself, where the cap only adds
// spills.
template <int STAGE, int FEAT>
void launch_feat(pmx_wind* w, dim3 grid, dim3 block, const StageArgs& a) {
  if (STAGE == 1 || (FEAT & FEAT_TAB) || (FEAT & 3) == 3)
    hipLaunchKernelGGL((wind_stage_kernel_w<STAGE, FEAT>), grid, block, 0, w->stream, a);
  else
    hipLaunchKernelGGL((wind_stage_kernel<STAGE, FEAT>), grid, block, 0, w->stream, a);
}

template <int STAGE>
void launch_stage(pmx_wind* w, dim3 grid, dim3 block, const StageArgs& a) {
  switch (features(w) & (FEAT_CHAR | FEAT_TAB)) {
    case 0: hipLaunchKernelGGL((wind_lforce_kernel<0>), grid, block, 0, w->stream, a); break;
    case FEAT_CHAR:
      hipLaunchKernelGGL((wind_lforce_kernel<FEAT_CHAR>), grid, block, 0, w->stream, a);
      break;
    default:
      hipLaunchKernelGGL((wind_lforce_kernel<FEAT_CHAR | FEAT_TAB>), grid, block, 0, w->stream, a);
  }
  switch (features(w)) {
    case 0: hipLaunchKernelGGL((wind_stage_kernel<STAGE, 0>), grid, block, 0, w->stream, a); break;
    case 1: launch_feat<STAGE, 1>(w, grid, block, a); break;
    case 2: launch_feat<STAGE, 2>(w, grid, block, a); break;
    case 3: launch_feat<STAGE, 3>(w, grid, block, a); break;
    case FEAT_ROE: launch_feat<STAGE, FEAT_ROE>(w, grid, block, a); break;
    // the tabulated force multiplier, or Roe with other features: one
    // kernel with every feature (grids still check each)
    default:
      if (features(w) & FEAT_ROE) launch_feat<STAGE, 15>(w, grid, block, a);
      else launch_feat<STAGE, 7>(w, grid, block, a);
  }
}

int advance(pmx_wind* w, bool loop) {
  dim3 grid((w->nr + BX - 1) / BX, (w->nt + BY - 1) / BY), block(BX, BY);
  w->stepped = true;
  hipEvent_t e[2];
  if (flag_shock(w, w->stream, loop)) return -1;
  if (timed_begin(w, e)) return -1;
  launch_stage<1>(w, grid, block, stage_args(w, w->V[0], w->V[1], loop));
  PMX_HIP_CHECK(hipGetLastError());
  if (timed_end(w, e, false)) return -1;
  if (timed_begin(w, e)) return -1;
  launch_stage<2>(w, grid, block, stage_args(w, w->V[1], w->V[0], loop));
  PMX_HIP_CHECK(hipGetLastError());
  return timed_end(w, e, false);
}

CoolArgs cool_args(pmx_wind* w) {
  CoolArgs a;
  a.V = w->V[0];
  a.nr = w->nr; a.nt = w->nt;
  a.sv = w->sv;
  a.k = w->k;
  a.x1 = w->T.x1;
  a.xi_s = w->xi_s;
  a.tr_s = w->tr_s;
  a.pre = w->pre;
  a.s = w->s;
  a.dtc = nullptr;
  a.loop = 0;
  return a;
}

int cool(pmx_wind* w, bool loop) {
  if (w->cooling == PMX_COOL_NONE) return 0;
  CoolArgs a = cool_args(w);
  a.loop = loop ? 1 : 0;
  dim3 grid((unsigned)((w->sv + 127) / 128)), block(128);
  hipEvent_t e[2];
  if (timed_begin(w, e)) return -1;
  if (w->mlp) {
    CoolBatch b;
    b.args = w->cargs; b.off = w->coff; b.G = 1; b.total = w->sv;
    b.loop = loop ? 1 : 0;
    if (mlp::surrogate_cool(w->mlp, b, w->sw, w->stream)) return -1;
  } else if (w->cooling == PMX_COOL_BLONDIN)
    hipLaunchKernelGGL(wind_cool_kernel<PMX_COOL_BLONDIN>, grid, block, 0, w->stream, a,
                       w->sw.cnt + SURR_CALLS);
  else
    hipLaunchKernelGGL(wind_cool_kernel<PMX_COOL_PLAIN>, grid, block, 0, w->stream, a,
                       w->sw.cnt + SURR_CALLS);
  PMX_HIP_CHECK(hipGetLastError());
  return timed_end(w, e, true);
}

// Src/set_geometry.c (SPHERICAL) and Src/States/plm_coeffs.c:30-91 on the
// ghost-inclusive 1-D arrays, evaluated on the host with the reference's
// expressions (the trigonometric ones with the host libm, as in the
// reference).
void plm_coeffs(const double* x, const double* xr, const double* dx, const double* xgc, int n,
                double* pc) {
  for (int q = 0; q < 6 * n; q++) pc[q] = 0.0;
  for (int i = 1; i <= n - 2; i++) {
    pc[2 * n + i] = dx[i] / (xgc[i + 1] - xgc[i]);
    pc[3 * n + i] = dx[i] / (xgc[i] - xgc[i - 1]);
    pc[0 * n + i] = (xgc[i + 1] - xgc[i]) / (xr[i] - xgc[i]);
    pc[1 * n + i] = (xgc[i] - xgc[i - 1]) / (xgc[i] - xr[i - 1]);
    pc[4 * n + i] = (xr[i] - xgc[i]) / dx[i];
    pc[5 * n + i] = (xgc[i] - xr[i - 1]) / dx[i];
  }
  (void)x;
}

}  // namespace

namespace pmx {
namespace wind {
// Problem constants from the descriptor's scalars (Src/pluto.h:700-720,
// the problem's definitions.h), with the reference's expressions.
Const make_const(const pmx_wind_desc* d) {
  Const k;
  k.gamma = d->gamma;
  k.gmm1 = d->gamma - 1.0;
  k.UD = d->unit_density; k.UL = d->unit_length; k.UV = d->unit_velocity;
  k.UT = (k.UL / k.UV);
  k.UP = (k.UD * k.UV * k.UV);
  k.UA = (k.UV * k.UV / k.UL);
  k.KELVIN = (k.UV * k.UV * C_amu / C_kB);
  k.mu = d->mu;
  {
    const double FHe = 0.2741 / C_AHe * C_AH / 0.7110;           // Src/pluto.h:718-720
    const double FZ = (1.0 - 0.7110 - 0.2741) / C_AZ * C_AH / 0.7110;
    k.mu_tab = (C_AH + FHe * C_AHe + FZ * C_AZ) / (2.0 + 3.0 * FHe + FZ * (1.0 + C_AZ * 0.5));
  }
  k.gm_code = C_G * d->cent_mass / (k.UL * k.UV * k.UV);
  k.sigma_e = C_sigmaT / C_amu / 1.18;
  k.krad = d->krad; k.alpharad = d->alpharad;
  k.lx = d->l_star * d->f_xray;
  k.t_x = d->t_x;
  k.E_cost = k.UL / k.UD / pow(k.UV, 3.0);
  // cv_idl/init.c:184-188: dfloor = DFLOOR / UNIT_DENSITY, pfloor with T = 500 K
  k.dfloor = d->dfloor / k.UD;
  k.pfloor = k.dfloor * 5.e2 / (k.KELVIN * d->mu);
  k.char_lim = d->char_limiting != 0;
  k.ent = d->entropy_switch != 0;
  k.hll = d->solver == PMX_SOLVER_HLL;
  k.roe = d->solver == PMX_SOLVER_ROE;
  k.first_dt = d->first_dt;
  k.fm_points = (d->krad == 999 && d->alpharad == 999) ? d->fm_points : 0;
  k.fast_div = 0;   // set with the line-force tables
  k.fm_fast = d->force_power == PMX_FORCE_POWER_FACTORED;
  k.midplane = 0;   // set with the geometry
  return k;
}

// div_rc against the division on the device: the count of x whose two
// quotients differ in any bit.
__global__ void div_check_kernel(const double* __restrict__ x, long long n, double c, double rc,
                                 unsigned long long* bad) {
  unsigned long long mine = 0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double q = div_rc(x[i], c, rc), r = x[i] / c;
    if (__double_as_longlong(q) != __double_as_longlong(r)) mine++;
  }
  if (mine) atomicAdd(bad, mine);
}

// The line force's reference-form power krad pow(num / dvds, alpha) as the
// line-force kernel evaluates it (force_power_shared), per element; 0 where
// dvds <= 0 (a bin the reference leaves at M = 0).
__global__ void power_check_kernel(const double* __restrict__ num, const double* __restrict__ dvds,
                                   long long n, Const k, double* __restrict__ out) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double d = dvds[i];
    out[i] = d > 0.0 ? force_power_shared(k, pow_num(num[i]), d, log2_dd(d)) : 0.0;
  }
}
}  // namespace wind
}  // namespace pmx

extern "C" {

int pmx_wind_destroy(pmx_wind* w) {
  if (!w) return 0;
  if (w->stream) hipStreamSynchronize(w->stream);
  for (auto e : w->ev_pool) hipEventDestroy(e);
  for (auto& p : w->pend_stage) { hipEventDestroy(p.first); hipEventDestroy(p.second); }
  for (auto& p : w->pend_cool) { hipEventDestroy(p.first); hipEventDestroy(p.second); }
  hipFree(w->V[0]); hipFree(w->V[1]); hipFree(w->U); hipFree(w->xi_s); hipFree(w->tr_s);
  hipFree(w->pre); hipFree(w->fm); hipFree(w->tab); hipFree(w->lf); hipFree(w->s); hipFree(w->shock); hipFree(w->Ue);
  hipFree(w->glf);
  hipFree(w->sw.list); hipFree(w->sw.fb); hipFree(w->sw.feat);
  hipFree(w->sw.f0); hipFree(w->sw.cnt); hipFree(w->cargs); hipFree(w->coff);
  if (w->copy_stream) {
    hipStreamSynchronize(w->copy_stream);
    hipStreamDestroy(w->copy_stream);
    hipEventDestroy(w->snap_ready);
    hipEventDestroy(w->snap_done);
  }
  hipFree(w->snap_dev);
  if (w->snap_host) hipHostFree(w->snap_host);
  if (w->have_graph) hipGraphExecDestroy(w->graph);
  if (w->stream) hipStreamDestroy(w->stream);
  delete w;
  return 0;
}

int pmx_wind_create(const pmx_wind_desc* d, pmx_wind** out) {
  if (!d || !out) { set_error("null argument"); return -1; }
  if (d->nx1 < 2 || d->nx2 < 2) { set_error("grid smaller than the ghost width"); return -1; }
  if (d->nangles < 1 || d->nangles > NANG_MAX) { set_error("nangles out of range"); return -1; }
  if (d->cooling < PMX_COOL_NONE || d->cooling > PMX_COOL_PLAIN) {
    set_error("unknown cooling mode"); return -1;
  }
  for (int q = 0; q < 4; q++)
    if (d->bc[q] < 0 || d->bc[q] > 4) { set_error("unknown boundary type"); return -1; }
  if (d->bc[0] == PMX_BC_WIND_POLE || d->bc[2] == PMX_BC_WIND_INNER || d->bc[1] > 2 ||
      d->bc[3] > 2) {
    set_error("PMX_BC_WIND_INNER is an r-beg and PMX_BC_WIND_POLE a theta-beg boundary");
    return -1;
  }
  if (!(d->dfloor >= 0.0)) { set_error("dfloor must be >= 0"); return -1; }
  if (d->solver != PMX_SOLVER_HLLC && d->solver != PMX_SOLVER_HLL &&
      d->solver != PMX_SOLVER_ROE) {
    set_error("the wind's solver is PMX_SOLVER_HLLC, PMX_SOLVER_HLL or PMX_SOLVER_ROE");
    return -1;
  }
  if (!d->x1 || !d->x1l || !d->x1r || !d->dx1 || !d->x2 || !d->x2l || !d->x2r || !d->dx2 ||
      !d->flux_r || !d->flux_t || !d->flux_p) {
    set_error("missing grid or flux arrays"); return -1;
  }
  const bool tabulated = d->krad == 999 && d->alpharad == 999;
  if (tabulated && (d->fm_points < 1 || d->fm_points > FM_POINTS_MAX || !d->fm_t || !d->fm_m)) {
    set_error("krad = alpharad = 999 needs the force multiplier fits (fm_points 1..4096, "
              "fm_t, fm_m)");
    return -1;
  }
  pmx_wind* w = new pmx_wind();
  w->nr = d->nx1; w->nt = d->nx2; w->na = d->nangles; w->cooling = d->cooling;
  for (int q = 0; q < 4; q++) w->bc[q] = d->bc[q];
  w->cfl = d->cfl; w->cfl_max_var = d->cfl_max_var;
  w->sv = (long long)w->nr * w->nt;
  w->step = 0;
  w->k = make_const(d);
  const Const& k = w->k;


  const int tr = w->nr + 2 * NG, tt = w->nt + 2 * NG, nr = w->nr, nt = w->nt, na = w->na;
  // host tables: radial 6 + 6 PLM rows, polar 9 + 6 PLM rows, centre weights, bin angles
  std::vector<double> x1(d->x1, d->x1 + tr), x1r(d->x1r, d->x1r + tr), x1l(d->x1l, d->x1l + tr),
      dx1(d->dx1, d->dx1 + tr);
  std::vector<double> x2(d->x2, d->x2 + tt), x2r(d->x2r, d->x2r + tt), x2l(d->x2l, d->x2l + tt),
      dx2(d->dx2, d->dx2 + tt);
  std::vector<double> rt(tr), inv_dx1(tr), dVr(tr), xgc1(tr), pr(6 * tr);
  for (int i = 0; i < tr; i++) {
    double xL = x1l[i], xR = x1r[i], x = x1[i], dx = dx1[i];
    xgc1[i] = x + 2.0 * x * dx * dx / (12.0 * x * x + dx * dx);
    rt[i] = (xR * xR * xR - xL * xL * xL) / (xR * xR - xL * xL) / 1.5;
    inv_dx1[i] = 1.0 / dx;
    dVr[i] = fabs(xR * xR * xR - xL * xL * xL) / 3.0;
  }
  std::vector<double> s(tt), sp(tt), dmu(tt), inv_dx2(tt), ct(tt), sin2(tt), cos2(tt), xgc2(tt),
      pt(6 * tt);
  for (int j = 0; j < tt; j++) {
    double xL = x2l[j], xR = x2r[j];
    xgc2[j] = sin(xR) - sin(xL) + xL * cos(xL) - xR * cos(xR);
    xgc2[j] /= cos(xL) - cos(xR);
    sp[j] = fabs(sin(xR));
    s[j] = fabs(sin(x2[j]));
    dmu[j] = fabs(cos(xL) - cos(xR));
    inv_dx2[j] = 1.0 / dx2[j];
    ct[j] = 1.0 / tan(x2[j]);
    sin2[j] = sin(x2[j]);
    cos2[j] = cos(x2[j]);
  }
  plm_coeffs(x1.data(), x1r.data(), dx1.data(), xgc1.data(), tr, pr.data());
  plm_coeffs(x2.data(), x2r.data(), dx2.data(), xgc2.data(), tt, pt.data());

  // The cv_idl side-0 pass's midplane row (cv_idl/init.c:276-311): with
  // SHOCK_FLATTENING MULTID the reference runs with 3 ghost zones, and the
  // pass's row j == np_int + 2 is the last interior polar row; it is reset
  // when that row's centre and volume centroid agree to 1e-6 (:280-281).
  // Per column: the disk density rho_0 (r / r_in)^-alpha, the Keplerian
  // v_phi and the pressure of the disk temperature at the zone's centroid,
  // with the reference's expressions and the host libm.
  std::vector<double> mid;
  if (d->dfloor > 0.0 && d->shock_flattening) {
    const int J = nt - 1 + NG;
    if ((fabs(x2[J] - xgc2[J]) / x2[J]) < 1e-6) {
      if (!(d->rho_0 > 0.0) || !(d->disk_mdot >= 0.0)) {
        set_error("the cv_idl midplane row needs rho_0 > 0 and disk_mdot >= 0 (RHO_0, "
                  "DISK_MDOT)");
        delete w;
        return -1;
      }
      mid.resize(3 * (size_t)nr);
      const double rho_0 = d->rho_0 / k.UD, rho_alpha = d->rho_alpha, disk_mdot = d->disk_mdot;
      const double r_WD = x1l[NG];                                   // g_domBeg[IDIR]
      const double gm_cgs = C_G * d->cent_mass, gm_code = gm_cgs / (k.UL * k.UV * k.UV);
      const double theta = xgc2[J];
      for (int i = 0; i < nr; i++) {
        const double r = xgc1[i + NG], rcyl = r * sin(theta);
        const double rho_mid = rho_0 * pow((r / r_WD), -1.0 * rho_alpha);
        double teff = pow(3.0 * gm_cgs * disk_mdot / (8.0 * C_PI * C_sigma), 0.25);
        teff *= pow(r_WD * k.UL, -0.75);
        const double temp = teff * pow(r_WD / rcyl, 0.75) * pow(1.0 - sqrt(r_WD / rcyl), 0.25);
        mid[i] = rho_mid;
        mid[nr + i] = sqrt(gm_code / r) * sin(theta);
        mid[2 * nr + i] = rho_mid * temp / (k.KELVIN * d->mu);
      }
      w->k.midplane = 1;
    }
  }

  // VGradCalc set-up (line_connect.c:534-625): offset point per (bin,
  // cell), turned into bilinear weights and its trigonometric factors.
  const double UL = k.UL;
  std::vector<double> f1c(nr), f2c(nt), sta(na), cta(na);
  std::vector<double> lf((size_t)NLF * na * w->sv, 0.0);
  for (int a = 0; a < na; a++) {
    double ta = (a + 0.5) * (2.0 * C_PI) / 36.0;
    sta[a] = sin(ta);
    cta[a] = cos(ta);
  }
  for (int i = 0; i < nr; i++) {
    int I = i + NG;
    double x11_0 = (x1[I - 1] + x1[I]) / 2.0 * UL, x22_0 = (x1[I + 1] + x1[I]) / 2.0 * UL;
    f1c[i] = (x1[I] * UL - x11_0) / (x22_0 - x11_0);
  }
  for (int j = 0; j < nt; j++) {
    int J = j + NG;
    double x11_1 = (x2[J - 1] + x2[J]) / 2.0, x22_1 = (x2[J + 1] + x2[J]) / 2.0;
    f2c[j] = (x2[J] - x11_1) / (x22_1 - x11_1);
  }
  const size_t sa = (size_t)na * w->sv;
  // div_rc's ranges hold for every line-force quotient when the fluxes are
  // 0 or of magnitude 2^-600..2^600 (the dividend is fm sigma_e f with fm
  // in 1..4401), UA lies in 2^-100..2^100 and the radii x1 UL in
  // 2^-200..2^200 (the radial weight's dividend and divisor are differences
  // of such radii, 0 or at least an ulp of them)
  auto div_rc_flux_ok = [](double f) {
    const double af = fabs(f);
    return af == 0.0 || (af >= 0x1p-600 && af <= 0x1p600);
  };
  bool fast_div = k.UA >= 0x1p-100 && k.UA <= 0x1p100;
  for (int I = 0; I < nr + 2 * NG; I++) {
    const double r = x1[I] * UL;
    if (!(r >= 0x1p-200 && r <= 0x1p200)) fast_div = false;
  }
  for (int a = 0; a < na; a++)
    for (int j = 0; j < nt; j++)
      for (int i = 0; i < nr; i++) {
        int I = i + NG, J = j + NG;
        size_t c = (size_t)j * nr + i, q = (size_t)a * w->sv + c;
        double x11_0 = (x1[I - 1] + x1[I]) / 2.0 * UL, x11_1 = (x2[J - 1] + x2[J]) / 2.0;
        double x22_0 = (x1[I + 1] + x1[I]) / 2.0 * UL, x22_1 = (x2[J + 1] + x2[J]) / 2.0;
        double maxds = fabs(x22_0 - x11_0);
        if (maxds > (x1[I] * UL * fabs(x22_1 - x11_1))) maxds = x1[I] * UL * fabs(x22_1 - x11_1);
        maxds /= 2.0;
        double fr = d->flux_r[q], ft = d->flux_t[q], fp = d->flux_p[q];
        if (!div_rc_flux_ok(fr) || !div_rc_flux_ok(ft)) fast_div = false;
        double mod_flux = sqrt(pow(fr, 2) + pow(ft, 2) + pow(fp, 2));
        lf[LF_FR * sa + q] = fr;
        lf[LF_FT * sa + q] = ft;
        if (mod_flux == 0.0) {
          lf[LF_F2 * sa + q] = -999;
          continue;
        }
        double x = x1[I] * sin(x2[J]) * UL, z = x1[I] * cos(x2[J]) * UL;
        double ta = (a + 0.5) * (2.0 * C_PI) / 36.0;
        double ddx1 = maxds * sin(ta), ddx2 = maxds * cos(ta);
        // ds and the radial weight (roff - x11_0) / (x22_0 - x11_0) are
        // recomputed per bin by the stage (vgrad_dvds), bit for bit
        double toff = atan((x + ddx1) / (z + ddx2));
        lf[LF_F2 * sa + q] = (toff - x11_1) / (x22_1 - x11_1);
        lf[LF_SIN * sa + q] = sin(toff);
        lf[LF_COS * sa + q] = cos(toff);
      }
  w->k.fast_div = fast_div ? 1 : 0;

  // pack the geometry tables into one buffer
  std::vector<std::pair<const double**, const std::vector<double>*>> parts = {
      {&w->T.x1, &x1}, {&w->T.dx1, &dx1}, {&w->T.x1r, &x1r}, {&w->T.rt, &rt},
      {&w->T.inv_dx1, &inv_dx1}, {&w->T.dVr, &dVr}, {&w->T.pr, &pr},
      {&w->T.x2, &x2}, {&w->T.dx2, &dx2}, {&w->T.s, &s}, {&w->T.sp, &sp}, {&w->T.dmu, &dmu},
      {&w->T.inv_dx2, &inv_dx2}, {&w->T.ct, &ct}, {&w->T.sin2, &sin2}, {&w->T.cos2, &cos2},
      {&w->T.pt, &pt}, {&w->T.f1c, &f1c}, {&w->T.f2c, &f2c}, {&w->T.sta, &sta},
      {&w->T.cta, &cta}, {&w->T.mid, &mid}};
  size_t ntab = 0;
  for (auto& p : parts) ntab += p.second->size();
  std::vector<double> packed;
  packed.reserve(ntab);
  std::vector<size_t> offs;
  for (auto& p : parts) {
    offs.push_back(packed.size());
    packed.insert(packed.end(), p.second->begin(), p.second->end());
  }

#define CK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { \
    set_error(std::string(#x) + ": " + hipGetErrorString(e_)); pmx_wind_destroy(w); return -1; } } while (0)
  CK(hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking));
  const size_t vbytes = sizeof(double) * NVW * w->sv;
  CK(hipMalloc(&w->V[0], vbytes));
  CK(hipMalloc(&w->V[1], vbytes));
  CK(hipMalloc(&w->U, vbytes));
  CK(hipMalloc(&w->xi_s, sizeof(double) * w->sv));
  CK(hipMalloc(&w->tr_s, sizeof(double) * w->sv));
  CK(hipMalloc(&w->tab, sizeof(double) * packed.size()));
  CK(hipMalloc(&w->lf, sizeof(double) * lf.size()));
  CK(hipMalloc(&w->s, sizeof(Scalars)));
  if (d->shock_flattening) CK(hipMalloc(&w->shock, (size_t)(w->nr + 4) * (w->nt + 4)));
  if (d->entropy_switch) CK(hipMalloc(&w->Ue, sizeof(double) * w->sv));
  CK(hipMalloc(&w->glf, sizeof(double) * 2 * w->sv));
  CK(hipMalloc(&w->pre, sizeof(double) * 6 * w->sv));
  {
    std::vector<double> ones((size_t)6 * w->sv, 1.0);
    CK(hipMemcpy(w->pre, ones.data(), sizeof(double) * ones.size(), hipMemcpyHostToDevice));
  }
  if (tabulated) {   // the reader keeps log10 of t and M (line_connect.c:215, :228)
    const size_t np = (size_t)d->fm_points, nm = np * (size_t)w->sv;
    std::vector<double> lg(np + nm);
    for (size_t q = 0; q < np; q++) lg[q] = log10(d->fm_t[q]);
    for (size_t q = 0; q < nm; q++) lg[np + q] = log10(d->fm_m[q]);
    CK(hipMalloc(&w->fm, sizeof(double) * lg.size()));
    CK(hipMemcpy(w->fm, lg.data(), sizeof(double) * lg.size(), hipMemcpyHostToDevice));
    w->T.fm_lt = w->fm;
    w->T.fm_lm = w->fm + np;
  } else {
    w->T.fm_lt = w->T.fm_lm = nullptr;
  }
  w->T.x1l0 = x1l[0];
  w->T.sin_x2l0 = fabs(sin(x2l[0]));
  CK(hipMalloc(&w->sw.cnt, sizeof(unsigned long long) * PMX_SURR_COUNTERS));
  CK(hipMemset(w->sw.cnt, 0, sizeof(unsigned long long) * PMX_SURR_COUNTERS));
  CK(hipMalloc(&w->cargs, sizeof(CoolArgs)));
  CK(hipMalloc(&w->coff, sizeof(long long) * 2));
  CK(hipMemcpy(w->tab, packed.data(), sizeof(double) * packed.size(), hipMemcpyHostToDevice));
  CK(hipMemcpy(w->lf, lf.data(), sizeof(double) * lf.size(), hipMemcpyHostToDevice));
  CK(hipMemset(w->V[0], 0, vbytes));
  CK(hipMemset(w->V[1], 0, vbytes));
  CK(hipMemset(w->U, 0, vbytes));
  CK(hipMemset(w->xi_s, 0, sizeof(double) * w->sv));
  CK(hipMemset(w->tr_s, 0, sizeof(double) * w->sv));
  for (size_t q = 0; q < parts.size(); q++) *parts[q].first = w->tab + offs[q];
  w->T.lf = w->lf;
  {
    CoolArgs ca = cool_args(w);
    const long long off[2] = {0, w->sv};
    CK(hipMemcpy(w->cargs, &ca, sizeof(ca), hipMemcpyHostToDevice));
    CK(hipMemcpy(w->coff, off, sizeof(off), hipMemcpyHostToDevice));
  }
  Scalars s0;
  memset(&s0, 0, sizeof(s0));
  red_clear(s0);
  s0.tstop = HUGE_VAL;
  s0.dt_cool = 1.e38;
  s0.dtcool_bits = (unsigned long long)0;
  memcpy(&s0.dtcool_bits, &s0.dt_cool, sizeof(double));
  CK(hipMemcpy(w->s, &s0, sizeof(Scalars), hipMemcpyHostToDevice));
#undef CK
  *out = w;
  return 0;
}

int pmx_wind_set_prim(pmx_wind* w, const double* V) {
  if (!w || !V) { set_error("null argument"); return -1; }
  PMX_HIP_CHECK(hipMemcpyAsync(w->V[0], V, sizeof(double) * NVW * w->sv, hipMemcpyHostToDevice,
                               w->stream));
  PMX_HIP_CHECK(hipStreamSynchronize(w->stream));
  return 0;
}

int pmx_wind_get_prim(pmx_wind* w, double* V) {
  if (!w || !V) { set_error("null argument"); return -1; }
  PMX_HIP_CHECK(hipMemcpyAsync(V, w->V[0], sizeof(double) * NVW * w->sv, hipMemcpyDeviceToHost,
                               w->stream));
  PMX_HIP_CHECK(hipStreamSynchronize(w->stream));
  return 0;
}

int pmx_wind_get_cons(pmx_wind* w, double* U) {
  if (!w || !U) { set_error("null argument"); return -1; }
  PMX_HIP_CHECK(hipMemcpyAsync(U, w->U, sizeof(double) * NVW * w->sv, hipMemcpyDeviceToHost,
                               w->stream));
  PMX_HIP_CHECK(hipStreamSynchronize(w->stream));
  return 0;
}

int pmx_wind_set_heatcool(pmx_wind* w, const double* xi, const double* t_r) {
  if (!w) { set_error("null argument"); return -1; }
  if (xi) {
    if (!t_r) { set_error("t_r missing"); return -1; }
    PMX_HIP_CHECK(hipMemcpyAsync(w->xi_s, xi, sizeof(double) * w->sv, hipMemcpyHostToDevice,
                                 w->stream));
    PMX_HIP_CHECK(hipMemcpyAsync(w->tr_s, t_r, sizeof(double) * w->sv, hipMemcpyHostToDevice,
                                 w->stream));
  } else {
    hipLaunchKernelGGL(heatcool_from_state_kernel, dim3((unsigned)((w->sv + 255) / 256)),
                       dim3(256), 0, w->stream, w->V[0], w->nr, w->sv, w->T.x1, w->k, w->xi_s,
                       w->tr_s);
    PMX_HIP_CHECK(hipGetLastError());
  }
  PMX_HIP_CHECK(hipStreamSynchronize(w->stream));
  return 0;
}

int pmx_wind_set_prefactors(pmx_wind* w, const double* pre) {
  if (!w) { set_error("null argument"); return -1; }
  if (pre) {
    PMX_HIP_CHECK(hipMemcpyAsync(w->pre, pre, sizeof(double) * 6 * w->sv,
                                 hipMemcpyHostToDevice, w->stream));
    PMX_HIP_CHECK(hipStreamSynchronize(w->stream));
  } else {
    std::vector<double> ones((size_t)6 * w->sv, 1.0);
    PMX_HIP_CHECK(hipMemcpyAsync(w->pre, ones.data(), sizeof(double) * ones.size(),
                                 hipMemcpyHostToDevice, w->stream));
    PMX_HIP_CHECK(hipStreamSynchronize(w->stream));
  }
  return 0;
}

int pmx_wind_snapshot(pmx_wind* w, int user_vars) {
  if (!w) { set_error("null argument"); return -1; }
  if (user_vars && !w->stepped) {
    set_error("user variables need a step first: the line-force quantities use the dv/ds "
              "of the last stage");
    return -1;
  }
  const size_t nsnap = (size_t)(NVW + PMX_NUV) * w->sv + 2;
  if (!w->snap_dev) {
    PMX_HIP_CHECK(hipStreamCreateWithFlags(&w->copy_stream, hipStreamNonBlocking));
    PMX_HIP_CHECK(hipEventCreateWithFlags(&w->snap_ready, hipEventDisableTiming));
    PMX_HIP_CHECK(hipEventCreateWithFlags(&w->snap_done, hipEventDisableTiming));
    PMX_HIP_CHECK(hipMalloc(&w->snap_dev, sizeof(double) * nsnap));
    PMX_HIP_CHECK(hipHostMalloc(&w->snap_host, sizeof(double) * nsnap));
  } else {
    // the staging copy is reused: the previous transfer must be out of it
    PMX_HIP_CHECK(hipStreamWaitEvent(w->stream, w->snap_done, 0));
  }
  const size_t nv = NVW + (user_vars ? PMX_NUV : 0);
  double* tail = w->snap_dev + (size_t)(NVW + PMX_NUV) * w->sv;
  PMX_HIP_CHECK(hipMemcpyAsync(w->snap_dev, w->V[0], sizeof(double) * NVW * w->sv,
                               hipMemcpyDeviceToDevice, w->stream));
  if (user_vars) {
    hipLaunchKernelGGL(wind_uservar_kernel, dim3((unsigned)((w->sv + 127) / 128)), dim3(128), 0,
                       w->stream, stage_args(w, w->V[1], w->V[0]), w->V[0], w->xi_s, w->tr_s,
                       w->pre, w->snap_dev + (size_t)NVW * w->sv);
    PMX_HIP_CHECK(hipGetLastError());
  }
  // t and dt: the first two doubles of the device scalars
  PMX_HIP_CHECK(hipMemcpyAsync(tail, w->s, sizeof(double) * 2, hipMemcpyDeviceToDevice,
                               w->stream));
  PMX_HIP_CHECK(hipEventRecord(w->snap_ready, w->stream));
  PMX_HIP_CHECK(hipStreamWaitEvent(w->copy_stream, w->snap_ready, 0));
  PMX_HIP_CHECK(hipMemcpyAsync(w->snap_host, w->snap_dev, sizeof(double) * nv * w->sv,
                               hipMemcpyDeviceToHost, w->copy_stream));
  PMX_HIP_CHECK(hipMemcpyAsync(w->snap_host + (size_t)(NVW + PMX_NUV) * w->sv, tail,
                               sizeof(double) * 2, hipMemcpyDeviceToHost, w->copy_stream));
  PMX_HIP_CHECK(hipEventRecord(w->snap_done, w->copy_stream));
  w->snap_pending = true;
  w->snap_nvar = (int)nv;
  w->snap_step = w->step;
  return 0;
}

int pmx_wind_snapshot_wait(pmx_wind* w, double* out, double* t, double* dt, long* step) {
  if (!w) { set_error("null argument"); return -1; }
  if (!w->snap_pending) { set_error("no snapshot taken"); return -1; }
  PMX_HIP_CHECK(hipEventSynchronize(w->snap_done));
  const double* tail = w->snap_host + (size_t)(NVW + PMX_NUV) * w->sv;
  if (out) memcpy(out, w->snap_host, sizeof(double) * (size_t)w->snap_nvar * w->sv);
  if (t) *t = tail[0];
  if (dt) *dt = tail[1];
  if (step) *step = w->snap_step;
  return 0;
}

int pmx_wind_snapshot_host(pmx_wind* w, const double** data) {
  if (!w || !data) { set_error("null argument"); return -1; }
  if (!w->snap_pending) { set_error("no snapshot taken"); return -1; }
  PMX_HIP_CHECK(hipEventSynchronize(w->snap_done));
  *data = w->snap_host;
  return 0;
}

int pmx_wind_set_time(pmx_wind* w, double t, double dt, long step) {
  if (!w) { set_error("null argument"); return -1; }
  Scalars s0;
  PMX_HIP_CHECK(hipMemcpyAsync(&s0, w->s, sizeof(Scalars), hipMemcpyDeviceToHost, w->stream));
  PMX_HIP_CHECK(hipStreamSynchronize(w->stream));
  s0.t = t;
  s0.dt = dt;
  s0.step = step;
  s0.halt = 0;
  s0.last = 0;
  clamp_last_step(s0);
  PMX_HIP_CHECK(hipMemcpyAsync(w->s, &s0, sizeof(Scalars), hipMemcpyHostToDevice, w->stream));
  PMX_HIP_CHECK(hipStreamSynchronize(w->stream));
  w->step = step;
  return 0;
}

int pmx_wind_set_tstop(pmx_wind* w, double tstop) {
  if (!w || !(tstop > -HUGE_VAL)) { set_error("bad argument"); return -1; }
  Scalars s0;
  PMX_HIP_CHECK(hipMemcpyAsync(&s0, w->s, sizeof(Scalars), hipMemcpyDeviceToHost, w->stream));
  PMX_HIP_CHECK(hipStreamSynchronize(w->stream));
  s0.tstop = tstop;
  s0.halt &= ~PMX_FLAG_TSTOP;
  if (s0.last) {           // undo a cut made for the previous tstop
    s0.dt = s0.dt_uncut;
    s0.last = 0;
  }
  clamp_last_step(s0);
  PMX_HIP_CHECK(hipMemcpyAsync(w->s, &s0, sizeof(Scalars), hipMemcpyHostToDevice, w->stream));
  PMX_HIP_CHECK(hipStreamSynchronize(w->stream));
  w->step = (long)s0.step;   // steps the halted run skipped do not count
  return 0;
}

int pmx_wind_set_dts(pmx_wind* w, double inv_dt_hyp, double dt_cool) {
  if (!w) { set_error("null argument"); return -1; }
  Scalars s0;
  PMX_HIP_CHECK(hipMemcpyAsync(&s0, w->s, sizeof(Scalars), hipMemcpyDeviceToHost, w->stream));
  PMX_HIP_CHECK(hipStreamSynchronize(w->stream));
  s0.inv_dt_hyp = inv_dt_hyp;
  s0.dt_cool = dt_cool;
  memcpy(&s0.dtcool_bits, &dt_cool, sizeof(double));
  red_clear(s0);
  PMX_HIP_CHECK(hipMemcpyAsync(w->s, &s0, sizeof(Scalars), hipMemcpyHostToDevice, w->stream));
  PMX_HIP_CHECK(hipStreamSynchronize(w->stream));
  return 0;
}

int pmx_wind_get_time(pmx_wind* w, double* t, double* dt, long* step, double* inv_dt_hyp,
                      double* dt_cool, int* flags) {
  if (!w) { set_error("null argument"); return -1; }
  Scalars s0;
  PMX_HIP_CHECK(hipMemcpyAsync(&s0, w->s, sizeof(Scalars), hipMemcpyDeviceToHost, w->stream));
  PMX_HIP_CHECK(hipStreamSynchronize(w->stream));
  if (t) *t = s0.t;
  if (dt) *dt = s0.dt;
  if (step) *step = (long)s0.step;
  if (inv_dt_hyp) *inv_dt_hyp = s0.inv_dt_hyp;
  if (dt_cool) {
    double c;
    fold_dtcool(&s0);
    memcpy(&c, &s0.dtcool_bits, sizeof(double));
    *dt_cool = c;
  }
  if (flags) *flags = s0.err | s0.halt;
  return 0;
}

int pmx_wind_advance_step(pmx_wind* w) {
  if (!w) { set_error("null argument"); return -1; }
  return advance(w, false);
}

int pmx_wind_split_source(pmx_wind* w) {
  if (!w) { set_error("null argument"); return -1; }
  return cool(w, false);
}

// One iteration of the main loop for step number st (Strang order by
// parity, NextTimeStep every second step with cooling).
int wind_step(pmx_wind* w, long st) {
  if (st % 2 == 0) {
    if (advance(w, true) || cool(w, true)) return -1;
  } else {
    if (cool(w, true) || advance(w, true)) return -1;
  }
  const int next = (w->cooling == PMX_COOL_NONE) || (st % 2 == 1);
  hipLaunchKernelGGL(wind_end_kernel, dim3(1), dim3(1), 0, w->stream, w->s, next, w->cfl,
                     w->cfl_max_var, w->k.first_dt);
  PMX_HIP_CHECK(hipGetLastError());
  return 0;
}

int pmx_wind_run(pmx_wind* w, int nsteps) {
  if (!w || nsteps < 0) { set_error("bad argument"); return -1; }
  for (int n = 0; n < nsteps; n++) {
    if (wind_step(w, w->step)) return -1;
    w->step++;
  }
  return 0;
}

// Steps per captured graph (an even number: the graph starts on an even
// step, whose split cooling comes after the hydro step): a graph launch
// costs more than the launches of one step pair save, so a graph holds
// several pairs.  Steps after the run halts are no-ops on the device.
constexpr int GRAPH_STEPS = 8;

int pmx_wind_run_graph(pmx_wind* w, int nsteps) {
  if (!w || nsteps < 0) { set_error("bad argument"); return -1; }
  if (w->timing) return pmx_wind_run(w, nsteps);
  int n = 0;
  if (w->step % 2 == 1 && n < nsteps) {       // align to an even step
    if (wind_step(w, w->step)) return -1;
    w->step++;
    n++;
  }
  if (nsteps - n >= GRAPH_STEPS && !w->have_graph) {
    hipGraph_t gph;
    PMX_HIP_CHECK(hipStreamBeginCapture(w->stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (int q = 0; q < GRAPH_STEPS; q++) rc |= wind_step(w, q);
    const hipError_t e = hipStreamEndCapture(w->stream, &gph);
    if (rc || e != hipSuccess) { set_error("graph capture failed"); return -1; }
    PMX_HIP_CHECK(hipGraphInstantiate(&w->graph, gph, nullptr, nullptr, 0));
    PMX_HIP_CHECK(hipGraphDestroy(gph));
    w->have_graph = true;
  }
  for (; nsteps - n >= GRAPH_STEPS; n += GRAPH_STEPS) {
    PMX_HIP_CHECK(hipGraphLaunch(w->graph, w->stream));
    w->step += GRAPH_STEPS;
  }
  for (; n < nsteps; n++) {
    if (wind_step(w, w->step)) return -1;
    w->step++;
  }
  if (nsteps > 0) w->stepped = true;
  return 0;
}

int pmx_wind_sync(pmx_wind* w) {
  if (!w) { set_error("null argument"); return -1; }
  PMX_HIP_CHECK(hipStreamSynchronize(w->stream));
  return 0;
}

int pmx_wind_set_surrogate(pmx_wind* w, pmx_mlp* m, int guess) {
  if (!w) { set_error("null argument"); return -1; }
  if (m && w->cooling != PMX_COOL_PLAIN) {
    set_error("the surrogate replaces the plain cooling integrator (PMX_COOL_PLAIN)");
    return -1;
  }
  if (m && !w->sw.list) {
    const size_t n = (size_t)w->sv;
    PMX_HIP_CHECK(hipMalloc(&w->sw.list, sizeof(long long) * n));
    PMX_HIP_CHECK(hipMalloc(&w->sw.fb, sizeof(long long) * n));
    PMX_HIP_CHECK(hipMalloc(&w->sw.feat, sizeof(float) * 4 * n));
    PMX_HIP_CHECK(hipMalloc(&w->sw.f0, sizeof(double) * n));
  }
  w->mlp = m;
  w->sw.guess = guess;
  if (w->have_graph) {                        // the captured step used the old cooling
    PMX_HIP_CHECK(hipGraphExecDestroy(w->graph));
    w->have_graph = false;
  }
  return 0;
}

int pmx_wind_surrogate_stats(pmx_wind* w, long long* listed, long long* accepted,
                             long long* fallback, long long* radiat_calls, int reset) {
  if (!w) { set_error("null argument"); return -1; }
  unsigned long long c[PMX_SURR_COUNTERS];
  {
    PMX_HIP_CHECK(hipMemcpyAsync(c, w->sw.cnt, sizeof(c), hipMemcpyDeviceToHost, w->stream));
    PMX_HIP_CHECK(hipStreamSynchronize(w->stream));
    if (reset)
      PMX_HIP_CHECK(hipMemset(w->sw.cnt + 4, 0, sizeof(unsigned long long) * (PMX_SURR_COUNTERS - 4)));
  }
  if (listed) *listed = (long long)c[4];
  if (accepted) *accepted = (long long)c[6];
  if (fallback) *fallback = (long long)c[5];
  if (radiat_calls) *radiat_calls = (long long)surr_calls(c);
  return 0;
}

int pmx_wind_timing(pmx_wind* w, int enable, int reset, double* stage_ms, int* stage_launches,
                    double* cool_ms, int* cool_launches) {
  if (!w) { set_error("null argument"); return -1; }
  if (collect(w)) return -1;
  if (stage_ms) *stage_ms = w->stage_ms;
  if (stage_launches) *stage_launches = w->stage_n;
  if (cool_ms) *cool_ms = w->cool_ms;
  if (cool_launches) *cool_launches = w->cool_n;
  if (reset) { w->stage_ms = w->cool_ms = 0.0; w->stage_n = w->cool_n = 0; }
  w->timing = enable != 0;
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------
// Batched sweep: many independent wind grids of uneven sizes advanced
// together (no counterpart in the reference, which runs one grid per
// process).  One persistent kernel per RK stage walks the list of (grid,
// tile) work items; the step loop is captured in a HIP graph; the tile
// shape and the number of persistent workgroups are chosen by timing.
// ---------------------------------------------------------------------
namespace pmx {
namespace wind {

__device__ __forceinline__ int find_item(const int* off, int G, int q) {
  int lo = 0, hi = G - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= q) lo = mid; else hi = mid - 1;
  }
  return lo;
}

template <int STAGE, int FEAT, int TBX, int TBY>
__global__ __launch_bounds__(TBX* TBY) __attribute__((amdgpu_waves_per_eu(PMX_WIND_WAVES, PMX_WIND_WAVES)))
void wind_batch_stage_kernel(const StageArgs* __restrict__ args, const int* __restrict__ tile_off,
                             int G, int total) {
  for (int item = blockIdx.x; item < total; item += gridDim.x) {
    // the grid index is the same for the whole workgroup: made explicit, the
    // grid's arguments are read with scalar loads instead of occupying
    // vector registers (76 -> 28 B of spills per lane, C5 440 -> 477
    // Mcell-updates/s)
    const int g = __builtin_amdgcn_readfirstlane(find_item(tile_off, G, item));
    const StageArgs& a = args[g];
    if (a.s->halt) continue;           // a halted grid (uniform over the workgroup)
    const int tx_n = (a.nr + TBX - 1) / TBX;
    const int local = item - tile_off[g];
    const int ty = local / tx_n, tx = local - ty * tx_n;
    const int i = tx * TBX + threadIdx.x, j = ty * TBY + threadIdx.y;
    double cdt = 0.0;
    int flags = 0;
    if (i < a.nr && j < a.nt) cell_stage<STAGE, FEAT>(a, i, j, cdt, flags);
    block_reduce_stage<STAGE, TBX * TBY>(a.s, cdt, flags);
  }
}

// The stages' line force (lforce_cell) over the same (grid, tile) items.
template <int FEAT, int TBX, int TBY>
__global__ __launch_bounds__(TBX* TBY) void wind_batch_lforce_kernel(
    const StageArgs* __restrict__ args, const int* __restrict__ tile_off, int G, int total) {
  for (int item = blockIdx.x; item < total; item += gridDim.x) {
    const int g = __builtin_amdgcn_readfirstlane(find_item(tile_off, G, item));
    const StageArgs& a = args[g];
    if (a.s->halt) continue;           // a halted grid (uniform over the workgroup)
    const int tx_n = (a.nr + TBX - 1) / TBX;
    const int local = item - tile_off[g];
    const int ty = local / tx_n, tx = local - ty * tx_n;
    const int i = tx * TBX + threadIdx.x, j = ty * TBY + threadIdx.y;
    if (i < a.nr && j < a.nt) lforce_cell<FEAT>(a, i, j);
  }
}

// Per grid: fold the stage-1 C_dt into invDt (update_stage.c:389-394),
// t += dt, and NextTimeStep + reset when `next` (see wind_end_kernel).
__global__ __launch_bounds__(64) void wind_batch_end_kernel(const StageArgs* args, int G, int next,
                                                            const double* cfl,
                                                            const double* cfl_max_var) {
  static_assert(RED_SLOTS == 32, "one lane per spread slot and reduction");
  const int g = blockIdx.x;          // one workgroup per grid
  if (g >= G) return;
  Scalars* s = args[g].s;
  if (s->halt) return;               // uniform over the workgroup
  // lanes 0-31 fold the stage-1 C_dt slots (max), lanes 32-63 the cooling
  // dt slots (min), each clearing its slot; the halves reduce separately
  const int l = threadIdx.x;
  const bool cdt_half = l < RED_SLOTS;
  const int w = red_index(cdt_half ? l : l - RED_SLOTS);
  unsigned long long v = cdt_half ? s->cdt_red[w] : s->dtcool_red[w];
  if (cdt_half) s->cdt_red[w] = 0ull;
  else s->dtcool_red[w] = ~0ull;
  for (int off = RED_SLOTS / 2; off > 0; off >>= 1) {
    const unsigned long long u = __shfl_xor(v, off);
    v = cdt_half ? (u > v ? u : v) : (u < v ? u : v);
  }
  const unsigned long long cmax = __shfl(v, 0), cmin = __shfl(v, RED_SLOTS);
  if (l != 0) return;
  const unsigned long long cb = s->cdt_bits > cmax ? s->cdt_bits : cmax;
  s->cdt_bits = 0ull;
  double inv = s->inv_dt_hyp;
  inv = dmax(inv, __longlong_as_double((long long)cb));
  inv /= (double)2;
  s->inv_dt_hyp = inv;
  const double dt = s->dt;
  s->t += dt;
  if (cmin < s->dtcool_bits) s->dtcool_bits = cmin;
  const double dt_cool = __longlong_as_double((long long)s->dtcool_bits);
  s->dt_cool = dt_cool;
  if (next) {
    double dt_hyp = 1.0 / s->inv_dt_hyp;
    dt_hyp *= cfl[g];
    double dtn = dt_hyp;
    dtn = dmin(dtn, dt_cool);
    dtn = dmin(dtn, cfl_max_var[g] * dt);
    if (dtn < args[g].k.first_dt * 1.e-9) s->halt |= PMX_FLAG_DT_TOO_SMALL;
    if (cfl_max_var[g] != 1.0) s->dt = dtn;
    s->inv_dt_hyp = 0.0;
    s->dtcool_bits = (unsigned long long)__double_as_longlong(1.e38);
  }
  end_of_step(s);
}

}  // namespace wind
}  // namespace pmx

struct pmx_wind_batch {
  int G;
  int cooling;
  std::vector<pmx_wind*> grids;
  StageArgs* args1;       // device arrays, G entries: stage 1 (V0 -> V1), stage 2 (V1 -> V0)
  StageArgs* args2;
  int* tile_off;          // device, G + 1, for the current tile shape
  int total_tiles;
  double* cfl;            // device, G
  double* cfl_max_var;
  CoolArgs* cargs;        // device, G
  long long* cell_off;    // device, G + 1 (multiples of 128)
  long long total_cells;  // padded
  long long cells;        // interior cells of all grids
  int feat_any;           // OR of the grids' features (characteristic limiting, the entropy
                          // switch, the tabulated force multiplier, Roe)
  int shape;              // index into the tile-shape table
  int blocks;             // persistent workgroups per stage launch
  hipStream_t stream;
  hipGraphExec_t graph;
  bool have_graph;
  pmx_mlp* mlp;
  SurrWork sw;
  unsigned long long* calls;
  long step;
  bool timing;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pend_stage, pend_cool;
  double stage_ms, cool_ms;
  int stage_n, cool_n;
};

namespace {

struct TileShape { int bx, by; };
constexpr TileShape SHAPES[] = {{64, 2}, {32, 4}, {16, 8}, {128, 1}, {64, 4}, {32, 8}};
constexpr int NSHAPES = sizeof(SHAPES) / sizeof(SHAPES[0]);

template <int STAGE>
void launch_batch_stage(pmx_wind_batch* b) {
  const StageArgs* args = STAGE == 1 ? b->args1 : b->args2;
  const dim3 grid(b->blocks);
#define PMX_SHAPE(q, X, Y)                                                                   \
  case q:                                                                                    \
    if (b->feat_any)                                                                         \
      hipLaunchKernelGGL((wind_batch_lforce_kernel<FEAT_CHAR | FEAT_TAB, X, Y>), grid,       \
                         dim3(X, Y), 0, b->stream, args, b->tile_off, b->G, b->total_tiles); \
    else                                                                                     \
      hipLaunchKernelGGL((wind_batch_lforce_kernel<0, X, Y>), grid, dim3(X, Y), 0,           \
                         b->stream, args, b->tile_off, b->G, b->total_tiles);                \
    if (b->feat_any & FEAT_ROE)                                                              \
      hipLaunchKernelGGL((wind_batch_stage_kernel<STAGE, 15, X, Y>), grid, dim3(X, Y), 0,    \
                         b->stream, args, b->tile_off, b->G, b->total_tiles);                \
    else if (b->feat_any)                                                                    \
      hipLaunchKernelGGL((wind_batch_stage_kernel<STAGE, 7, X, Y>), grid, dim3(X, Y), 0,     \
                         b->stream, args, b->tile_off, b->G, b->total_tiles);                \
    else                                                                                     \
      hipLaunchKernelGGL((wind_batch_stage_kernel<STAGE, 0, X, Y>), grid, dim3(X, Y), 0,     \
                         b->stream, args, b->tile_off, b->G, b->total_tiles);                \
    break;
  switch (b->shape) {
    PMX_SHAPE(0, 64, 2)
    PMX_SHAPE(1, 32, 4)
    PMX_SHAPE(2, 16, 8)
    PMX_SHAPE(3, 128, 1)
    PMX_SHAPE(4, 64, 4)
    PMX_SHAPE(5, 32, 8)
  }
#undef PMX_SHAPE
}

int batch_set_shape(pmx_wind_batch* b, int shape, int blocks) {
  std::vector<int> off(b->G + 1, 0);
  for (int g = 0; g < b->G; g++) {
    const pmx_wind* w = b->grids[g];
    const int tx = (w->nr + SHAPES[shape].bx - 1) / SHAPES[shape].bx;
    const int ty = (w->nt + SHAPES[shape].by - 1) / SHAPES[shape].by;
    off[g + 1] = off[g] + tx * ty;
  }
  PMX_HIP_CHECK(hipMemcpy(b->tile_off, off.data(), sizeof(int) * (b->G + 1), hipMemcpyHostToDevice));
  b->total_tiles = off[b->G];
  b->shape = shape;
  b->blocks = blocks > 0 ? std::min(blocks, b->total_tiles) : b->total_tiles;
  if (b->have_graph) {
    hipGraphExecDestroy(b->graph);
    b->have_graph = false;
  }
  return 0;
}

int batch_timed(pmx_wind_batch* b, hipEvent_t* e, bool begin, bool cool) {
  if (!b->timing) return 0;
  if (begin) {
    if (b->ev_pool.size() < 2) {
      for (int q = 0; q < 64; q++) {
        hipEvent_t ev;
        PMX_HIP_CHECK(hipEventCreate(&ev));
        b->ev_pool.push_back(ev);
      }
    }
    e[0] = b->ev_pool.back(); b->ev_pool.pop_back();
    e[1] = b->ev_pool.back(); b->ev_pool.pop_back();
    PMX_HIP_CHECK(hipEventRecord(e[0], b->stream));
  } else {
    PMX_HIP_CHECK(hipEventRecord(e[1], b->stream));
    (cool ? b->pend_cool : b->pend_stage).push_back({e[0], e[1]});
  }
  return 0;
}

int batch_advance(pmx_wind_batch* b) {
  hipEvent_t e[2];
  for (pmx_wind* w : b->grids)
    if (flag_shock(w, b->stream, true)) return -1;
  if (batch_timed(b, e, true, false)) return -1;
  launch_batch_stage<1>(b);
  PMX_HIP_CHECK(hipGetLastError());
  if (batch_timed(b, e, false, false)) return -1;
  if (batch_timed(b, e, true, false)) return -1;
  launch_batch_stage<2>(b);
  PMX_HIP_CHECK(hipGetLastError());
  return batch_timed(b, e, false, false);
}

int batch_cool(pmx_wind_batch* b) {
  if (b->cooling == PMX_COOL_NONE) return 0;
  CoolBatch cb;
  cb.args = b->cargs; cb.off = b->cell_off; cb.G = b->G; cb.total = b->total_cells;
  cb.loop = 1;
  hipEvent_t e[2];
  if (batch_timed(b, e, true, true)) return -1;
  if (b->mlp) {
    if (mlp::surrogate_cool(b->mlp, cb, b->sw, b->stream)) return -1;
  } else if (mlp::plain_cool(cb, b->calls, b->stream)) {
    return -1;
  }
  return batch_timed(b, e, false, true);
}

// One main-loop iteration of every grid (the same step parity for all).
int batch_step(pmx_wind_batch* b, long st) {
  if (st % 2 == 0) {
    if (batch_advance(b) || batch_cool(b)) return -1;
  } else {
    if (batch_cool(b) || batch_advance(b)) return -1;
  }
  const int next = (b->cooling == PMX_COOL_NONE) || (st % 2 == 1);
  hipLaunchKernelGGL(wind_batch_end_kernel, dim3(b->G), dim3(64), 0, b->stream,
                     b->args1, b->G, next, b->cfl, b->cfl_max_var);
  PMX_HIP_CHECK(hipGetLastError());
  return 0;
}

}  // namespace

extern "C" {

int pmx_wind_batch_destroy(pmx_wind_batch* b) {
  if (!b) return 0;
  if (b->stream) hipStreamSynchronize(b->stream);
  if (b->have_graph) hipGraphExecDestroy(b->graph);
  for (auto e : b->ev_pool) hipEventDestroy(e);
  for (auto& p : b->pend_stage) { hipEventDestroy(p.first); hipEventDestroy(p.second); }
  for (auto& p : b->pend_cool) { hipEventDestroy(p.first); hipEventDestroy(p.second); }
  for (auto w : b->grids) pmx_wind_destroy(w);
  hipFree(b->args1); hipFree(b->args2); hipFree(b->tile_off); hipFree(b->cfl);
  hipFree(b->cfl_max_var); hipFree(b->cargs); hipFree(b->cell_off); hipFree(b->calls);
  hipFree(b->sw.list); hipFree(b->sw.fb); hipFree(b->sw.feat);
  hipFree(b->sw.f0); hipFree(b->sw.cnt);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return 0;
}

int pmx_wind_batch_create(const pmx_wind_desc* descs, int ngrids, pmx_wind_batch** out) {
  if (!descs || ngrids <= 0 || !out) { set_error("bad argument"); return -1; }
  pmx_wind_batch* b = new pmx_wind_batch();
  b->G = ngrids;
  b->cooling = descs[0].cooling;
  for (int g = 0; g < ngrids; g++) {
    if (descs[g].cooling != b->cooling || descs[g].cooling == PMX_COOL_BLONDIN) {
      set_error("batched grids share one cooling mode: PMX_COOL_NONE or PMX_COOL_PLAIN");
      pmx_wind_batch_destroy(b);
      return -1;
    }
    pmx_wind* w = nullptr;
    if (pmx_wind_create(&descs[g], &w)) { pmx_wind_batch_destroy(b); return -1; }
    b->grids.push_back(w);
  }
#define CK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { \
    set_error(std::string(#x) + ": " + hipGetErrorString(e_)); pmx_wind_batch_destroy(b); return -1; } } while (0)
  CK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  const int G = ngrids;
  std::vector<StageArgs> a1(G), a2(G);
  std::vector<CoolArgs> ca(G);
  std::vector<double> cfl(G), cmv(G);
  std::vector<long long> coff(G + 1, 0);
  b->cells = 0;
  for (int g = 0; g < G; g++) {
    pmx_wind* w = b->grids[g];
    a1[g] = stage_args(w, w->V[0], w->V[1], true);
    a2[g] = stage_args(w, w->V[1], w->V[0], true);
    ca[g] = cool_args(w);
    cfl[g] = w->cfl;
    cmv[g] = w->cfl_max_var;
    coff[g + 1] = coff[g] + (w->sv + 127) / 128 * 128;
    b->cells += w->sv;
    b->feat_any |= features(w);
  }
  b->total_cells = coff[G];
  CK(hipMalloc(&b->args1, sizeof(StageArgs) * G));
  CK(hipMalloc(&b->args2, sizeof(StageArgs) * G));
  CK(hipMalloc(&b->tile_off, sizeof(int) * (G + 1)));
  CK(hipMalloc(&b->cfl, sizeof(double) * G));
  CK(hipMalloc(&b->cfl_max_var, sizeof(double) * G));
  CK(hipMalloc(&b->cargs, sizeof(CoolArgs) * G));
  CK(hipMalloc(&b->cell_off, sizeof(long long) * (G + 1)));
  CK(hipMalloc(&b->calls, sizeof(unsigned long long) * RED_SLOTS * RED_STRIDE));
  CK(hipMemset(b->calls, 0, sizeof(unsigned long long) * RED_SLOTS * RED_STRIDE));
  CK(hipMemcpy(b->args1, a1.data(), sizeof(StageArgs) * G, hipMemcpyHostToDevice));
  CK(hipMemcpy(b->args2, a2.data(), sizeof(StageArgs) * G, hipMemcpyHostToDevice));
  CK(hipMemcpy(b->cfl, cfl.data(), sizeof(double) * G, hipMemcpyHostToDevice));
  CK(hipMemcpy(b->cfl_max_var, cmv.data(), sizeof(double) * G, hipMemcpyHostToDevice));
  CK(hipMemcpy(b->cargs, ca.data(), sizeof(CoolArgs) * G, hipMemcpyHostToDevice));
  CK(hipMemcpy(b->cell_off, coff.data(), sizeof(long long) * (G + 1), hipMemcpyHostToDevice));
#undef CK
  if (batch_set_shape(b, 0, 0)) { pmx_wind_batch_destroy(b); return -1; }
  *out = b;
  return 0;
}

int pmx_wind_batch_grid(pmx_wind_batch* b, int g, pmx_wind** w) {
  if (!b || g < 0 || g >= b->G || !w) { set_error("bad argument"); return -1; }
  *w = b->grids[g];
  return 0;
}

int pmx_wind_batch_set_step(pmx_wind_batch* b, long step) {
  if (!b) { set_error("null argument"); return -1; }
  b->step = step;
  return 0;
}

int pmx_wind_batch_set_surrogate(pmx_wind_batch* b, pmx_mlp* m, int guess) {
  if (!b) { set_error("null argument"); return -1; }
  if (m && b->cooling != PMX_COOL_PLAIN) { set_error("the surrogate needs PMX_COOL_PLAIN"); return -1; }
  if (m && !b->sw.list) {
    const size_t n = (size_t)b->total_cells;
    PMX_HIP_CHECK(hipMalloc(&b->sw.list, sizeof(long long) * n));
    PMX_HIP_CHECK(hipMalloc(&b->sw.fb, sizeof(long long) * n));
    PMX_HIP_CHECK(hipMalloc(&b->sw.feat, sizeof(float) * 4 * n));
    PMX_HIP_CHECK(hipMalloc(&b->sw.f0, sizeof(double) * n));
    PMX_HIP_CHECK(hipMalloc(&b->sw.cnt, sizeof(unsigned long long) * PMX_SURR_COUNTERS));
    PMX_HIP_CHECK(hipMemset(b->sw.cnt, 0, sizeof(unsigned long long) * PMX_SURR_COUNTERS));
  }
  b->mlp = m;
  b->sw.guess = guess;
  if (b->have_graph) { hipGraphExecDestroy(b->graph); b->have_graph = false; }
  return 0;
}

// Times stage-1 launches (they read V^n and write only V^(1), U and the
// step reductions, which are saved and restored) for every tile shape and
// persistent workgroup count, and keeps the fastest.
int pmx_wind_batch_autotune(pmx_wind_batch* b, int* shape_out, int* blocks_out, double* ms_out) {
  if (!b) { set_error("null argument"); return -1; }
  int dev = 0, ncu = 0;
  PMX_HIP_CHECK(hipGetDevice(&dev));
  PMX_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  std::vector<Scalars> saved(b->G);
  for (int g = 0; g < b->G; g++)
    PMX_HIP_CHECK(hipMemcpy(&saved[g], b->grids[g]->s, sizeof(Scalars), hipMemcpyDeviceToHost));
  hipEvent_t e0, e1;
  PMX_HIP_CHECK(hipEventCreate(&e0));
  PMX_HIP_CHECK(hipEventCreate(&e1));
  double best = 1e300;
  int bs = 0, bb = 0;
  const int mult[] = {0, 1, 2, 4, 8};   // 0: one workgroup per tile
  for (int sh = 0; sh < NSHAPES; sh++) {
    for (int mq : mult) {
      if (batch_set_shape(b, sh, mq * ncu)) return -1;
      launch_batch_stage<1>(b);                 // warm-up
      PMX_HIP_CHECK(hipEventRecord(e0, b->stream));
      for (int r = 0; r < 3; r++) launch_batch_stage<1>(b);
      PMX_HIP_CHECK(hipEventRecord(e1, b->stream));
      PMX_HIP_CHECK(hipGetLastError());
      PMX_HIP_CHECK(hipEventSynchronize(e1));
      float ms = 0.f;
      PMX_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) { best = ms; bs = sh; bb = mq * ncu; }
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  for (int g = 0; g < b->G; g++)
    PMX_HIP_CHECK(hipMemcpy(b->grids[g]->s, &saved[g], sizeof(Scalars), hipMemcpyHostToDevice));
  if (batch_set_shape(b, bs, bb)) return -1;
  if (shape_out) *shape_out = bs;
  if (blocks_out) *blocks_out = b->blocks;
  if (ms_out) *ms_out = best / 3.0;
  return 0;
}

int pmx_wind_batch_set_shape(pmx_wind_batch* b, int shape, int blocks) {
  if (!b || shape < 0 || shape >= NSHAPES) { set_error("bad argument"); return -1; }
  return batch_set_shape(b, shape, blocks);
}

int pmx_wind_batch_run(pmx_wind_batch* b, int nsteps, int use_graph) {
  if (!b || nsteps < 0) { set_error("bad argument"); return -1; }
  int n = 0;
  if (nsteps > 0)
    for (auto* g : b->grids) g->stepped = true;
  if (use_graph && !b->timing) {
    // a captured pair of steps (even, odd) replayed
    if (b->step % 2 == 1 && n < nsteps) {
      if (batch_step(b, b->step)) return -1;
      b->step++;
      n++;
    }
    if (nsteps - n >= 2 && !b->have_graph) {
      hipGraph_t gph;
      PMX_HIP_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      int rc = batch_step(b, 0) | batch_step(b, 1);
      hipError_t e = hipStreamEndCapture(b->stream, &gph);
      if (rc || e != hipSuccess) { set_error("graph capture failed"); return -1; }
      PMX_HIP_CHECK(hipGraphInstantiate(&b->graph, gph, nullptr, nullptr, 0));
      PMX_HIP_CHECK(hipGraphDestroy(gph));
      b->have_graph = true;
    }
    for (; nsteps - n >= 2; n += 2) {
      PMX_HIP_CHECK(hipGraphLaunch(b->graph, b->stream));
      b->step += 2;
    }
  }
  for (; n < nsteps; n++) {
    if (batch_step(b, b->step)) return -1;
    b->step++;
  }
  return 0;
}

int pmx_wind_batch_sync(pmx_wind_batch* b) {
  if (!b) { set_error("null argument"); return -1; }
  PMX_HIP_CHECK(hipStreamSynchronize(b->stream));
  return 0;
}

int pmx_wind_batch_timing(pmx_wind_batch* b, int enable, int reset, double* stage_ms,
                          int* stage_launches, double* cool_ms, int* cool_launches) {
  if (!b) { set_error("null argument"); return -1; }
  for (int c = 0; c < 2; c++) {
    auto& v = c ? b->pend_cool : b->pend_stage;
    for (auto& p : v) {
      float ms = 0.f;
      PMX_HIP_CHECK(hipEventSynchronize(p.second));
      PMX_HIP_CHECK(hipEventElapsedTime(&ms, p.first, p.second));
      if (c) { b->cool_ms += ms; b->cool_n++; } else { b->stage_ms += ms; b->stage_n++; }
      b->ev_pool.push_back(p.first);
      b->ev_pool.push_back(p.second);
    }
    v.clear();
  }
  if (stage_ms) *stage_ms = b->stage_ms;
  if (stage_launches) *stage_launches = b->stage_n;
  if (cool_ms) *cool_ms = b->cool_ms;
  if (cool_launches) *cool_launches = b->cool_n;
  if (reset) { b->stage_ms = b->cool_ms = 0.0; b->stage_n = b->cool_n = 0; }
  b->timing = enable != 0;
  return 0;
}

int pmx_wind_batch_cells(pmx_wind_batch* b, long long* cells) {
  if (!b || !cells) { set_error("bad argument"); return -1; }
  *cells = b->cells;
  return 0;
}

int pmx_check_div_rc(const double* x, long long n, double c, long long* mismatches) {
  if (!x || n <= 0 || !(c > 0.0) || !mismatches) { set_error("bad argument"); return -1; }
  double* d = nullptr;
  unsigned long long* bad = nullptr;
  PMX_HIP_CHECK(hipMalloc(&d, sizeof(double) * n));
  if (hipMalloc(&bad, sizeof(unsigned long long)) != hipSuccess) {
    hipFree(d);
    set_error("hipMalloc");
    return -1;
  }
  unsigned long long h = 0;
  hipError_t e = hipMemcpy(d, x, sizeof(double) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(bad, &h, sizeof(h), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    const long long nb = std::min<long long>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(pmx::wind::div_check_kernel, dim3((unsigned)nb), dim3(256), 0, 0, d, n, c,
                       1.0 / c, bad);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(&h, bad, sizeof(h), hipMemcpyDeviceToHost);
  hipFree(d);
  hipFree(bad);
  if (e != hipSuccess) { set_error(hipGetErrorString(e)); return -1; }
  *mismatches = (long long)h;
  return 0;
}

int pmx_check_force_power(const double* num, const double* dvds, long long n, double krad,
                          double alpha, double* out) {
  if (!num || !dvds || !out || n <= 0) { set_error("bad argument"); return -1; }
  double* d = nullptr;
  PMX_HIP_CHECK(hipMalloc(&d, sizeof(double) * 3 * n));
  pmx::wind::Const k{};
  k.krad = krad;
  k.alpharad = alpha;
  hipError_t e = hipMemcpy(d, num, sizeof(double) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d + n, dvds, sizeof(double) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    const long long nb = std::min<long long>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(pmx::wind::power_check_kernel, dim3((unsigned)nb), dim3(256), 0, 0, d,
                       d + n, n, k, d + 2 * n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(out, d + 2 * n, sizeof(double) * n, hipMemcpyDeviceToHost);
  hipFree(d);
  if (e != hipSuccess) { set_error(hipGetErrorString(e)); return -1; }
  return 0;
}

}  // extern "C"
