// Fused method-of-lines stage kernel for HD on uniform Cartesian grids.
//
// One launch performs one RK stage of the reference's UpdateStage
// (Src/Time_Stepping/update_stage.c:35-410) for every direction at once,
// followed by the RK combination (rk_step.c:214-237, 281-305) and the
// cons->prim map (mappers3D.c:16-141):
//
//   U_s = U_in + rhs_x(V) + rhs_y(V) + rhs_z(V)        (added in that order)
//   U   = combine(U0, U_s);   V_out = C2P(U)            (U floored in place)
//
// Work decomposition (2.5-D): a workgroup owns a TX x TY tile of the x-y
// plane and marches along z over a chunk of CZ planes.
//   * z: each thread keeps a register window of its column, the PPM
//     interface value and left state of the previous cell, and the flux of
//     the previous z-face, so every z-face and every z-state is computed
//     exactly once.
//   * x, y: the plane (with H ghost cells in x and y) is staged in LDS;
//     per-cell left/right states are computed once into LDS, then the
//     face fluxes, then the update.  Only the one-cell ring around the
//     tile is computed twice (by the neighbouring tile).
// Ghost cells are never stored: halo reads map out-of-domain indices with
// the boundary rule (map_bc), so no separate boundary kernel runs.
//
// Layout in HBM: V, U are SoA [nv][nz][ny][nx] float64, interior only.
#pragma once
#include "pmx_common.h"
#include "pmx_hd_math.h"

namespace pmx {

enum { MODE_FIRST = 0, MODE_WEIGHTED = 1, MODE_THIRD = 2 };

#ifndef PMX_WAVES_PER_EU
#define PMX_WAVES_PER_EU 2   // occupancy the register allocation targets: 2 fits the stage
                             // without spills (~215 VGPRs) and measured 3% faster than 3
#endif
#ifndef PMX_MERGE_XY
#define PMX_MERGE_XY 1     // x and y sweeps share phases, 4 barriers per plane instead of 6
#endif
#ifndef PMX_ZUNROLL
#define PMX_ZUNROLL 1   // unroll of the z-plane loop
#endif
#ifndef PMX_LDS_PAD
#define PMX_LDS_PAD 0   // extra doubles per LDS plane row
#endif
#ifndef PMX_PREFETCH_U
#define PMX_PREFETCH_U 0   // load U and V0 of the cell at the top of the plane loop
#endif

#ifndef PMX_MAX_ZCHUNKS
#define PMX_MAX_ZCHUNKS 48
#endif

// Tile of the x-y plane per workgroup: every per-plane phase (x-states,
// x-faces, y-states, y-faces) has at most 256 items, so each is a single
// pass of the 256-thread workgroup.
#ifndef PMX_TILE_X
#define PMX_TILE_X 16
#endif
#ifndef PMX_TILE_Y
#define PMX_TILE_Y 14
#endif
constexpr int TILE_X = PMX_TILE_X;
constexpr int TILE_Y = PMX_TILE_Y;
// 2-D grids take shorter tiles: KH 256^2 makes 512 tiles of 16 x 8, two
// workgroups on every CU, where 16 x 14 made 304, two on 48 CUs and one on
// the rest (21.6-22.6 vs 22.9-23.8 us per step, bit-identical)
#ifndef PMX_TILE_Y_2D
#define PMX_TILE_Y_2D 8
#endif
template <int DIMS>
constexpr int tile_y() { return DIMS == 3 ? TILE_Y : PMX_TILE_Y_2D; }
constexpr int NTHREADS = 256;

struct StageArgs {
  const double* Vin;   // stage input, halo source
  const double* V0;    // V at the start of the step (U0 = P2C(V0)); may alias Vout
  double* U;           // U of the previous stage (in) / this stage (out)
  double* Vout;        // V after this stage
  int nx, ny, nz;
  long long sv;        // stride between variables = nx*ny*nz
  int bc[6];
  double gamma, gmm1;
  const double* dtdx[3];    // dt/dx per cell and direction (this step)
  const double* inv_dl[3];  // 1/dx per cell and direction
  double w0, wc;            // RK weights for MODE_WEIGHTED
  int cz;                   // z planes per workgroup (uniform chunks, nzc == 0)
  int tiles_x, tiles_y, tiles_z;
  // z chunks of unequal length (3-D, nzc > 0): chunk c is planes
  // [zc0[c], zc0[c+1]); each XCD's workgroups take its tiles chunk by
  // chunk, so the long chunks are dispatched first and the short ones
  // fill the end of the launch
  int nzc;
  int zc0[PMX_MAX_ZCHUNKS + 1];
  int write_u;              // store U (always needed except after the last stage)
  unsigned long long* cdt_bits;  // MODE_FIRST: max over cells of C_dt (bit pattern), spread
                                 // over the words red_index(k) by workgroup
  int* err;                      // OR of cons->prim floor flags
  const int* halt;               // main-loop steps: skip the stage once nonzero; null for a
                                 // step the caller drives (explicit dt)
};

template <int REC> struct RecTraits {
  static constexpr int H = (REC == REC_PPM) ? 3 : 2;  // ghost width, Src/get_nghost.c:19-42
};

// Boundary kinds known at compile time.  The stage kernel sits at its
// register limit, and the run-time boundary switch of map_bc costs it
// about 40% more spilled registers; a domain with the same rule on every
// side takes a specialised load instead.
enum { BCK_ANY = 0, BCK_OUTFLOW = 1, BCK_PERIODIC = 2 };

template <int BCK>
__device__ __forceinline__ int map_bck(int g, int n, int bcl, int bcr, bool& flip) {
  if (BCK == BCK_OUTFLOW) return min(max(g, 0), n - 1);
  if (BCK == BCK_PERIODIC) {
    g += g < 0 ? n : 0;
    g -= g >= n ? n : 0;
    return min(max(g, 0), n - 1);
  }
  return map_bc(g, n, bcl, bcr, flip);
}

// Load the 5 primitive variables of global cell (i, j, k) with boundary
// mapping; flip the sign of the normal velocity of a reflected direction.
template <int DIMS, int BCK>
__device__ __forceinline__ void load_cell(const StageArgs& a, const double* __restrict__ V,
                                          int i, int j, int k, double* q) {
  bool fx = false, fy = false, fz = false;
  int ii = map_bck<BCK>(i, a.nx, a.bc[0], a.bc[1], fx);
  int jj = map_bck<BCK>(j, a.ny, a.bc[2], a.bc[3], fy);
  int kk = (DIMS == 3) ? map_bck<BCK>(k, a.nz, a.bc[4], a.bc[5], fz) : 0;
  long long o = ((long long)kk * a.ny + jj) * a.nx + ii;
#pragma unroll
  for (int nv = 0; nv < NV; nv++) q[nv] = V[o + nv * a.sv];
  if (fx) q[VX1] = -q[VX1];
  if (fy) q[VX2] = -q[VX2];
  if (fz) q[VX3] = -q[VX3];
}

// Per-cell states from a 5-point (PPM) or 3-point (PLM) stencil s[0..4]
// centred at s[2]; vp = right-face state, vm = left-face state.
template <int REC>
__device__ __forceinline__ void cell_states(const double (*s)[NV], double* vp, double* vm) {
  if (REC == REC_PPM) {
#pragma unroll
    for (int nv = 0; nv < NV; nv++) {
      double Ip = ppm_iface(s[1][nv], s[2][nv], s[3][nv], s[4][nv]);
      double Im = ppm_iface(s[0][nv], s[1][nv], s[2][nv], s[3][nv]);
      ppm_limit(s[2][nv], Ip, Im, vp[nv], vm[nv]);
    }
  } else {
    plm_states<REC>(s[1], s[2], s[3], vp, vm);
  }
}

// Map halo item `it` of a tile plane to LDS coordinates (r, c): first the
// x-halo (H columns left and right of every tile row), then the y-halo
// (H rows below and above every tile column).
template <int H, int TX, int TY>
__device__ __forceinline__ void halo_item(int it, int& r, int& c) {
  constexpr int NHX = 2 * H * TY;
  if (it < NHX) {
    int cc = it % (2 * H);
    r = H + it / (2 * H);
    c = cc < H ? cc : TX + cc;
  } else {
    int jt = it - NHX;
    int rr = jt / TX;
    c = H + jt % TX;
    r = rr < H ? rr : TY + rr;
  }
}

template <int DIMS, int REC, int SOLVER, int MODE, int BCK, int TX, int TY, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(PMX_WAVES_PER_EU, PMX_WAVES_PER_EU))) void stage_kernel(StageArgs a) {
  constexpr int H = RecTraits<REC>::H;
  constexpr int NC = TX * TY;          // cells of a tile plane
  constexpr int LWR = TX + 2 * H;                // plane row width
  constexpr int LW = LWR + PMX_LDS_PAD;           // row stride in LDS
  constexpr int LH = TY + 2 * H;
  constexpr int NSX = (TX + 2) * TY;   // x-states per tile plane
  constexpr int NSY = TX * (TY + 2);   // y-states
  constexpr int NFX = (TX + 1) * TY;   // x-faces
  constexpr int NFY = TX * (TY + 1);   // y-faces
  constexpr int NST = NSX > NSY ? NSX : NSY;
  constexpr int NPL = NV * LH * LW;
  constexpr int NFYD = 7 * NFY;
  constexpr int NPLANE = NPL > NFYD ? NPL : NFYD;
  constexpr int NHALO = 2 * H * (TX + TY);
  constexpr int HPT = (NHALO + NT - 1) / NT;  // halo cells loaded per thread
  static_assert(NC <= NT, "one thread per tile cell");

  // LDS arrays are component-major so that consecutive lanes touch
  // consecutive doubles (no bank conflicts).
  __shared__ double sPlane[NPLANE];    // V plane [nv][LH][LW]; later y-fluxes [7][NFY]
  // states [vp(5) vm(5)][item]; x and y each have their own when the two
  // directions' phases share barriers (72 KB, still 2 workgroups per CU;
  // in 2-D, KH 256^2, 2% faster than separate phases with less LDS)
  constexpr bool MERGE = PMX_MERGE_XY;
  __shared__ double sSt[(MERGE ? 2 : 1) * NST * 10];
  double* const sStX = sSt;
  double* const sStY = sSt + (MERGE ? NST * 10 : 0);
  __shared__ double sFx[NFX * 7];      // x-fluxes [flux(5) press cmax][face]
  __shared__ double sRed[NT / 64];

  if (a.halt && *a.halt) return;       // the run has stopped (uniform over the launch)

  const int tid = threadIdx.x;
  const bool owner = tid < NC;         // thread owns tile cell (tx, ty)
  const int tx = tid % TX, ty = tid / TX;
  const int ntile = a.tiles_x * a.tiles_y;
  const bool zsched = DIMS == 3 && a.nzc > 0;
  int tile, bz;
  if (zsched) {
    // Workgroup b runs on XCD b % 8 (round-robin dispatch): each XCD owns
    // a contiguous eighth of the tiles, so neighbouring tiles share its L2,
    // and walks them chunk by chunk, long chunks first.  The launch has
    // 8 ceil(ntile / 8) nzc workgroups; the few past an XCD's share end here.
    const int x = blockIdx.x % 8, r = blockIdx.x / 8;
    const int t0 = x * ntile / 8, nt = (x + 1) * ntile / 8 - t0;
    if (r >= nt * a.nzc) return;
    bz = r / nt;
    tile = t0 + r % nt;
  } else {
    const int lb = xcd_remap(blockIdx.x, ntile * a.tiles_z);
    bz = lb / ntile;
    tile = lb % ntile;
  }
  tile = __builtin_amdgcn_readfirstlane(tile);
  bz = __builtin_amdgcn_readfirstlane(bz);
  const int bx = tile % a.tiles_x;
  const int by = tile / a.tiles_x;
  const int i0 = bx * TX, j0 = by * TY;
  const int i = i0 + tx, j = j0 + ty;
  const bool valid = owner && (i < a.nx) && (j < a.ny);
  int k0 = 0, k1 = 1;
  if (DIMS == 3) {
    if (zsched) {
      k0 = a.zc0[bz];
      k1 = a.zc0[bz + 1];
    } else {
      k0 = bz * a.cz;
      k1 = min(k0 + a.cz, a.nz);
    }
    k0 = __builtin_amdgcn_readfirstlane(k0);
    k1 = __builtin_amdgcn_readfirstlane(k1);
  }
  const double gamma = a.gamma, gmm1 = a.gmm1;

  double cdt_max = 0.0;
  int fail = 0;

  // ---- z-direction register state -----------------------------------
  // PPM window w[0..3] = v[k..k+3]; PLM window w[0..2] = v[k..k+2]
  constexpr int WZ = (REC == REC_PPM) ? 4 : 3;
  double w[WZ][NV];
  double Iz[NV];          // PPM interface value at k+1/2
  double vpz[NV];         // right-face state of cell k
  double Fz[7];           // flux at k-1/2 (5), press, cmax

  if (DIMS == 3 && owner) {
    // prologue: states of cells k0-1 and k0, flux at k0-1/2
    if (REC == REC_PPM) {
      double p[6][NV];  // v[k0-3 .. k0+2]
#pragma unroll
      for (int s = 0; s < 6; s++) load_cell<DIMS, BCK>(a, a.Vin, i, j, k0 - 3 + s, p[s]);
      double vm0[NV], vpm[NV];
#pragma unroll
      for (int nv = 0; nv < NV; nv++) {
        double Ia = ppm_iface(p[0][nv], p[1][nv], p[2][nv], p[3][nv]);  // I(k0-3/2)
        double Ib = ppm_iface(p[1][nv], p[2][nv], p[3][nv], p[4][nv]);  // I(k0-1/2)
        Iz[nv] = ppm_iface(p[2][nv], p[3][nv], p[4][nv], p[5][nv]);     // I(k0+1/2)
        double vmm;
        ppm_limit(p[2][nv], Ib, Ia, vpm[nv], vmm);                       // cell k0-1
        ppm_limit(p[3][nv], Iz[nv], Ib, vpz[nv], vm0[nv]);               // cell k0
      }
      riemann<SOLVER, 2, DIMS>(vpm, vm0, gamma, gmm1, Fz, Fz[5], Fz[6]);
#pragma unroll
      for (int s = 0; s < 3; s++)
#pragma unroll
        for (int nv = 0; nv < NV; nv++) w[s][nv] = p[s + 3][nv];  // v[k0..k0+2]
    } else {
      double p[4][NV];  // v[k0-2 .. k0+1]
#pragma unroll
      for (int s = 0; s < 4; s++) load_cell<DIMS, BCK>(a, a.Vin, i, j, k0 - 2 + s, p[s]);
      double vpm[NV], vmm[NV], vm0[NV];
      plm_states<REC>(p[0], p[1], p[2], vpm, vmm);   // cell k0-1
      plm_states<REC>(p[1], p[2], p[3], vpz, vm0);   // cell k0
      riemann<SOLVER, 2, DIMS>(vpm, vm0, gamma, gmm1, Fz, Fz[5], Fz[6]);
#pragma unroll
      for (int s = 0; s < 2; s++)
#pragma unroll
        for (int nv = 0; nv < NV; nv++) w[s][nv] = p[s + 2][nv];  // v[k0..k0+1]
    }
  }

  // halo of the first plane; later planes are prefetched one plane ahead
  double hq[HPT][NV];
#pragma unroll
  for (int q = 0; q < HPT; q++) {
    int it = tid + q * NT, r, c;
    if (it < NHALO) {
      halo_item<H, TX, TY>(it, r, c);
      load_cell<DIMS, BCK>(a, a.Vin, i0 + c - H, j0 + r - H, k0, hq[q]);
    }
  }

#pragma unroll PMX_ZUNROLL
  for (int k = k0; k < k1; k++) {
    double vc[NV];  // own cell value at plane k
    if (owner) {
      if (DIMS == 3) {
        load_cell<DIMS, BCK>(a, a.Vin, i, j, k + WZ - 1, w[WZ - 1]);
#pragma unroll
        for (int nv = 0; nv < NV; nv++) vc[nv] = w[0][nv];
      } else {
        load_cell<DIMS, BCK>(a, a.Vin, i, j, 0, vc);
      }
    }
    // U of this stage and V at the start of the step, needed by the update
    double u[NV], v0[NV];
    const long long o = ((long long)k * a.ny + (valid ? j : 0)) * a.nx + (valid ? i : 0);
    if (PMX_PREFETCH_U && MODE != MODE_FIRST && valid) {
#pragma unroll
      for (int nv = 0; nv < NV; nv++) {
        u[nv] = a.U[o + nv * a.sv];
        v0[nv] = a.V0[o + nv * a.sv];
      }
    }

    // ---- P0: stage the plane with its x/y halo in LDS ----------------
    if (owner) {
#pragma unroll
      for (int nv = 0; nv < NV; nv++) sPlane[(nv * LH + H + ty) * LW + H + tx] = vc[nv];
    }
#pragma unroll
    for (int q = 0; q < HPT; q++) {
      int it = tid + q * NT, r, c;
      if (it < NHALO) {
        halo_item<H, TX, TY>(it, r, c);
#pragma unroll
        for (int nv = 0; nv < NV; nv++) sPlane[(nv * LH + r) * LW + c] = hq[q][nv];
        if (k + 1 < k1) load_cell<DIMS, BCK>(a, a.Vin, i0 + c - H, j0 + r - H, k + 1, hq[q]);
      }
    }
    __syncthreads();

    // x-states of cells -1 .. TX (per row)
    auto x_state = [&](int it) {
      int r = it / (TX + 2), c = it % (TX + 2) - 1;
      double s[5][NV];
#pragma unroll
      for (int q = 0; q < 5; q++) {
        int cc = H + c - 2 + q;
        cc = cc < 0 ? 0 : (cc >= LWR ? LWR - 1 : cc);  // PLM (H=2) never reads the outer pair
#pragma unroll
        for (int nv = 0; nv < NV; nv++) s[q][nv] = sPlane[(nv * LH + H + r) * LW + cc];
      }
      double vp[NV], vm[NV];
      cell_states<REC>(s, vp, vm);
#pragma unroll
      for (int nv = 0; nv < NV; nv++) {
        sStX[nv * NST + it] = vp[nv];
        sStX[(5 + nv) * NST + it] = vm[nv];
      }
    };
    // x-faces
    auto x_face = [&](int f) {
      int r = f / (TX + 1), fi = f % (TX + 1);  // face between cells fi-1 and fi
      int sl = r * (TX + 2) + fi, sr = sl + 1;
      double vL[NV], vR[NV], F[7];
#pragma unroll
      for (int nv = 0; nv < NV; nv++) {
        vL[nv] = sStX[nv * NST + sl];
        vR[nv] = sStX[(5 + nv) * NST + sr];
      }
      riemann<SOLVER, 0, DIMS>(vL, vR, gamma, gmm1, F, F[5], F[6]);
#pragma unroll
      for (int q = 0; q < 7; q++) sFx[q * NFX + f] = F[q];
    };
    // y-states of rows -1 .. TY (per column)
    auto y_state = [&](int it) {
      int c = it % TX, r = it / TX - 1;
      double s[5][NV];
#pragma unroll
      for (int q = 0; q < 5; q++) {
        int rr = H + r - 2 + q;
        rr = rr < 0 ? 0 : (rr >= LH ? LH - 1 : rr);
#pragma unroll
        for (int nv = 0; nv < NV; nv++) s[q][nv] = sPlane[(nv * LH + rr) * LW + H + c];
      }
      double vp[NV], vm[NV];
      cell_states<REC>(s, vp, vm);
#pragma unroll
      for (int nv = 0; nv < NV; nv++) {
        sStY[nv * NST + it] = vp[nv];
        sStY[(5 + nv) * NST + it] = vm[nv];
      }
    };
    // y-faces (the fluxes overwrite the plane buffer, whose reads are done)
    auto y_face = [&](int f) {
      int c = f % TX, fj = f / TX;  // face between rows fj-1 and fj
      int sl = fj * TX + c, sr = sl + TX;
      double vL[NV], vR[NV], F[7];
#pragma unroll
      for (int nv = 0; nv < NV; nv++) {
        vL[nv] = sStY[nv * NST + sl];
        vR[nv] = sStY[(5 + nv) * NST + sr];
      }
      riemann<SOLVER, 1, DIMS>(vL, vR, gamma, gmm1, F, F[5], F[6]);
#pragma unroll
      for (int q = 0; q < 7; q++) sPlane[q * NFY + f] = F[q];
    };

    if constexpr (MERGE) {
    // ---- P1: x- and y-states; P2: x- and y-faces (two barriers) -------
    for (int it = tid; it < NST; it += NT) {
      if (it < NSX) x_state(it);
      if (it < NSY) y_state(it);
    }
    __syncthreads();
    for (int f = tid; f < (NFX > NFY ? NFX : NFY); f += NT) {
      if (f < NFX) x_face(f);
      if (f < NFY) y_face(f);
    }
    __syncthreads();
    } else {
    // ---- P1: x-states; P2: x-faces; P3: y-states; P4: y-faces ----------
    for (int it = tid; it < NSX; it += NT) x_state(it);
    __syncthreads();
    for (int f = tid; f < NFX; f += NT) x_face(f);
    __syncthreads();
    for (int it = tid; it < NSY; it += NT) y_state(it);
    __syncthreads();
    for (int f = tid; f < NFY; f += NT) y_face(f);
    __syncthreads();
    }

    // ---- P5: z-face, update, RK combination, cons->prim ---------------
    if (owner) {
      double Fzn[7];
      if (DIMS == 3) {
        double vmn[NV], vpn[NV];
        if (REC == REC_PPM) {
#pragma unroll
          for (int nv = 0; nv < NV; nv++) {
            double In = ppm_iface(w[0][nv], w[1][nv], w[2][nv], w[3][nv]);  // I(k+3/2)
            ppm_limit(w[1][nv], In, Iz[nv], vpn[nv], vmn[nv]);             // cell k+1
            Iz[nv] = In;
          }
        } else {
          plm_states<REC>(w[0], w[1], w[2], vpn, vmn);                     // cell k+1
        }
        riemann<SOLVER, 2, DIMS>(vpz, vmn, gamma, gmm1, Fzn, Fzn[5], Fzn[6]);
#pragma unroll
        for (int nv = 0; nv < NV; nv++) vpz[nv] = vpn[nv];
      }

      if (valid) {
        if (MODE == MODE_FIRST) prim_to_cons(vc, u, gmm1);
        if (!PMX_PREFETCH_U && MODE != MODE_FIRST) {
#pragma unroll
          for (int nv = 0; nv < NV; nv++) {
            u[nv] = a.U[o + nv * a.sv];
            v0[nv] = a.V0[o + nv * a.sv];
          }
        }
        double cdt = 0.0;
        // x
        {
          const double* Fm = &sFx[ty * (TX + 1) + tx];   // face i-1/2; i+1/2 is Fm[1]
          double scrh = a.dtdx[0][i];
          double rhs[NV];
#pragma unroll
          for (int nv = 0; nv < NV; nv++) rhs[nv] = -scrh * (Fm[nv * NFX + 1] - Fm[nv * NFX]);
          rhs[VX1] -= scrh * (Fm[5 * NFX + 1] - Fm[5 * NFX]);
#pragma unroll
          for (int nv = 0; nv < NV; nv++) u[nv] += rhs[nv];
          if (MODE == MODE_FIRST)
            cdt = 0.0 + 0.5 * (Fm[6 * NFX] + Fm[6 * NFX + 1]) * a.inv_dl[0][i] * 1.0;
        }
        // y
        {
          const double* Fm = &sPlane[ty * TX + tx];       // face j-1/2; j+1/2 is Fm[TX]
          double scrh = a.dtdx[1][j];
          double rhs[NV];
#pragma unroll
          for (int nv = 0; nv < NV; nv++) rhs[nv] = -scrh * (Fm[nv * NFY + TX] - Fm[nv * NFY]);
          rhs[VX2] -= scrh * (Fm[5 * NFY + TX] - Fm[5 * NFY]);
#pragma unroll
          for (int nv = 0; nv < NV; nv++) u[nv] += rhs[nv];
          if (MODE == MODE_FIRST)
            cdt += 0.5 * (Fm[6 * NFY] + Fm[6 * NFY + TX]) * a.inv_dl[1][j] * 1.0;
        }
        // z
        if (DIMS == 3) {
          double scrh = a.dtdx[2][k];
          double rhs[NV];
#pragma unroll
          for (int nv = 0; nv < NV; nv++) rhs[nv] = -scrh * (Fzn[nv] - Fz[nv]);
          rhs[VX3] -= scrh * (Fzn[5] - Fz[5]);
#pragma unroll
          for (int nv = 0; nv < NV; nv++) u[nv] += rhs[nv];
          if (MODE == MODE_FIRST) cdt += 0.5 * (Fz[6] + Fzn[6]) * a.inv_dl[2][k] * 1.0;
        }
        if (MODE != MODE_FIRST) {
          double u0[NV];
          prim_to_cons(v0, u0, gmm1);
          if (MODE == MODE_WEIGHTED) {
#pragma unroll
            for (int nv = 0; nv < NV; nv++) u[nv] = a.w0 * u0[nv] + a.wc * u[nv];
          } else {
            const double one_third = 1.0 / 3.0;
#pragma unroll
            for (int nv = 0; nv < NV; nv++) u[nv] = one_third * (u0[nv] + 2.0 * u[nv]);
          }
        }
        double vn[NV];
        fail |= cons_to_prim(u, vn, gmm1);
#pragma unroll
        for (int nv = 0; nv < NV; nv++) a.Vout[o + nv * a.sv] = vn[nv];
        if (a.write_u) {
#pragma unroll
          for (int nv = 0; nv < NV; nv++) a.U[o + nv * a.sv] = u[nv];
        }
        if (MODE == MODE_FIRST) cdt_max = dmax(cdt_max, cdt);
      }
      if (DIMS == 3) {
#pragma unroll
        for (int q = 0; q < 7; q++) Fz[q] = Fzn[q];
#pragma unroll
        for (int s = 0; s < WZ - 1; s++)
#pragma unroll
          for (int nv = 0; nv < NV; nv++) w[s][nv] = w[s + 1][nv];
      }
    }
    __syncthreads();
  }

  // ---- block reductions: max C_dt (stage 1) and cons->prim flags -------
  if (MODE == MODE_FIRST) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) cdt_max = dmax(cdt_max, __shfl_xor(cdt_max, off));
    if ((tid & 63) == 0) sRed[tid >> 6] = cdt_max;
    __syncthreads();
    if (tid == 0) {
      double m = sRed[0];
      for (int q = 1; q < NT / 64; q++) m = dmax(m, sRed[q]);
      // non-negative doubles order like their bit patterns
      atomicMax(a.cdt_bits + red_index(blockIdx.x), (unsigned long long)__double_as_longlong(m));
    }
  }
  if (fail) atomicOr(a.err, fail);
}

// Launches stage_kernel<DIMS, REC, SOLVER, MODE, BCK_ANY, ...> (a domain
// whose sides do not all share one rule) on `stream`; those instantiations
// live in their own translation unit, pmx_hydro_any.hip.
void launch_stage_any(int dims, int rec, int solver, int mode, const StageArgs& a, int nblocks,
                      hipStream_t stream);

}  // namespace pmx
