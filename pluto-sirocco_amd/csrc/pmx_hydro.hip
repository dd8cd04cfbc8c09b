// libplutomx: hydro context, step driver and C ABI (see include/plutomx.h).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/plutomx.h"
#include "pmx_common.h"
#include "pmx_hydro_kernels.h"

namespace pmx {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

struct DevScalars {
  double t;                      // g_time
  double dt;                     // g_dt of the step being taken
  double inv_dt_hyp;             // Dts->invDt_hyp of the last step
  double tstop;                  // runtime.tstop
  double dt_uncut;               // the dt a cut to tstop replaced
  int err;                       // cons->prim flags of the step under way
  int step_err;                  // cons->prim flags of the last completed step
  int halt;                      // PMX_FLAG_DT_TOO_SMALL / PMX_FLAG_TSTOP once set (sticky):
                                 // later main-loop steps are no-ops
  int last;                      // the next step is the last one (dt clamped to tstop - t)
  unsigned long long cdt_red[RED_SLOTS * RED_STRIDE];   // stage-1 C_dt maxima, spread
                                                        // (words red_index(k), 0 = none)
};

// Src/main.c:226-229: the step that would reach tstop (to 1e-8 relative)
// is the last one, and its dt is cut to land on tstop.
__host__ __device__ inline void clamp_last_step(DevScalars& s) {
  if ((s.t + s.dt) >= s.tstop * (1.0 - 1.e-8)) {
    s.dt_uncut = s.dt;
    s.dt = (s.tstop - s.t);
    s.last = 1;
  }
}

// dt/dx for every cell and direction, and reset of the step reductions.
// A main-loop step (dt_override null) does nothing once the run has halted.
__global__ void prep_kernel(DevScalars* s, const double* dx, double* dtdx, int ntot,
                            const double* dt_override) {
  if (!dt_override && s->halt) return;
  double dt = dt_override ? *dt_override : s->dt;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  for (int q = t; q < ntot; q += gridDim.x * blockDim.x) dtdx[q] = dt / dx[q];
  if (t < RED_SLOTS) s->cdt_red[red_index(t)] = 0ull;
  if (t == 0) {
    if (dt_override) s->dt = dt;
    s->err = 0;
  }
}

// Src/update_stage.c:389-394 then Src/main.c:596-697 (NextTimeStep, no
// parabolic terms, no cooling) and the time update of main.c:296.  In the
// main loop (loop != 0) a halted run does nothing; NextTimeStep's "dt is
// too small" quit (main.c:672-683) halts it, and so does the end of the
// step that reached tstop; otherwise the next step's dt is cut to land on
// tstop when it would pass it (main.c:226-229).  One thread.  Returns
// whether the run goes on.
__device__ bool finish_step(DevScalars* s, int ndim, double cfl, double cfl_max_var,
                            double first_dt, int loop, const double* inv_in) {
  if (loop && s->halt) return false;
  double inv = inv_in ? *inv_in : 0.0;   // Dts->invDt_hyp the stage-1 maximum folds into
  unsigned long long m = 0ull;   // non-negative doubles order like their bit patterns
#pragma unroll
  for (int k = 0; k < RED_SLOTS; k++) {
    const unsigned long long v = s->cdt_red[red_index(k)];
    m = v > m ? v : m;
  }
  double c = __longlong_as_double((long long)m);
  inv = dmax(inv, c);
  inv /= (double)ndim;
  s->inv_dt_hyp = inv;
  s->step_err = s->err;
  s->err = 0;
  double dt = s->dt;
  s->t += dt;
  double dt_hyp = 1.0 / inv;
  dt_hyp *= cfl;
  double dtnext = dt_hyp;
  dtnext = dmin(dtnext, cfl_max_var * dt);
  if (cfl_max_var != 1.0) s->dt = dtnext;
  if (!loop) return false;
  if (dtnext < first_dt * 1.e-9) s->halt |= PMX_FLAG_DT_TOO_SMALL;   // main.c:672-683
  if (s->last) {
    s->halt |= PMX_FLAG_TSTOP;
    s->last = 0;
  } else {
    DevScalars q = *s;
    clamp_last_step(q);
    s->dt = q.dt;
    s->dt_uncut = q.dt_uncut;
    s->last = q.last;
  }
  return !s->halt;
}

// finish_step, then, in the main loop while the run goes on, what
// prep_kernel would do for the next step: dt/dx with the next dt and the
// cleared C_dt slots (the next step's own prep launch is then left out).
__global__ __launch_bounds__(256) void finish_kernel(DevScalars* s, int ndim, double cfl,
                                                     double cfl_max_var, double first_dt,
                                                     int loop, const double* inv_in,
                                                     const double* dx, double* dtdx, int ntot) {
  __shared__ int go;
  __shared__ double dt_next;
  if (threadIdx.x == 0) {
    go = finish_step(s, ndim, cfl, cfl_max_var, first_dt, loop, inv_in) ? 1 : 0;
    dt_next = s->dt;
  }
  __syncthreads();
  if (!go) return;
  for (int q = threadIdx.x; q < ntot; q += blockDim.x) dtdx[q] = dt_next / dx[q];
  if (threadIdx.x < RED_SLOTS) s->cdt_red[red_index(threadIdx.x)] = 0ull;
}

// Plain streaming copy, used to calibrate the HBM byte counters and to
// measure the bandwidth a simple kernel reaches on this device.
__global__ void copy_probe_kernel(const double2* __restrict__ src, double2* __restrict__ dst,
                                  long long n2) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (long long q = t; q < n2; q += stride) dst[q] = src[q];
}

}  // namespace pmx

using namespace pmx;

struct pmx_hydro {
  pmx_hydro_desc d;
  int nx, ny, nz;
  long long ncell;
  double* V[3];
  double* U;
  double* dx;       // [nx | ny | nz]
  double* dtdx;
  double* inv_dl;
  DevScalars* s;
  double* dt_in;    // device slots of an explicit step: dt, and Dts->invDt_hyp on entry
  hipStream_t stream;
  hipGraphExec_t graph;
  bool have_graph;
  bool timing;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pending;
  double timed_ms;
  int timed_launches;
  int cz;
};

namespace {

int z_schedule(int nz, int* zc0);

// The boundary kind of a domain: one rule on every side of the active
// directions is compiled in (BCK_OUTFLOW, BCK_PERIODIC), else BCK_ANY.
int boundary_kind(const pmx_hydro_desc& d) {
  const int ns = 2 * d.dims;
  bool out = true, per = true;
  for (int q = 0; q < ns; q++) {
    out = out && d.bc[q] == PMX_BC_OUTFLOW;
    per = per && d.bc[q] == PMX_BC_PERIODIC;
  }
  return out ? BCK_OUTFLOW : (per ? BCK_PERIODIC : BCK_ANY);
}

template <int DIMS, int REC, int SOLVER, int MODE>
void launch_stage(pmx_hydro* h, const StageArgs& a, int nblocks) {
  const dim3 g(nblocks), b(NTHREADS);
  switch (boundary_kind(h->d)) {
    case BCK_OUTFLOW:
      hipLaunchKernelGGL((stage_kernel<DIMS, REC, SOLVER, MODE, BCK_OUTFLOW, TILE_X,
                                       tile_y<DIMS>(), NTHREADS>), g, b, 0, h->stream, a);
      break;
    case BCK_PERIODIC:
      hipLaunchKernelGGL((stage_kernel<DIMS, REC, SOLVER, MODE, BCK_PERIODIC, TILE_X,
                                       tile_y<DIMS>(), NTHREADS>), g, b, 0, h->stream, a);
      break;
    default:   // mixed sides: compiled in pmx_hydro_any.hip
      launch_stage_any(DIMS, REC, SOLVER, MODE, a, nblocks, h->stream);
  }
}

template <int DIMS, int REC, int SOLVER>
void dispatch_mode(pmx_hydro* h, const StageArgs& a, int nblocks, int mode) {
  if (mode == MODE_FIRST) launch_stage<DIMS, REC, SOLVER, MODE_FIRST>(h, a, nblocks);
  else if (mode == MODE_WEIGHTED) launch_stage<DIMS, REC, SOLVER, MODE_WEIGHTED>(h, a, nblocks);
  else launch_stage<DIMS, REC, SOLVER, MODE_THIRD>(h, a, nblocks);
}

template <int DIMS, int REC>
void dispatch_solver(pmx_hydro* h, const StageArgs& a, int nblocks, int mode) {
  if (h->d.solver == PMX_SOLVER_ROE) dispatch_mode<DIMS, REC, SOL_ROE>(h, a, nblocks, mode);
  else dispatch_mode<DIMS, REC, SOL_HLLC>(h, a, nblocks, mode);
}

template <int DIMS>
void dispatch_rec(pmx_hydro* h, const StageArgs& a, int nblocks, int mode) {
  switch (h->d.recon) {
    case PMX_REC_PLM_VANLEER: dispatch_solver<DIMS, REC_PLM_VL>(h, a, nblocks, mode); break;
    case PMX_REC_PLM_MC: dispatch_solver<DIMS, REC_PLM_MC>(h, a, nblocks, mode); break;
    case PMX_REC_PLM_MINMOD: dispatch_solver<DIMS, REC_PLM_MM>(h, a, nblocks, mode); break;
    case PMX_REC_PLM_DEFAULT: dispatch_solver<DIMS, REC_PLM_DEFAULT>(h, a, nblocks, mode); break;
    default: dispatch_solver<DIMS, REC_PPM>(h, a, nblocks, mode); break;
  }
}

int stage(pmx_hydro* h, const double* Vin, const double* V0, double* Vout, int mode, double w0,
          double wc, int write_u, bool loop) {
  StageArgs a;
  a.Vin = Vin;
  a.V0 = V0;
  a.U = h->U;
  a.Vout = Vout;
  a.nx = h->nx; a.ny = h->ny; a.nz = h->nz;
  a.sv = h->ncell;
  for (int q = 0; q < 6; q++) a.bc[q] = h->d.bc[q];
  a.gamma = h->d.gamma;
  a.gmm1 = h->d.gamma - 1.0;
  a.dtdx[0] = h->dtdx; a.dtdx[1] = h->dtdx + h->nx; a.dtdx[2] = h->dtdx + h->nx + h->ny;
  a.inv_dl[0] = h->inv_dl; a.inv_dl[1] = h->inv_dl + h->nx; a.inv_dl[2] = h->inv_dl + h->nx + h->ny;
  a.w0 = w0; a.wc = wc;
  a.cz = h->cz;
  a.tiles_x = (h->nx + TILE_X - 1) / TILE_X;
  const int ty = h->d.dims == 3 ? tile_y<3>() : tile_y<2>();
  a.tiles_y = (h->ny + ty - 1) / ty;
  a.nzc = h->d.dims == 3 ? z_schedule(h->nz, a.zc0) : 0;
  a.tiles_z = (h->d.dims == 3) ? (a.nzc ? a.nzc : (h->nz + h->cz - 1) / h->cz) : 1;
  a.write_u = write_u;
  a.cdt_bits = h->s->cdt_red;
  a.err = &h->s->err;
  a.halt = loop ? &h->s->halt : nullptr;
  // with the z schedule, 8 ceil(tiles / 8) workgroups per chunk (a whole
  // number per XCD; see stage_kernel)
  int nblocks = a.nzc ? 8 * ((a.tiles_x * a.tiles_y + 7) / 8) * a.nzc
                      : a.tiles_x * a.tiles_y * a.tiles_z;

  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (h->timing) {
    if (h->ev_pool.size() < 2) {
      for (int q = 0; q < 64; q++) {
        hipEvent_t e;
        PMX_HIP_CHECK(hipEventCreate(&e));
        h->ev_pool.push_back(e);
      }
    }
    e0 = h->ev_pool.back(); h->ev_pool.pop_back();
    e1 = h->ev_pool.back(); h->ev_pool.pop_back();
    PMX_HIP_CHECK(hipEventRecord(e0, h->stream));
  }
  if (h->d.dims == 3) dispatch_rec<3>(h, a, nblocks, mode);
  else dispatch_rec<2>(h, a, nblocks, mode);
  PMX_HIP_CHECK(hipGetLastError());
  if (h->timing) {
    PMX_HIP_CHECK(hipEventRecord(e1, h->stream));
    h->ev_pending.push_back({e0, e1});
  }
  return 0;
}

// One AdvanceStep: Src/Time_Stepping/rk_step.c:117-330.  V[0] holds V^n on
// entry and V^{n+1} on exit.  In the main loop the previous step's
// finish_kernel has prepared this one (dt/dx, cleared slots) unless the
// run is just starting (with_prep).
int advance(pmx_hydro* h, const double* dt_override, bool with_prep) {
  int ntot = h->nx + h->ny + h->nz;
  const bool loop = dt_override == nullptr;
  if (with_prep || !loop) {
    hipLaunchKernelGGL(prep_kernel, dim3((ntot + 255) / 256), dim3(256), 0, h->stream, h->s,
                       h->dx, h->dtdx, ntot, dt_override);
    PMX_HIP_CHECK(hipGetLastError());
  }
  if (h->d.rk == 3) {
    if (stage(h, h->V[0], h->V[0], h->V[1], MODE_FIRST, 0, 0, 1, loop)) return -1;
    if (stage(h, h->V[1], h->V[0], h->V[2], MODE_WEIGHTED, 0.75, 0.25, 1, loop)) return -1;
    if (stage(h, h->V[2], h->V[0], h->V[0], MODE_THIRD, 0, 0, 1, loop)) return -1;
  } else {
    if (stage(h, h->V[0], h->V[0], h->V[1], MODE_FIRST, 0, 0, 1, loop)) return -1;
    if (stage(h, h->V[1], h->V[0], h->V[0], MODE_WEIGHTED, 0.5, 0.5, 1, loop)) return -1;
  }
  hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(256), 0, h->stream, h->s, h->d.dims,
                     h->d.cfl, h->d.cfl_max_var, h->d.first_dt, loop ? 1 : 0,
                     loop ? nullptr : dt_override + 1, h->dx, h->dtdx, ntot);
  PMX_HIP_CHECK(hipGetLastError());
  return 0;
}

int collect_timing(pmx_hydro* h) {
  for (auto& p : h->ev_pending) {
    float ms = 0.f;
    PMX_HIP_CHECK(hipEventSynchronize(p.second));
    PMX_HIP_CHECK(hipEventElapsedTime(&ms, p.first, p.second));
    h->timed_ms += ms;
    h->timed_launches++;
    h->ev_pool.push_back(p.first);
    h->ev_pool.push_back(p.second);
  }
  h->ev_pending.clear();
  return 0;
}

// strided copy helpers between PLUTO's ghosted Vc and the interior layout
void copy_interior(const pmx_hydro* h, const double* src, double* dst, bool src_ghost, int ng) {
  int gx = ng, gy = ng, gz = (h->d.dims == 3) ? ng : 0;
  long long tx = h->nx + 2 * gx, ty = h->ny + 2 * gy, tz = h->nz + 2 * gz;
  for (int nv = 0; nv < NV; nv++)
    for (int k = 0; k < h->nz; k++)
      for (int j = 0; j < h->ny; j++) {
        long long oi = ((long long)nv * h->nz + k) * h->ny * h->nx + (long long)j * h->nx;
        long long og = ((nv * tz + k + gz) * ty + j + gy) * tx + gx;
        if (src_ghost) memcpy(dst + oi, src + og, sizeof(double) * h->nx);
        else memcpy(dst + og, src + oi, sizeof(double) * h->nx);
      }
}

int host_map(int g, int n, int bl, int br, bool& flip) {
  if (g < 0) {
    if (bl == BC_PERIODIC) g += n;
    else if (bl == BC_OUTFLOW) g = 0;
    else { g = -1 - g; flip = !flip; }
  } else if (g >= n) {
    if (br == BC_PERIODIC) g -= n;
    else if (br == BC_OUTFLOW) g = n - 1;
    else { g = 2 * n - 1 - g; flip = !flip; }
  }
  return g;
}

// The checks pmx_hydro_create makes of a descriptor.
int check_desc(const pmx_hydro_desc* desc) {
  if (desc->dims != 2 && desc->dims != 3) { set_error("dims must be 2 or 3"); return -1; }
  if (desc->rk != 2 && desc->rk != 3) { set_error("rk must be 2 or 3"); return -1; }
  if (desc->recon != PMX_REC_PLM_VANLEER && desc->recon != PMX_REC_PLM_MC &&
      desc->recon != PMX_REC_PLM_MINMOD && desc->recon != PMX_REC_PLM_DEFAULT &&
      desc->recon != PMX_REC_PPM) {
    set_error("unknown reconstruction"); return -1;
  }
  if (desc->solver != PMX_SOLVER_HLLC && desc->solver != PMX_SOLVER_ROE) {
    set_error("unknown Riemann solver"); return -1;
  }
  for (int q = 0; q < 6; q++)
    if (desc->bc[q] < 0 || desc->bc[q] > 2) { set_error("unknown boundary type"); return -1; }
  int nz = desc->dims == 3 ? desc->nx3 : 1;
  int H = desc->recon == PMX_REC_PPM ? 3 : 2;
  if (desc->nx1 < H || desc->nx2 < H || (desc->dims == 3 && desc->nx3 < H)) {
    set_error("grid smaller than the ghost width"); return -1;
  }
  if (!desc->dx1 || !desc->dx2 || (desc->dims == 3 && !desc->dx3)) {
    set_error("missing cell widths"); return -1;
  }
  return 0;
}

// z planes per workgroup: long enough to amortise the column prologue,
// short enough to give several workgroups per CU.  PMX_CZ overrides.
int z_chunk(int nz) {
  int cz = nz < 64 ? 8 : 16;
  if (const char* e = getenv("PMX_CZ")) cz = atoi(e) > 0 ? atoi(e) : cz;
  return cz;
}

// z chunks of decreasing length for a 3-D grid: 64 planes while more than
// 128 remain, then 32 while more than 64, 16 while more than 32, 8 while
// more than 16, then 8-plane pieces (nz = 256: 64, 64, 32, 32, 16, 16, 8,
// 8, 8, 8).  Each chunk re-reads 3 + 3 ghost planes and redoes two z
// states and a z face per column, so long chunks cost less; but a launch
// of equal chunks leaves its last round of workgroups partly empty (Sedov
// 256^3 in 16-plane chunks: 4,864 workgroups, 512 resident at a time, 9.5
// rounds), whereas short chunks dispatched last let every CU finish
// together: 1.265 vs 1.311 ms per stage (32-plane chunks first: 1.273;
// equal 32-plane chunks: 1.354).  Writes zc0[0..n] (zc0[n] = nz) and
// returns the number n of chunks, or 0 for the uniform chunks of z_chunk
// (PMX_CZ set, or more chunks than fit).
int z_schedule(int nz, int* zc0) {
  if (getenv("PMX_CZ")) return 0;
  int n = 0, z = 0, rem = nz;
  auto push = [&](int len) { zc0[n++] = z; z += len; rem -= len; };
  for (int size = 64; size >= 8; size /= 2)
    while (rem > 2 * size && n < PMX_MAX_ZCHUNKS) push(size);
  while (rem > 0 && n < PMX_MAX_ZCHUNKS) push(rem < 8 ? rem : 8);
  zc0[n] = nz;
  return rem == 0 ? n : 0;
}

// dx per cell of every direction ([nx | ny | nz]) and its reciprocal
// (Src/set_geometry.c:305).
void host_tables(const pmx_hydro_desc* desc, int nx, int ny, int nz, std::vector<double>& dx,
                 std::vector<double>& inv) {
  const int ntot = nx + ny + nz;
  dx.resize(ntot);
  inv.resize(ntot);
  for (int i = 0; i < nx; i++) dx[i] = desc->dx1[i];
  for (int j = 0; j < ny; j++) dx[nx + j] = desc->dx2[j];
  for (int k = 0; k < nz; k++) dx[nx + ny + k] = desc->dims == 3 ? desc->dx3[k] : 1.0;
  for (int q = 0; q < ntot; q++) inv[q] = 1.0 / dx[q];
}

}  // namespace

extern "C" {

const char* pmx_last_error(void) { return g_err.c_str(); }

int pmx_hydro_create(const pmx_hydro_desc* desc, pmx_hydro** out) {
  if (!desc || !out) { set_error("null argument"); return -1; }
  if (check_desc(desc)) return -1;
  int nz = desc->dims == 3 ? desc->nx3 : 1;
  pmx_hydro* h = new pmx_hydro();
  h->d = *desc;
  h->nx = desc->nx1; h->ny = desc->nx2; h->nz = nz;
  h->ncell = (long long)h->nx * h->ny * h->nz;
  h->cz = z_chunk(h->nz);
  size_t vbytes = sizeof(double) * NV * h->ncell;
  int ntot = h->nx + h->ny + h->nz;
#define CK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { \
    set_error(std::string(#x) + ": " + hipGetErrorString(e_)); pmx_hydro_destroy(h); return -1; } } while (0)
  CK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  for (int q = 0; q < 3; q++) CK(hipMalloc(&h->V[q], vbytes));
  CK(hipMalloc(&h->U, vbytes));
  CK(hipMalloc(&h->dx, sizeof(double) * ntot));
  CK(hipMalloc(&h->dtdx, sizeof(double) * ntot));
  CK(hipMalloc(&h->inv_dl, sizeof(double) * ntot));
  CK(hipMalloc(&h->s, sizeof(DevScalars)));
  CK(hipMalloc(&h->dt_in, 2 * sizeof(double)));
  {
    std::vector<double> dx, inv;
    host_tables(desc, h->nx, h->ny, h->nz, dx, inv);
    CK(hipMemcpy(h->dx, dx.data(), sizeof(double) * ntot, hipMemcpyHostToDevice));
    CK(hipMemcpy(h->inv_dl, inv.data(), sizeof(double) * ntot, hipMemcpyHostToDevice));
    DevScalars s0;
    memset(&s0, 0, sizeof(s0));
    s0.tstop = HUGE_VAL;
    CK(hipMemcpy(h->s, &s0, sizeof(s0), hipMemcpyHostToDevice));
  }
  for (int q = 0; q < 3; q++) CK(hipMemset(h->V[q], 0, vbytes));
  CK(hipMemset(h->U, 0, vbytes));
#undef CK
  *out = h;
  return 0;
}

int pmx_hydro_destroy(pmx_hydro* h) {
  if (!h) return 0;
  if (h->stream) hipStreamSynchronize(h->stream);
  if (h->have_graph) hipGraphExecDestroy(h->graph);
  for (auto& p : h->ev_pending) { h->ev_pool.push_back(p.first); h->ev_pool.push_back(p.second); }
  for (auto e : h->ev_pool) hipEventDestroy(e);
  for (int q = 0; q < 3; q++) if (h->V[q]) hipFree(h->V[q]);
  if (h->U) hipFree(h->U);
  if (h->dx) hipFree(h->dx);
  if (h->dtdx) hipFree(h->dtdx);
  if (h->inv_dl) hipFree(h->inv_dl);
  if (h->s) hipFree(h->s);
  if (h->dt_in) hipFree(h->dt_in);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
  return 0;
}

int pmx_hydro_set_prim(pmx_hydro* h, const double* V, int ghosts) {
  if (!h || !V) { set_error("null argument"); return -1; }
  size_t n = (size_t)NV * h->ncell;
  if (ghosts) {
    std::vector<double> tmp(n);
    copy_interior(h, V, tmp.data(), true, h->d.recon == PMX_REC_PPM ? 3 : 2);
    PMX_HIP_CHECK(hipMemcpyAsync(h->V[0], tmp.data(), n * sizeof(double), hipMemcpyHostToDevice,
                                 h->stream));
    PMX_HIP_CHECK(hipStreamSynchronize(h->stream));
  } else {
    PMX_HIP_CHECK(hipMemcpyAsync(h->V[0], V, n * sizeof(double), hipMemcpyHostToDevice,
                                 h->stream));
    PMX_HIP_CHECK(hipStreamSynchronize(h->stream));
  }
  return 0;
}

int pmx_hydro_get_prim(pmx_hydro* h, double* V, int ghosts) {
  if (!h || !V) { set_error("null argument"); return -1; }
  size_t n = (size_t)NV * h->ncell;
  if (!ghosts) {
    PMX_HIP_CHECK(hipMemcpyAsync(V, h->V[0], n * sizeof(double), hipMemcpyDeviceToHost,
                                 h->stream));
    PMX_HIP_CHECK(hipStreamSynchronize(h->stream));
    return 0;
  }
  std::vector<double> tmp(n);
  PMX_HIP_CHECK(hipMemcpyAsync(tmp.data(), h->V[0], n * sizeof(double), hipMemcpyDeviceToHost,
                               h->stream));
  PMX_HIP_CHECK(hipStreamSynchronize(h->stream));
  int ng = h->d.recon == PMX_REC_PPM ? 3 : 2;
  int gz = h->d.dims == 3 ? ng : 0;
  long long tx = h->nx + 2 * ng, ty = h->ny + 2 * ng, tz = h->nz + 2 * gz;
  for (int nv = 0; nv < NV; nv++)
    for (long long k = 0; k < tz; k++)
      for (long long j = 0; j < ty; j++)
        for (long long i = 0; i < tx; i++) {
          bool flip = false;
          int ii = host_map((int)i - ng, h->nx, h->d.bc[0], h->d.bc[1], flip);
          int jj = host_map((int)j - ng, h->ny, h->d.bc[2], h->d.bc[3], flip);
          bool fx = false, fy = false, fz = false;
          ii = host_map((int)i - ng, h->nx, h->d.bc[0], h->d.bc[1], fx);
          jj = host_map((int)j - ng, h->ny, h->d.bc[2], h->d.bc[3], fy);
          int kk = h->d.dims == 3 ? host_map((int)k - gz, h->nz, h->d.bc[4], h->d.bc[5], fz) : 0;
          double val = tmp[(((long long)nv * h->nz + kk) * h->ny + jj) * h->nx + ii];
          if ((nv == 1 && fx) || (nv == 2 && fy) || (nv == 3 && fz)) val = -val;
          V[((nv * tz + k) * ty + j) * tx + i] = val;
          (void)flip;
        }
  return 0;
}

int pmx_hydro_get_cons(pmx_hydro* h, double* U, int pluto_layout) {
  if (!h || !U) { set_error("null argument"); return -1; }
  size_t n = (size_t)NV * h->ncell;
  if (!pluto_layout) {
    PMX_HIP_CHECK(hipMemcpyAsync(U, h->U, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    PMX_HIP_CHECK(hipStreamSynchronize(h->stream));
    return 0;
  }
  std::vector<double> tmp(n);
  PMX_HIP_CHECK(hipMemcpyAsync(tmp.data(), h->U, n * sizeof(double), hipMemcpyDeviceToHost,
                               h->stream));
  PMX_HIP_CHECK(hipStreamSynchronize(h->stream));
  for (long long c = 0; c < h->ncell; c++)
    for (int nv = 0; nv < NV; nv++) U[c * NV + nv] = tmp[nv * h->ncell + c];
  return 0;
}

int pmx_hydro_set_time(pmx_hydro* h, double t, double dt) {
  if (!h) { set_error("null argument"); return -1; }
  DevScalars s0;
  PMX_HIP_CHECK(hipMemcpyAsync(&s0, h->s, sizeof(s0), hipMemcpyDeviceToHost, h->stream));
  PMX_HIP_CHECK(hipStreamSynchronize(h->stream));
  s0.t = t;
  s0.dt = dt;
  s0.halt = 0;
  s0.last = 0;
  clamp_last_step(s0);
  PMX_HIP_CHECK(hipMemcpyAsync(h->s, &s0, sizeof(s0), hipMemcpyHostToDevice, h->stream));
  PMX_HIP_CHECK(hipStreamSynchronize(h->stream));
  return 0;
}

int pmx_hydro_set_tstop(pmx_hydro* h, double tstop) {
  if (!h || !(tstop > -HUGE_VAL)) { set_error("bad argument"); return -1; }
  DevScalars s0;
  PMX_HIP_CHECK(hipMemcpyAsync(&s0, h->s, sizeof(s0), hipMemcpyDeviceToHost, h->stream));
  PMX_HIP_CHECK(hipStreamSynchronize(h->stream));
  s0.tstop = tstop;
  s0.halt &= ~PMX_FLAG_TSTOP;
  if (s0.last) {           // undo a cut made for the previous tstop
    s0.dt = s0.dt_uncut;
    s0.last = 0;
  }
  clamp_last_step(s0);
  PMX_HIP_CHECK(hipMemcpyAsync(h->s, &s0, sizeof(s0), hipMemcpyHostToDevice, h->stream));
  PMX_HIP_CHECK(hipStreamSynchronize(h->stream));
  return 0;
}

int pmx_hydro_get_time(pmx_hydro* h, double* t, double* dt, double* inv_dt_hyp, int* flags) {
  if (!h) { set_error("null argument"); return -1; }
  DevScalars s0;
  PMX_HIP_CHECK(hipMemcpyAsync(&s0, h->s, sizeof(s0), hipMemcpyDeviceToHost, h->stream));
  PMX_HIP_CHECK(hipStreamSynchronize(h->stream));
  if (t) *t = s0.t;
  if (dt) *dt = s0.dt;
  if (inv_dt_hyp) *inv_dt_hyp = s0.inv_dt_hyp;
  if (flags) *flags = s0.step_err | s0.err | s0.halt;
  return 0;
}

int pmx_hydro_advance_step(pmx_hydro* h, double dt, double* inv_dt_hyp) {
  if (!h) { set_error("null argument"); return -1; }
  const double in[2] = {dt, inv_dt_hyp ? *inv_dt_hyp : 0.0};
  PMX_HIP_CHECK(hipMemcpyAsync(h->dt_in, in, sizeof(in), hipMemcpyHostToDevice, h->stream));
  if (advance(h, h->dt_in, true)) return -1;
  if (h->timing) collect_timing(h);
  DevScalars s0;
  PMX_HIP_CHECK(hipMemcpyAsync(&s0, h->s, sizeof(s0), hipMemcpyDeviceToHost, h->stream));
  PMX_HIP_CHECK(hipStreamSynchronize(h->stream));
  if (inv_dt_hyp) *inv_dt_hyp = s0.inv_dt_hyp;
  return 0;
}

int pmx_hydro_run(pmx_hydro* h, int nsteps) {
  if (!h || nsteps < 0) { set_error("bad argument"); return -1; }
  for (int n = 0; n < nsteps; n++)
    if (advance(h, nullptr, n == 0)) return -1;
  return 0;
}

// Steps per captured graph: a graph launch costs more than the few kernel
// launches of one step saves (KH 256^2: 33 us per step replaying one step
// per graph against 26 us eager), so a graph holds several steps.  Steps
// after the run halts are no-ops on the device, as in the eager loop.
constexpr int GRAPH_STEPS = 8;

int pmx_hydro_run_graph(pmx_hydro* h, int nsteps) {
  if (!h || nsteps < 0) { set_error("bad argument"); return -1; }
  if (h->timing) return pmx_hydro_run(h, nsteps);
  if (nsteps >= GRAPH_STEPS && !h->have_graph) {
    hipGraph_t g;
    PMX_HIP_CHECK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (int q = 0; q < GRAPH_STEPS && !rc; q++) rc = advance(h, nullptr, q == 0);
    hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (rc) return -1;
    if (e != hipSuccess) { set_error(std::string("capture: ") + hipGetErrorString(e)); return -1; }
    PMX_HIP_CHECK(hipGraphInstantiate(&h->graph, g, nullptr, nullptr, 0));
    PMX_HIP_CHECK(hipGraphDestroy(g));
    h->have_graph = true;
  }
  int n = 0;
  for (; nsteps - n >= GRAPH_STEPS; n += GRAPH_STEPS)
    PMX_HIP_CHECK(hipGraphLaunch(h->graph, h->stream));
  for (const int n0 = n; n < nsteps; n++)
    if (advance(h, nullptr, n == n0)) return -1;
  return 0;
}

int pmx_hydro_sync(pmx_hydro* h) {
  if (!h) { set_error("null argument"); return -1; }
  PMX_HIP_CHECK(hipStreamSynchronize(h->stream));
  if (h->timing) collect_timing(h);
  return 0;
}

int pmx_hydro_stage_timing(pmx_hydro* h, int enable, int reset, double* total_ms, int* launches) {
  if (!h) { set_error("null argument"); return -1; }
  PMX_HIP_CHECK(hipStreamSynchronize(h->stream));
  collect_timing(h);
  if (total_ms) *total_ms = h->timed_ms;
  if (launches) *launches = h->timed_launches;
  if (reset) { h->timed_ms = 0.0; h->timed_launches = 0; }
  h->timing = enable != 0;
  return 0;
}

// ---------------------------------------------------------------------
// The step without allocation: pmx_hydro_advance_step for callers that
// cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).  The caller passes the
// state; everything else the step needs (the intermediate stage states,
// the dt/dx tables, the time scalars) lives in `scratch`, a fixed array
// in device memory that every GPU holds for the life of the process.
// ---------------------------------------------------------------------
}  // extern "C"

namespace {

constexpr long long NOALLOC_MAX_CELLS = 1ll << 19;   // 2 stage states of 5 doubles
constexpr int NOALLOC_MAX_LINE = 8192;               // nx + ny + nz
constexpr size_t NOALLOC_SCALARS = 16 + RED_SLOTS * RED_STRIDE;   // doubles for DevScalars
// [V1 | V2] stage states, then dx, dt/dx, 1/dx tables, the dt slots and
// the time scalars (with their spread reduction words): 40 MiB + 196 KiB
constexpr size_t NOALLOC_DOUBLES = 2 * (size_t)NV * NOALLOC_MAX_CELLS +
                                   3 * (size_t)NOALLOC_MAX_LINE + 2 + NOALLOC_SCALARS;
static_assert(sizeof(DevScalars) <= NOALLOC_SCALARS * sizeof(double), "scalar slots");

__device__ __align__(256) double scratch[NOALLOC_DOUBLES];

constexpr int NOALLOC_MAX_DEVICES = 64;
double* g_scratch[NOALLOC_MAX_DEVICES];   // scratch's address on each device, once looked up

// scratch's device address on the current device (host code never passes
// the array's name itself to a kernel: on the host it is a stand-in).
int noalloc_scratch(double** out) {
  int dev = 0;
  PMX_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) { set_error("device index out of range"); return -1; }
  if (!g_scratch[dev]) {
    void* p = nullptr;
    PMX_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch[dev] = static_cast<double*>(p);
  }
  *out = g_scratch[dev];
  return 0;
}

}  // namespace

extern "C" {

int pmx_hydro_advance_step_noalloc(const pmx_hydro_desc* desc, double* V, double* U, double dt,
                                   double* inv_dt_hyp) {
  if (!desc || !V || !U) { set_error("null argument"); return -1; }
  if (check_desc(desc)) return -1;
  const int nx = desc->nx1, ny = desc->nx2, nz = desc->dims == 3 ? desc->nx3 : 1;
  const long long ncell = (long long)nx * ny * nz;
  if (ncell > NOALLOC_MAX_CELLS || nx + ny + nz > NOALLOC_MAX_LINE) {
    set_error("pmx_hydro_advance_step_noalloc: the grid (" + std::to_string(ncell) +
              " cells, nx+ny+nz = " + std::to_string(nx + ny + nz) + ") exceeds the limit of " +
              std::to_string(NOALLOC_MAX_CELLS) + " cells and nx+ny+nz <= " +
              std::to_string(NOALLOC_MAX_LINE) + " (use pmx_hydro_create)");
    return -1;
  }
  double* base = nullptr;
  if (noalloc_scratch(&base)) return -1;
  // a context on the stack whose device buffers are the caller's state and
  // slices of scratch; the step runs on the null stream and is waited for
  pmx_hydro h;
  h.d = *desc;
  h.nx = nx; h.ny = ny; h.nz = nz;
  h.ncell = ncell;
  h.cz = z_chunk(nz);
  h.V[0] = V;
  h.V[1] = base;
  h.V[2] = base + (size_t)NV * ncell;
  h.U = U;
  double* tab = base + 2 * (size_t)NV * NOALLOC_MAX_CELLS;
  const int ntot = nx + ny + nz;
  h.dx = tab;
  h.dtdx = tab + ntot;
  h.inv_dl = tab + 2 * ntot;
  h.dt_in = tab + 3 * NOALLOC_MAX_LINE;
  h.s = reinterpret_cast<DevScalars*>(h.dt_in + 2);
  h.stream = nullptr;
  h.have_graph = false;
  h.timing = false;
  h.timed_ms = 0.0;
  h.timed_launches = 0;
  std::vector<double> dx, inv;
  host_tables(desc, nx, ny, nz, dx, inv);
  DevScalars s0;
  memset(&s0, 0, sizeof(s0));       // what the last call left is cleared
  s0.tstop = HUGE_VAL;
  const double in[2] = {dt, inv_dt_hyp ? *inv_dt_hyp : 0.0};
  PMX_HIP_CHECK(hipMemcpyAsync(h.dx, dx.data(), sizeof(double) * ntot, hipMemcpyHostToDevice, 0));
  PMX_HIP_CHECK(hipMemcpyAsync(h.inv_dl, inv.data(), sizeof(double) * ntot,
                               hipMemcpyHostToDevice, 0));
  PMX_HIP_CHECK(hipMemcpyAsync(h.s, &s0, sizeof(s0), hipMemcpyHostToDevice, 0));
  PMX_HIP_CHECK(hipMemcpyAsync(h.dt_in, in, sizeof(in), hipMemcpyHostToDevice, 0));
  if (advance(&h, h.dt_in, true)) return -1;
  PMX_HIP_CHECK(hipMemcpyAsync(&s0, h.s, sizeof(s0), hipMemcpyDeviceToHost, 0));
  PMX_HIP_CHECK(hipStreamSynchronize(0));
  if (inv_dt_hyp) *inv_dt_hyp = s0.inv_dt_hyp;
  return 0;
}

int pmx_copy_probe(long long nbytes, int reps, double* ms_per_copy) {
  if (nbytes <= 0 || nbytes % 16 != 0 || reps <= 0) { set_error("bad argument"); return -1; }
  double2 *a = nullptr, *b = nullptr;
  hipEvent_t e0, e1;
  PMX_HIP_CHECK(hipMalloc(&a, nbytes));
  if (hipMalloc(&b, nbytes) != hipSuccess) { hipFree(a); set_error("hipMalloc"); return -1; }
  hipMemset(a, 0, nbytes);
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  long long n2 = nbytes / 16;
  // 1024 workgroups of 256 threads striding over the buffer: the best of a
  // sweep of grid sizes (1024 .. 65536 workgroups) and per-thread unrolling
  // (1 .. 8 loads in flight) on MI355X, ~5.5 TB/s for 2 GiB
  hipLaunchKernelGGL(copy_probe_kernel, dim3(1024), dim3(256), 0, 0, a, b, n2);
  hipEventRecord(e0, 0);
  for (int r = 0; r < reps; r++)
    hipLaunchKernelGGL(copy_probe_kernel, dim3(1024), dim3(256), 0, 0, a, b, n2);
  hipEventRecord(e1, 0);
  hipError_t err = hipEventSynchronize(e1);
  float ms = 0.f;
  hipEventElapsedTime(&ms, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(a);
  hipFree(b);
  if (err != hipSuccess) { set_error(hipGetErrorString(err)); return -1; }
  if (ms_per_copy) *ms_per_copy = ms / reps;
  return 0;
}

int pmx_copy_probe_runtime(long long nbytes, int reps, double* ms_per_copy) {
  if (nbytes <= 0 || reps <= 0) { set_error("bad argument"); return -1; }
  void *a = nullptr, *b = nullptr;
  hipEvent_t e0, e1;
  PMX_HIP_CHECK(hipMalloc(&a, nbytes));
  if (hipMalloc(&b, nbytes) != hipSuccess) { hipFree(a); set_error("hipMalloc"); return -1; }
  hipMemset(a, 0, nbytes);
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipError_t err = hipMemcpyAsync(b, a, nbytes, hipMemcpyDeviceToDevice, 0);
  if (err == hipSuccess) err = hipEventRecord(e0, 0);
  for (int r = 0; r < reps && err == hipSuccess; r++)
    err = hipMemcpyAsync(b, a, nbytes, hipMemcpyDeviceToDevice, 0);
  if (err == hipSuccess) err = hipEventRecord(e1, 0);
  if (err == hipSuccess) err = hipEventSynchronize(e1);
  float ms = 0.f;
  if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(a);
  hipFree(b);
  if (err != hipSuccess) { set_error(hipGetErrorString(err)); return -1; }
  if (ms_per_copy) *ms_per_copy = ms / reps;
  return 0;
}

int pmx_hydro_device_state(pmx_hydro* h, double** V, double** U) {
  if (!h) { set_error("null argument"); return -1; }
  if (V) *V = h->V[0];
  if (U) *U = h->U;
  return 0;
}

}  // extern "C"
