// libplutomx: MLP cooling surrogate (no counterpart in the reference).
//
// A 4 -> 64 -> 64 -> 64 -> 1 network with SiLU activations, BF16 weights
// and activations as matrix-core operands, FP32 accumulation and FP32
// biases.  Inputs are (log10 rho[g/cm^3], log10 T[K], log10 xi,
// log10(dt / t_cool)) with t_cool = E / |dE/dt| of the cell's initial
// state; the output times YS is log10(T_final / T_initial) over dt.
//
// Kernels
//   mlp_sample_kernel   training samples labelled by the plain cooling
//                       integrator (the same device code as the wind's
//                       SplitSource), log-uniform over the training hull
//   mlp_train_kernel    forward + backward of 16-sample tiles on
//                       v_mfma_f32_16x16x16_bf16; per-workgroup partial
//                       gradients (deterministic reduction)
//   mlp_update_kernel   gradient reduction in a fixed order + Adam (FP32
//                       master weights) + BF16 copies for the next step
//   mlp_predict_kernel  forward only
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/plutomx.h"
#include "pmx_common.h"
#include "pmx_mlp.h"
#include "pmx_wind_host.h"

namespace pmx {
namespace mlp {

using wind::Const;
using wind::RadCell;

// ---------------------------------------------------------------------
// samples
// ---------------------------------------------------------------------
__device__ __forceinline__ unsigned long long mix64(unsigned long long x) {
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__device__ __forceinline__ double unif(unsigned long long seed, unsigned long long idx) {
  unsigned long long h = mix64(seed + (idx + 1ull) * 0x9E3779B97F4A7C15ull);
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

// One training sample: a single cell with density, temperature, ionisation
// parameter and dt / t_cool drawn log-uniformly over the hull, integrated
// over dt by the plain integrator (g_time > 3, so Radiat takes xi and T_r
// from the SIROCCO arrays; tracer 1).
__global__ void mlp_sample_kernel(Const k, long long n, unsigned long long seed, float* X,
                                  float* Y, int* nev, double* aux) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  double lr = HULL_LO[0] + (HULL_HI[0] - HULL_LO[0]) * unif(seed, 4 * q + 0);
  double lT = HULL_LO[1] + (HULL_HI[1] - HULL_LO[1]) * unif(seed, 4 * q + 1);
  double lx = HULL_LO[2] + (HULL_HI[2] - HULL_LO[2]) * unif(seed, 4 * q + 2);
  double lf = HULL_LO[3] + (HULL_HI[3] - HULL_LO[3]) * unif(seed, 4 * q + 3);
  double rho = pow(10.0, lr) / k.UD;
  double p0 = rho * pow(10.0, lT) / (k.KELVIN * k.mu);
  RadCell R;
  R.rho = rho; R.trc = 1.0; R.x1 = 1.0; R.xi_s = pow(10.0, lx); R.tr_s = k.t_x; R.t = 5.0;
  wind::load_pre(nullptr, 0, 0, R.pre);
  R.nev = 0;
  double T0 = p0 / rho * k.KELVIN * k.mu;
  double E0 = p0 / (k.gamma - 1.0);
  double k1 = wind::radiat(k, R, E0);
  float* x = X + 4 * q;
  double* ax = aux + 4 * q;
  ax[0] = rho; ax[1] = p0; ax[2] = R.xi_s; ax[3] = 0.0;
  x[0] = (float)log10(rho * k.UD);
  x[1] = (float)log10(T0);
  x[2] = (float)lx;
  if (!(k1 != 0.0) || !isfinite(k1)) {
    x[3] = (float)lf;
    Y[q] = NAN;
    nev[q] = 0;
    return;
  }
  double tau = E0 / fabs(k1);
  double dt = pow(10.0, lf) * tau;
  ax[3] = dt;
  x[3] = (float)log10(dt / tau);
  R.nev = 0;
  double dtc;
  int bad = 0;
  double pf = wind::plain_cell(k, R, p0, dt, dtc, bad);
  double Tf = pf / rho * k.KELVIN * k.mu;
  Y[q] = bad ? NAN : (float)log10(Tf / T0);
  nev[q] = R.nev;
}

// ---------------------------------------------------------------------
// training
// ---------------------------------------------------------------------
struct alignas(16) TrainLds {
  unsigned short Xb[16 * 16], XT[16 * 16];
  float Z1[16 * 64], Z2[16 * 64], Z3[16 * 64];
  unsigned short H1[16 * 64], H2[16 * 64], H3[16 * 64];
  unsigned short H1T[64 * 16], H2T[64 * 16];
  unsigned short dA[16 * 64], dAT[64 * 16];
  unsigned short dB[16 * 64], dBT[64 * 16];
  float y[16], t[16], dy[16];
};

// One 64-lane workgroup per block; each runs `tiles` tiles of 16 samples
// starting at sample `start + blockIdx.x * tiles * 16` (mod n), and writes
// its summed gradient of  L = sum (y - t)^2 / batch  to partial[block],
// where t = label / YS is the scaled target.
__global__ __launch_bounds__(64) void mlp_train_kernel(Weights w, const float* X, const float* Y,
                                                       long long n, long long start, int tiles,
                                                       float inv_batch, long long nvalid_limit,
                                                       float* partial, float* lossp) {
  __shared__ TrainLds S;
  const int l = threadIdx.x, r = l & 15, c = l & 15, r0 = (l >> 4) * 4, kq = (l >> 4) * 4;
  f4 acc1[4], acc2[4][4], acc3[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++) {
    acc1[a] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int b = 0; b < 4; b++) acc2[a][b] = acc3[a][b] = f4{0.f, 0.f, 0.f, 0.f};
  }
  float gb1[4] = {0.f, 0.f, 0.f, 0.f}, gb2[4] = {0.f, 0.f, 0.f, 0.f}, gb3[4] = {0.f, 0.f, 0.f, 0.f};
  float g4 = 0.f, gb4 = 0.f, loss = 0.f;
  const long long base = start + (long long)blockIdx.x * tiles * 16;
  for (int tile = 0; tile < tiles; tile++) {
    // ---- load 16 samples -------------------------------------------
    const long long lin = base + tile * 16 + r;   // position inside the batch
    const long long sidx = lin % n;
    const float label = Y[sidx];
    // samples past the batch and unlabelled ones (NaN) carry no loss
    const bool valid = (lin - start) < nvalid_limit && label == label;
    float xin[4];
    load_features(X + 4 * sidx, xin);
#pragma unroll
    for (int f = 0; f < 4; f++) {
      int col = kq + f;
      float v = (col < 4 && valid) ? xin[col] : 0.f;
      unsigned short b = f2bf(v);
      S.Xb[r * 16 + col] = b;
      S.XT[col * 16 + r] = b;
    }
    if (l < 16) {
      S.t[r] = valid ? label * (1.f / YS) : 0.f;   // the network fits log10(T_f/T_0) / YS
      S.dy[r] = valid ? 1.f : 0.f;   // validity mask until the loss is known
    }
    __syncthreads();
    // ---- forward ---------------------------------------------------
    f4 acc[4];
    dense16<16>(S.Xb, w.W1p, acc, l);
    epilogue(acc, w.P + OFF_B1, S.Z1, S.H1, S.H1T, l);
    __syncthreads();
    dense16<64>(S.H1, w.W2, acc, l);
    epilogue(acc, w.P + OFF_B2, S.Z2, S.H2, S.H2T, l);
    __syncthreads();
    dense16<64>(S.H2, w.W3, acc, l);
    epilogue(acc, w.P + OFF_B3, S.Z3, S.H3, nullptr, l);
    __syncthreads();
    float y = output_unit(S.H3, w, l);
    if (l < 16) {
      float m = S.dy[r];
      float res = (y - S.t[r]) * m;
      S.y[r] = y;
      S.dy[r] = 2.f * res * inv_batch;
      loss += res * res;
    }
    __syncthreads();
    // ---- backward --------------------------------------------------
    {
      float s4sum = 0.f, db = 0.f;
#pragma unroll
      for (int s = 0; s < 16; s++) {
        s4sum += bf2f(S.H3[s * 64 + l]) * S.dy[s];
        db += S.dy[s];
      }
      g4 += s4sum;
      gb4 += db;
    }
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
      const int nn = nt * 16 + c;
      const float wv = bf2f(w.w4[nn]);
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int s = r0 + i;
        float v = S.dy[s] * wv * dsilu(S.Z3[s * 64 + nn]);
        unsigned short b = f2bf(v);
        S.dA[s * 64 + nn] = b;
        S.dAT[nn * 16 + s] = b;
        gb3[nt] += v;
      }
    }
    __syncthreads();
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
      s4 a = frag(S.dAT + (mt * 16 + r) * 16 + kq);
#pragma unroll
      for (int nt = 0; nt < 4; nt++)
        acc3[mt][nt] = mma(a, frag(S.H2T + (nt * 16 + r) * 16 + kq), acc3[mt][nt]);
    }
    dense16<64>(S.dA, w.W3T, acc, l);   // dH2 = dZ3 W3
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
      const int nn = nt * 16 + c;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int s = r0 + i;
        float v = acc[nt][i] * dsilu(S.Z2[s * 64 + nn]);
        unsigned short b = f2bf(v);
        S.dB[s * 64 + nn] = b;
        S.dBT[nn * 16 + s] = b;
        gb2[nt] += v;
      }
    }
    __syncthreads();
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
      s4 a = frag(S.dBT + (mt * 16 + r) * 16 + kq);
#pragma unroll
      for (int nt = 0; nt < 4; nt++)
        acc2[mt][nt] = mma(a, frag(S.H1T + (nt * 16 + r) * 16 + kq), acc2[mt][nt]);
    }
    dense16<64>(S.dB, w.W2T, acc, l);   // dH1 = dZ2 W2
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
      const int nn = nt * 16 + c;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int s = r0 + i;
        float v = acc[nt][i] * dsilu(S.Z1[s * 64 + nn]);
        S.dAT[nn * 16 + s] = f2bf(v);
        gb1[nt] += v;
      }
    }
    __syncthreads();
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
      acc1[mt] = mma(frag(S.dAT + (mt * 16 + r) * 16 + kq), frag(S.XT + r * 16 + kq), acc1[mt]);
    __syncthreads();
  }
  // ---- partial gradient of this workgroup ---------------------------
  float* out = partial + (long long)blockIdx.x * NPP;
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int m = mt * 16 + r0 + i;
#pragma unroll
      for (int nt = 0; nt < 4; nt++) {
        out[OFF_W3 + m * 64 + nt * 16 + c] = acc3[mt][nt][i];
        out[OFF_W2 + m * 64 + nt * 16 + c] = acc2[mt][nt][i];
      }
      if (c < 4) out[OFF_W1 + m * 4 + c] = acc1[mt][i];
    }
#pragma unroll
  for (int nt = 0; nt < 4; nt++) {
    float a1 = gb1[nt], a2 = gb2[nt], a3 = gb3[nt];
    a1 += __shfl_xor(a1, 16); a1 += __shfl_xor(a1, 32);
    a2 += __shfl_xor(a2, 16); a2 += __shfl_xor(a2, 32);
    a3 += __shfl_xor(a3, 16); a3 += __shfl_xor(a3, 32);
    if (l < 16) {
      out[OFF_B1 + nt * 16 + l] = a1;
      out[OFF_B2 + nt * 16 + l] = a2;
      out[OFF_B3 + nt * 16 + l] = a3;
    }
  }
  out[OFF_W4 + l] = g4;
  loss += __shfl_xor(loss, 1); loss += __shfl_xor(loss, 2); loss += __shfl_xor(loss, 4);
  loss += __shfl_xor(loss, 8);
  if (l == 0) {
    out[OFF_B4] = gb4;
    lossp[blockIdx.x] = loss;
  }
}

// Writes the BF16 operand copies of parameter p.
__device__ __forceinline__ void write_bf16(const Weights& w, int p, float v) {
  unsigned short b = f2bf(v);
  auto W = const_cast<unsigned short*>(w.W1p);
  if (p < OFF_B1) {
    int n = p / 4, kk = p % 4;
    W[n * 16 + kk] = b;
  } else if (p >= OFF_W2 && p < OFF_B2) {
    int q = p - OFF_W2, n = q / 64, kk = q % 64;
    const_cast<unsigned short*>(w.W2)[n * 64 + kk] = b;
    const_cast<unsigned short*>(w.W2T)[kk * 64 + n] = b;
  } else if (p >= OFF_W3 && p < OFF_B3) {
    int q = p - OFF_W3, n = q / 64, kk = q % 64;
    const_cast<unsigned short*>(w.W3)[n * 64 + kk] = b;
    const_cast<unsigned short*>(w.W3T)[kk * 64 + n] = b;
  } else if (p >= OFF_W4 && p < OFF_B4) {
    const_cast<unsigned short*>(w.w4)[p - OFF_W4] = b;
  }
}

// Sum of the partial gradients in block order, then either the gradient
// is stored (grad != nullptr) or one Adam step is applied (FP32 master
// weights) and the BF16 copies refreshed.  Thread 0 also records the
// summed loss of the step.
__global__ void mlp_update_kernel(Weights w, float* m1, float* m2, const float* partial,
                                  int nwg, const float* lossp, float* loss_out, float lr,
                                  float bc1, float bc2, float* grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p == 0 && loss_out) {
    float s = 0.f;
    for (int q = 0; q < nwg; q++) s += lossp[q];
    *loss_out = s;
  }
  if (p >= NP) return;
  float g = 0.f;
  for (int q = 0; q < nwg; q++) g += partial[(long long)q * NPP + p];
  if (grad) {
    grad[p] = g;
    return;
  }
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float a = b1 * m1[p] + (1.f - b1) * g;
  float v = b2 * m2[p] + (1.f - b2) * g * g;
  m1[p] = a;
  m2[p] = v;
  float* P = const_cast<float*>(w.P);
  float np = P[p] - lr * (a / bc1) / (sqrtf(v / bc2) + eps);
  P[p] = np;
  write_bf16(w, p, np);
}

__global__ void mlp_sync_kernel(Weights w) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < NP) write_bf16(w, p, w.P[p]);
}

struct alignas(16) FwdLds {
  unsigned short Xb[16 * 16];
  unsigned short H1[16 * 64], H2[16 * 64], H3[16 * 64];
};

// Forward pass over n samples (raw features X[n][4]); Yout = YS * y.
__global__ __launch_bounds__(64) void mlp_predict_kernel(Weights w, const float* X, long long n,
                                                         float* Yout) {
  __shared__ FwdLds S;
  const int l = threadIdx.x, r = l & 15, kq = (l >> 4) * 4;
  const long long s0 = (long long)blockIdx.x * 16;
  const long long sidx = s0 + r;
  const bool valid = sidx < n;
  float xin[4] = {0.f, 0.f, 0.f, 0.f};
  if (valid) load_features(X + 4 * sidx, xin);
#pragma unroll
  for (int f = 0; f < 4; f++) {
    int col = kq + f;
    S.Xb[r * 16 + col] = f2bf((col < 4 && valid) ? xin[col] : 0.f);
  }
  __syncthreads();
  f4 acc[4];
  dense16<16>(S.Xb, w.W1p, acc, l);
  epilogue(acc, w.P + OFF_B1, nullptr, S.H1, nullptr, l);
  __syncthreads();
  dense16<64>(S.H1, w.W2, acc, l);
  epilogue(acc, w.P + OFF_B2, nullptr, S.H2, nullptr, l);
  __syncthreads();
  dense16<64>(S.H2, w.W3, acc, l);
  epilogue(acc, w.P + OFF_B3, nullptr, S.H3, nullptr, l);
  __syncthreads();
  float y = output_unit(S.H3, w, l);
  if (l < 16 && valid) Yout[sidx] = y * YS;
}

// ---------------------------------------------------------------------
// surrogate-cooled SplitSource
// ---------------------------------------------------------------------
using wind::CoolArgs;
using wind::SurrWork;
using wind::CoolBatch;
using wind::find_grid;

// Appends the lanes with `pred` to a list whose tail is *ctr, and adds
// their number to the running total *total (if given): one atomic each per
// wavefront.  Every lane of the wavefront must call it.
__device__ __forceinline__ long long wave_append(unsigned long long* ctr, bool pred,
                                                 unsigned long long* total = nullptr) {
  const unsigned long long mask = __ballot(pred);
  if (mask == 0ull) return -1;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((long long)mask) - 1;
  unsigned long long base = 0ull;
  if (lane == leader) {
    base = atomicAdd(ctr, (unsigned long long)__popcll(mask));
    if (total) atomicAdd(total, (unsigned long long)__popcll(mask));
  }
  base = __shfl(base, leader);
  return pred ? (long long)(base + __popcll(mask & ((1ull << lane) - 1ull))) : -1;
}

__device__ __forceinline__ RadCell cell_state(const CoolArgs& a, long long cell) {
  RadCell R;
  R.rho = a.V[RHO * a.sv + cell];
  R.trc = a.V[wind::TRC * a.sv + cell];
  R.x1 = a.x1[(int)(cell % a.nr) + 2];
  R.xi_s = a.xi_s[cell];
  R.tr_s = a.tr_s[cell];
  wind::load_pre(a.pre, a.sv, cell, R.pre);
  R.t = a.s->t;
  R.nev = 0;
  return R;
}

// dt_cool candidates and fallback flags of lanes that may belong to
// different grids: one reduction per wavefront when all its lanes share a
// grid, per-lane atomics otherwise.
__device__ __forceinline__ void reduce_dtcool_any(double cand, int bad, wind::Scalars* s) {
  const unsigned long long key = (unsigned long long)s;
  const unsigned long long k0 = __shfl(key, 0);
  if (__all(key == k0)) {
    for (int off = 32; off > 0; off >>= 1) {
      cand = dmin(cand, __shfl_xor(cand, off));
      bad |= __shfl_xor(bad, off);
    }
    if ((threadIdx.x & 63) == 0 && s) {
      if (cand >= 0.0 && cand < 1.0e300)
        atomicMin(&s->dtcool_red[red_index(blockIdx.x)],
                  (unsigned long long)__double_as_longlong(cand));
      if (bad) atomicOr(&s->err, PMX_FLAG_COOL_SUBSTEPS);
    }
  } else if (s) {
    if (cand >= 0.0 && cand < 1.0e300)
      atomicMin(&s->dtcool_red[red_index(blockIdx.x)],
                (unsigned long long)__double_as_longlong(cand));
    if (bad) atomicOr(&s->err, PMX_FLAG_COOL_SUBSTEPS);
  }
}

// Ends a call of the surrogate path: clears the list tails cnt[0] and
// cnt[1] for the next call (one workgroup, after the call's last kernel;
// a last-workgroup ticket in that kernel costs more than this launch).
__global__ void surr_end_call_kernel(unsigned long long* cnt) {
  if (threadIdx.x < 2) cnt[threadIdx.x] = 0ull;
}

// Adds the workgroup's Radiat call counts to its spread slot of the
// counter slots at ctr: one device atomic per workgroup.  Every thread of
// the workgroup must call it.
__device__ __forceinline__ void count_calls(unsigned long long* ctr, int n) {
  __shared__ unsigned long long sc;
  unsigned long long v = (unsigned long long)n;
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if (threadIdx.x == 0) sc = 0ull;
  __syncthreads();
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(&sc, v);
  __syncthreads();
  if (threadIdx.x == 0 && sc) atomicAdd(ctr + red_index(blockIdx.x), sc);
}

// Non-stiff cells (|dE/dt| <= 0.5 E/dt) take the reference integrator here
// and keep its exact result; stiff cells inside the training hull are
// listed for the network, the others queued for the plain integrator.
// Threads run over the global cell index of all grids in `b`.
__global__ __launch_bounds__(128) void surr_classify_kernel(CoolBatch b, SurrWork w) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double cand = 1.0e300;
  int bad = 0;
  bool to_list = false, to_fb = false;
  float f[4] = {0.f, 0.f, 0.f, 0.f};
  double k1 = 0.0;
  int ncalls = 0;
  wind::Scalars* sg = nullptr;
  if (q < b.total) {
    const int g = find_grid(b.off, b.G, q);
    const CoolArgs& a = b.args[g];
    const long long cell = q - b.off[g];
    sg = a.s;
    if (cell < a.sv && !(b.loop && a.s->halt)) {
      const Const& k = a.k;
      const double dt = a.dtc ? a.dtc[cell] : a.s->dt, t = a.s->t;
      RadCell R = cell_state(a, cell);
      const double p = a.V[PRS * a.sv + cell];
      double E0 = p / (k.gamma - 1.0);
      k1 = wind::radiat(k, R, E0);
      if (!(fabs(k1) > 0.5 / dt)) {
        // the integrator continues from this first rate: a non-stiff cell
        // costs what the plain kernel's does
        a.V[PRS * a.sv + cell] = wind::plain_cell_from(k, R, p, dt, E0, k1, cand, bad);
        ncalls = R.nev;
      } else {
        ncalls = R.nev;
        double xi, tx;
        if (t < 3.0) {
          const double nH = R.rho * k.UD / (1.43 * wind::C_mp), r = R.x1 * k.UL;
          xi = k.lx / nH / r / r;
          tx = k.t_x;
        } else {
          xi = R.xi_s;
          tx = R.tr_s;
        }
        const double T0 = p / R.rho * k.KELVIN * k.mu;
        const double x[4] = {log10(R.rho * k.UD), log10(T0), log10(xi),
                             log10(dt / (E0 / fabs(k1)))};
        // the network is trained on unit prefactors and T_r = T_x
        bool in = tx == k.t_x && E0 > 0.0 && R.pre[0] == 1.0 && R.pre[1] == 1.0 &&
                  R.pre[2] == 1.0 && R.pre[3] == 1.0 && R.pre[4] == 1.0;
        for (int c = 0; c < 4; c++) {
          in = in && x[c] >= HULL_LO[c] && x[c] <= HULL_HI[c];
          f[c] = (float)x[c];
        }
        to_list = in;
        to_fb = !in;
      }
    }
  }
  const long long li = wave_append(&w.cnt[0], to_list, &w.cnt[4]);
  const long long fi = wave_append(&w.cnt[1], to_fb, &w.cnt[5]);
  if (to_list) {
    w.list[li] = q;
    for (int c = 0; c < 4; c++) w.feat[4 * li + c] = f[c];
    w.f0[li] = k1;
  }
  if (to_fb) w.fb[fi] = q;
  count_calls(w.cnt + wind::SURR_CALLS, ncalls);
  reduce_dtcool_any(cand, bad, sg);
}

__device__ __forceinline__ void surr_mlp_group(const Weights& wt, const SurrWork& w, FwdLds& S,
                                               long long n, long long s0, float* ylds) {
  const int l = threadIdx.x, r = l & 15, kq = (l >> 4) * 4;
  const long long sidx = s0 + r;
  const bool valid = sidx < n;
  float xin[4] = {0.f, 0.f, 0.f, 0.f};
  if (valid) load_features(w.feat + 4 * sidx, xin);
#pragma unroll
  for (int f = 0; f < 4; f++) {
    int col = kq + f;
    S.Xb[r * 16 + col] = f2bf((col < 4 && valid) ? xin[col] : 0.f);
  }
  __syncthreads();
  f4 acc[4];
  dense16<16>(S.Xb, wt.W1p, acc, l);
  epilogue(acc, wt.P + OFF_B1, nullptr, S.H1, nullptr, l);
  __syncthreads();
  dense16<64>(S.H1, wt.W2, acc, l);
  epilogue(acc, wt.P + OFF_B2, nullptr, S.H2, nullptr, l);
  __syncthreads();
  dense16<64>(S.H2, wt.W3, acc, l);
  epilogue(acc, wt.P + OFF_B3, nullptr, S.H3, nullptr, l);
  __syncthreads();
  float y = output_unit(S.H3, wt, l);
  if (l < 16) ylds[r] = y * YS;
  __syncthreads();   // S is reused by the next group
}

// Newton refinement of the network's guess yhat; failures join the queue.
__device__ __forceinline__ void surr_refine_cell(const CoolBatch& b, const SurrWork& w,
                                                 long long qi, long long n, float yhat) {
  double cand = 1.0e300;
  bool fail = false, ok = false;
  long long q = 0;
  int ncalls = 0;
  wind::Scalars* sg = nullptr;
  if (qi < n) {
    q = w.list[qi];
    const int g = find_grid(b.off, b.G, q);
    const CoolArgs& a = b.args[g];
    const long long cell = q - b.off[g];
    sg = a.s;
    const Const& k = a.k;
    RadCell R = cell_state(a, cell);
    const double p = a.V[PRS * a.sv + cell], dt = a.dtc ? a.dtc[cell] : a.s->dt;
    const double E0 = p / (k.gamma - 1.0);
    // guess: the network's, or (ablation) the explicit step limited to one
    // e-fold, which is what a solver without the network starts from
    const double f0 = w.f0[qi];
    const double Eg = w.guess == 0 ? E0 * pow(10.0, (double)yhat)
                                   : E0 * exp(fmax(-1.0, fmin(1.0, f0 * dt / E0)));
    double Ef;
    if (wind::refine_cell(k, R, E0, f0, dt, Eg, Ef) > 0) {
      a.V[PRS * a.sv + cell] = wind::finish_cell(k, R, p, Ef, dt, cand);
      ok = true;
    } else {
      fail = true;
    }
    ncalls = R.nev;
  }
  const long long fi = wave_append(&w.cnt[1], fail, &w.cnt[5]);
  if (fail) w.fb[fi] = q;
  wave_append(&w.cnt[6], ok);
  count_calls(w.cnt + wind::SURR_CALLS, ncalls);
  reduce_dtcool_any(cand, 0, sg);
}

// The listed cells, 64 per wavefront and pass: the network's forward pass
// (four MFMA groups of 16 cells), then one cell per lane refined from its
// guess.  The grid strides over the list, whose length is known only on
// the device; the loop bound is uniform over the (one-wavefront)
// workgroup, so every lane takes part in the wavefront-wide appends and
// reductions of every pass.
__global__ __launch_bounds__(64) void surr_mlp_refine_kernel(Weights wt, CoolBatch b, SurrWork w) {
  __shared__ FwdLds S;
  __shared__ float ylds[64];
  const long long n = (long long)w.cnt[0];
  for (long long s0 = (long long)blockIdx.x * 64; s0 < n; s0 += (long long)gridDim.x * 64) {
    if (w.guess == 0) {
#pragma unroll 1
      for (int g = 0; g < 4 && s0 + 16 * g < n; g++)
        surr_mlp_group(wt, w, S, n, s0 + 16 * g, ylds + 16 * g);
    }
    surr_refine_cell(b, w, s0 + threadIdx.x, n, w.guess == 0 ? ylds[threadIdx.x] : 0.f);
    __syncthreads();   // ylds is rewritten by the next pass
  }
}

// The plain integrator for the queued cells.
__device__ __forceinline__ void surr_fallback_cell(const CoolBatch& b, const SurrWork& w,
                                                   long long qi, long long n) {
  double cand = 1.0e300;
  int bad = 0, ncalls = 0;
  wind::Scalars* sg = nullptr;
  if (qi < n) {
    const long long q = w.fb[qi];
    const int g = find_grid(b.off, b.G, q);
    const CoolArgs& a = b.args[g];
    const long long cell = q - b.off[g];
    sg = a.s;
    RadCell R = cell_state(a, cell);
    const double p = a.V[PRS * a.sv + cell];
    a.V[PRS * a.sv + cell] = wind::plain_cell(a.k, R, p, a.dtc ? a.dtc[cell] : a.s->dt, cand, bad);
    ncalls = R.nev;
  }
  count_calls(w.cnt + wind::SURR_CALLS, ncalls);
  reduce_dtcool_any(cand, bad, sg);
}

__global__ __launch_bounds__(128) void surr_fallback_kernel(CoolBatch b, SurrWork w) {
  const long long n = (long long)w.cnt[1];
  for (long long base = (long long)blockIdx.x * blockDim.x; base < n;
       base += (long long)gridDim.x * blockDim.x)
    surr_fallback_cell(b, w, base + threadIdx.x, n);
}

// The plain integrator on every cell of every grid in `b`.
__global__ __launch_bounds__(128) void plain_cells_kernel(CoolBatch b, unsigned long long* calls) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double cand = 1.0e300;
  int bad = 0, ncalls = 0;
  wind::Scalars* sg = nullptr;
  if (q < b.total) {
    const int g = find_grid(b.off, b.G, q);
    const CoolArgs& a = b.args[g];
    const long long cell = q - b.off[g];
    sg = a.s;
    if (cell < a.sv && !(b.loop && a.s->halt)) {
      RadCell R = cell_state(a, cell);
      const double p = a.V[PRS * a.sv + cell];
      a.V[PRS * a.sv + cell] =
          wind::plain_cell(a.k, R, p, a.dtc ? a.dtc[cell] : a.s->dt, cand, bad);
      ncalls = R.nev;
    }
  }
  count_calls(calls, ncalls);
  reduce_dtcool_any(cand, bad, sg);
}

}  // namespace mlp
}  // namespace pmx

using namespace pmx;
using namespace pmx::mlp;

struct pmx_mlp {
  float* P;            // FP32 parameters
  float* m1;
  float* m2;
  unsigned short* bf;  // BF16 operand copies
  Weights w;
  float* X;            // dataset [n][4]
  float* Y;
  int* nev;
  double* aux;      // per sample: rho (code), p0 (code), xi, dt (code)
  long long n;
  float* partial;
  float* lossp;
  float* loss_hist;
  int max_wg;
  long long adam_t;
  hipStream_t stream;
};

namespace {

int ensure_partials(pmx_mlp* m, int nwg) {
  if (nwg <= m->max_wg) return 0;
  hipFree(m->partial);
  hipFree(m->lossp);
  m->partial = nullptr;
  m->lossp = nullptr;
  PMX_HIP_CHECK(hipMalloc(&m->partial, sizeof(float) * (size_t)NPP * nwg));
  PMX_HIP_CHECK(hipMalloc(&m->lossp, sizeof(float) * nwg));
  m->max_wg = nwg;
  return 0;
}

int ensure_dataset(pmx_mlp* m, long long n) {
  if (n <= m->n && m->X) return 0;
  hipFree(m->X); hipFree(m->Y); hipFree(m->nev); hipFree(m->aux);
  m->X = nullptr; m->Y = nullptr; m->nev = nullptr; m->aux = nullptr;
  PMX_HIP_CHECK(hipMalloc(&m->aux, sizeof(double) * 4 * n));
  PMX_HIP_CHECK(hipMalloc(&m->X, sizeof(float) * 4 * n));
  PMX_HIP_CHECK(hipMalloc(&m->Y, sizeof(float) * n));
  PMX_HIP_CHECK(hipMalloc(&m->nev, sizeof(int) * n));
  m->n = n;
  return 0;
}

}  // namespace

namespace pmx {
namespace mlp {
int plain_cool(const wind::CoolBatch& b, unsigned long long* calls, hipStream_t st) {
  hipLaunchKernelGGL(plain_cells_kernel, dim3((unsigned)((b.total + 127) / 128)), dim3(128), 0, st,
                     b, calls);
  PMX_HIP_CHECK(hipGetLastError());
  return 0;
}

constexpr int SURR_GRID = 2048;

int surrogate_cool(pmx_mlp* m, const wind::CoolBatch& b, const wind::SurrWork& w,
                   hipStream_t st) {
  const unsigned nb = (unsigned)((b.total + 127) / 128);
  // the list and queue lengths are known only on the device: those kernels
  // stride over them with at most SURR_GRID workgroups
  const unsigned nr = std::min(nb, (unsigned)SURR_GRID);
  const unsigned nm = (unsigned)std::min((b.total + 63) / 64, (long long)SURR_GRID);
  hipLaunchKernelGGL(surr_classify_kernel, dim3(nb), dim3(128), 0, st, b, w);
  hipLaunchKernelGGL(surr_mlp_refine_kernel, dim3(nm), dim3(64), 0, st, m->w, b, w);
  hipLaunchKernelGGL(surr_fallback_kernel, dim3(nr), dim3(128), 0, st, b, w);
  hipLaunchKernelGGL(surr_end_call_kernel, dim3(1), dim3(64), 0, st, w.cnt);
  PMX_HIP_CHECK(hipGetLastError());
  return 0;
}
}  // namespace mlp
}  // namespace pmx

extern "C" {

int pmx_mlp_destroy(pmx_mlp* m) {
  if (!m) return 0;
  if (m->stream) hipStreamSynchronize(m->stream);
  hipFree(m->P); hipFree(m->m1); hipFree(m->m2); hipFree(m->bf);
  hipFree(m->X); hipFree(m->Y); hipFree(m->nev); hipFree(m->aux);
  hipFree(m->partial); hipFree(m->lossp); hipFree(m->loss_hist);
  if (m->stream) hipStreamDestroy(m->stream);
  delete m;
  return 0;
}

int pmx_mlp_create(unsigned long long seed, pmx_mlp** out) {
  if (!out) { set_error("null argument"); return -1; }
  pmx_mlp* m = new pmx_mlp();
  memset(m, 0, sizeof(*m));
#define CK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { \
    set_error(std::string(#x) + ": " + hipGetErrorString(e_)); pmx_mlp_destroy(m); return -1; } } while (0)
  CK(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
  CK(hipMalloc(&m->P, sizeof(float) * NP));
  CK(hipMalloc(&m->m1, sizeof(float) * NP));
  CK(hipMalloc(&m->m2, sizeof(float) * NP));
  const size_t nbf = 64 * 16 + 4 * 64 * 64 + 64;
  CK(hipMalloc(&m->bf, sizeof(unsigned short) * nbf));
  CK(hipMemset(m->bf, 0, sizeof(unsigned short) * nbf));
  CK(hipMemset(m->m1, 0, sizeof(float) * NP));
  CK(hipMemset(m->m2, 0, sizeof(float) * NP));
  CK(hipMalloc(&m->loss_hist, sizeof(float) * 1));
  m->w.P = m->P;
  m->w.W1p = m->bf;
  m->w.W2 = m->bf + 64 * 16;
  m->w.W3 = m->w.W2 + 64 * 64;
  m->w.W2T = m->w.W3 + 64 * 64;
  m->w.W3T = m->w.W2T + 64 * 64;
  m->w.w4 = m->w.W3T + 64 * 64;
  // He-style uniform initialisation with a fixed-seed generator (host)
  std::vector<float> P(NP, 0.f);
  unsigned long long st = seed * 0x9E3779B97F4A7C15ull + 1ull;
  auto next = [&st]() {
    st ^= st << 13; st ^= st >> 7; st ^= st << 17;
    return (double)(st >> 11) * (1.0 / 9007199254740992.0);
  };
  auto fill = [&](int off, int cnt, int fan_in) {
    double a = sqrt(6.0 / fan_in);
    for (int q = 0; q < cnt; q++) P[off + q] = (float)((2.0 * next() - 1.0) * a);
  };
  fill(OFF_W1, 256, 4);
  fill(OFF_W2, 4096, 64);
  fill(OFF_W3, 4096, 64);
  fill(OFF_W4, 64, 64);
  CK(hipMemcpy(m->P, P.data(), sizeof(float) * NP, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(mlp_sync_kernel, dim3((NP + 255) / 256), dim3(256), 0, m->stream, m->w);
  CK(hipGetLastError());
  CK(hipStreamSynchronize(m->stream));
#undef CK
  *out = m;
  return 0;
}

int pmx_mlp_set_params(pmx_mlp* m, const float* P) {
  if (!m || !P) { set_error("null argument"); return -1; }
  PMX_HIP_CHECK(hipMemcpyAsync(m->P, P, sizeof(float) * NP, hipMemcpyHostToDevice, m->stream));
  hipLaunchKernelGGL(mlp_sync_kernel, dim3((NP + 255) / 256), dim3(256), 0, m->stream, m->w);
  PMX_HIP_CHECK(hipGetLastError());
  PMX_HIP_CHECK(hipMemsetAsync(m->m1, 0, sizeof(float) * NP, m->stream));
  PMX_HIP_CHECK(hipMemsetAsync(m->m2, 0, sizeof(float) * NP, m->stream));
  m->adam_t = 0;
  PMX_HIP_CHECK(hipStreamSynchronize(m->stream));
  return 0;
}

int pmx_mlp_get_params(pmx_mlp* m, float* P) {
  if (!m || !P) { set_error("null argument"); return -1; }
  PMX_HIP_CHECK(hipMemcpyAsync(P, m->P, sizeof(float) * NP, hipMemcpyDeviceToHost, m->stream));
  PMX_HIP_CHECK(hipStreamSynchronize(m->stream));
  return 0;
}

int pmx_mlp_sample(pmx_mlp* m, const pmx_wind_desc* d, long long n, unsigned long long seed) {
  if (!m || !d || n <= 0) { set_error("bad argument"); return -1; }
  if (ensure_dataset(m, n)) return -1;
  Const k = wind::make_const(d);
  hipLaunchKernelGGL(mlp_sample_kernel, dim3((unsigned)((n + 127) / 128)), dim3(128), 0,
                     m->stream, k, n, seed, m->X, m->Y, m->nev, m->aux);
  PMX_HIP_CHECK(hipGetLastError());
  PMX_HIP_CHECK(hipStreamSynchronize(m->stream));
  m->n = n;
  return 0;
}

int pmx_mlp_set_dataset(pmx_mlp* m, const float* X, const float* Y, long long n) {
  if (!m || !X || !Y || n <= 0) { set_error("bad argument"); return -1; }
  if (ensure_dataset(m, n)) return -1;
  PMX_HIP_CHECK(hipMemcpyAsync(m->X, X, sizeof(float) * 4 * n, hipMemcpyHostToDevice, m->stream));
  PMX_HIP_CHECK(hipMemcpyAsync(m->Y, Y, sizeof(float) * n, hipMemcpyHostToDevice, m->stream));
  PMX_HIP_CHECK(hipMemsetAsync(m->nev, 0, sizeof(int) * n, m->stream));
  PMX_HIP_CHECK(hipMemsetAsync(m->aux, 0, sizeof(double) * 4 * n, m->stream));
  PMX_HIP_CHECK(hipStreamSynchronize(m->stream));
  m->n = n;
  return 0;
}

int pmx_mlp_get_dataset(pmx_mlp* m, float* X, float* Y, int* nev, double* aux, long long n) {
  if (!m || n > m->n) { set_error("bad argument"); return -1; }
  if (aux) PMX_HIP_CHECK(hipMemcpyAsync(aux, m->aux, sizeof(double) * 4 * n, hipMemcpyDeviceToHost, m->stream));
  if (X) PMX_HIP_CHECK(hipMemcpyAsync(X, m->X, sizeof(float) * 4 * n, hipMemcpyDeviceToHost, m->stream));
  if (Y) PMX_HIP_CHECK(hipMemcpyAsync(Y, m->Y, sizeof(float) * n, hipMemcpyDeviceToHost, m->stream));
  if (nev) PMX_HIP_CHECK(hipMemcpyAsync(nev, m->nev, sizeof(int) * n, hipMemcpyDeviceToHost, m->stream));
  PMX_HIP_CHECK(hipStreamSynchronize(m->stream));
  return 0;
}

// Loss and gradient of the first `count` dataset samples (one batch).
int pmx_mlp_loss_grad(pmx_mlp* m, long long count, float* loss, float* grad) {
  if (!m || count <= 0 || count > m->n || !grad) { set_error("bad argument"); return -1; }
  const int tiles = 4;
  const int nwg = (int)((count + 16 * tiles - 1) / (16 * tiles));
  if (ensure_partials(m, nwg)) return -1;
  float* dgrad = nullptr;
  PMX_HIP_CHECK(hipMalloc(&dgrad, sizeof(float) * (NP + 1)));
  hipLaunchKernelGGL(mlp_train_kernel, dim3(nwg), dim3(64), 0, m->stream, m->w, m->X, m->Y,
                     m->n, 0ll, tiles, 1.f / (float)count, count, m->partial, m->lossp);
  hipLaunchKernelGGL(mlp_update_kernel, dim3((NP + 255) / 256), dim3(256), 0, m->stream, m->w,
                     m->m1, m->m2, m->partial, nwg, m->lossp, dgrad + NP, 0.f, 1.f, 1.f, dgrad);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(grad, dgrad, sizeof(float) * NP, hipMemcpyDeviceToHost, m->stream);
  float l = 0.f;
  if (e == hipSuccess) e = hipMemcpyAsync(&l, dgrad + NP, sizeof(float), hipMemcpyDeviceToHost, m->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  hipFree(dgrad);
  if (e != hipSuccess) { set_error(hipGetErrorString(e)); return -1; }
  if (loss) *loss = l / (float)count;
  return 0;
}

// `steps` Adam steps over the dataset with batches of `batch` samples
// (a multiple of 64), learning rate decaying as a cosine from lr0 to lr1.
// loss_hist (optional, `steps` entries) receives each step's mean loss.
int pmx_mlp_train(pmx_mlp* m, int steps, int batch, float lr0, float lr1, float* loss_hist) {
  if (!m || steps <= 0 || batch <= 0 || batch % 64 || !m->X) { set_error("bad argument"); return -1; }
  const int tiles = 4;
  const int nwg = batch / (16 * tiles);
  if (ensure_partials(m, nwg)) return -1;
  float* dl = nullptr;
  PMX_HIP_CHECK(hipMalloc(&dl, sizeof(float) * steps));
  long long start = (m->adam_t * (long long)batch) % m->n;
  for (int s = 0; s < steps; s++) {
    m->adam_t++;
    const double ph = steps > 1 ? (double)s / (steps - 1) : 0.0;
    const float lr = (float)(lr1 + 0.5 * (lr0 - lr1) * (1.0 + cos(3.14159265358979 * ph)));
    const float bc1 = (float)(1.0 - pow(0.9, (double)m->adam_t));
    const float bc2 = (float)(1.0 - pow(0.999, (double)m->adam_t));
    hipLaunchKernelGGL(mlp_train_kernel, dim3(nwg), dim3(64), 0, m->stream, m->w, m->X, m->Y,
                       m->n, start, tiles, 1.f / (float)batch, (long long)batch, m->partial,
                       m->lossp);
    hipLaunchKernelGGL(mlp_update_kernel, dim3((NP + 255) / 256), dim3(256), 0, m->stream,
                       m->w, m->m1, m->m2, m->partial, nwg, m->lossp, dl + s, lr, bc1, bc2,
                       (float*)nullptr);
    start = (start + batch) % m->n;
  }
  hipError_t e = hipGetLastError();
  std::vector<float> h(steps);
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), dl, sizeof(float) * steps, hipMemcpyDeviceToHost, m->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  hipFree(dl);
  if (e != hipSuccess) { set_error(hipGetErrorString(e)); return -1; }
  if (loss_hist)
    for (int s = 0; s < steps; s++) loss_hist[s] = h[s] / (float)batch;
  return 0;
}

int pmx_mlp_cool_cells(pmx_mlp* m, const pmx_wind_desc* d, long long n, const double* rho,
                       const double* p, const double* xi, const double* dt, double* p_out,
                       int mode, long long* stats, double* ms) {
  if (!m || !d || n <= 0 || !rho || !p || !xi || !dt || !p_out || mode < 0 || mode > 2) {
    set_error("bad argument");
    return -1;
  }
  const Const k = wind::make_const(d);
  std::vector<double> V((size_t)6 * n, 0.0), tr((size_t)n, d->t_x), x1(4, 1.0);
  for (long long q = 0; q < n; q++) {
    V[RHO * n + q] = rho[q];
    V[PRS * n + q] = p[q];
    V[wind::TRC * n + q] = 1.0;
  }
  double *dV = nullptr, *dxi = nullptr, *dtr = nullptr, *ddt = nullptr, *dx1 = nullptr;
  wind::Scalars* ds = nullptr;
  wind::SurrWork sw;
  memset(&sw, 0, sizeof(sw));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipSuccess;
  auto A = [&](void** ptr, size_t bytes) { if (e == hipSuccess) e = hipMalloc(ptr, bytes); };
  A((void**)&dV, sizeof(double) * 6 * n);
  A((void**)&dxi, sizeof(double) * n);
  A((void**)&dtr, sizeof(double) * n);
  A((void**)&ddt, sizeof(double) * n);
  A((void**)&dx1, sizeof(double) * 4);
  A((void**)&ds, sizeof(wind::Scalars));
  A((void**)&sw.list, sizeof(long long) * n);
  A((void**)&sw.fb, sizeof(long long) * n);
  A((void**)&sw.feat, sizeof(float) * 4 * n);
  A((void**)&sw.f0, sizeof(double) * n);
  A((void**)&sw.cnt, sizeof(unsigned long long) * PMX_SURR_COUNTERS);
  wind::CoolArgs* dargs = nullptr;
  long long* doff = nullptr;
  A((void**)&dargs, sizeof(wind::CoolArgs));
  A((void**)&doff, sizeof(long long) * 2);
  wind::Scalars s0;
  memset(&s0, 0, sizeof(s0));
  wind::red_clear(s0);
  s0.tstop = HUGE_VAL;
  s0.t = 5.0;    // g_time > 3: Radiat takes xi and T_r from the per-cell arrays
  s0.dt = 0.0;
  {
    double big = 1.e38;
    memcpy(&s0.dtcool_bits, &big, sizeof(double));
  }
  if (e == hipSuccess) e = hipMemcpy(dV, V.data(), sizeof(double) * 6 * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dxi, xi, sizeof(double) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dtr, tr.data(), sizeof(double) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(ddt, dt, sizeof(double) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dx1, x1.data(), sizeof(double) * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(ds, &s0, sizeof(s0), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(sw.cnt, 0, sizeof(unsigned long long) * PMX_SURR_COUNTERS);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) {
    wind::CoolArgs a;
    a.V = dV; a.nr = 1; a.nt = (int)n; a.sv = n; a.k = k; a.x1 = dx1; a.xi_s = dxi; a.tr_s = dtr;
    a.pre = nullptr;
    a.s = ds; a.dtc = ddt; a.loop = 0;
    const long long off[2] = {0, n};
    e = hipMemcpy(dargs, &a, sizeof(a), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(doff, off, sizeof(off), hipMemcpyHostToDevice);
    wind::CoolBatch b;
    b.args = dargs; b.off = doff; b.G = 1; b.total = n; b.loop = 0;
    sw.guess = mode == 2 ? 1 : 0;
    if (e == hipSuccess) e = hipEventRecord(e0, m->stream);
    if (e == hipSuccess) {
      if (mode == 0)
        hipLaunchKernelGGL(plain_cells_kernel, dim3((unsigned)((n + 127) / 128)), dim3(128), 0,
                           m->stream, b, sw.cnt + wind::SURR_CALLS);
      else if (surrogate_cool(m, b, sw, m->stream))
        e = hipErrorUnknown;
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(e1, m->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
  }
  float fms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&fms, e0, e1);
  if (e == hipSuccess) e = hipMemcpy(V.data(), dV, sizeof(double) * 6 * n, hipMemcpyDeviceToHost);
  unsigned long long c[PMX_SURR_COUNTERS] = {};
  if (e == hipSuccess) e = hipMemcpy(c, sw.cnt, sizeof(c), hipMemcpyDeviceToHost);
  hipFree(dV); hipFree(dxi); hipFree(dtr); hipFree(ddt); hipFree(dx1); hipFree(ds);
  hipFree(sw.list); hipFree(sw.fb); hipFree(sw.feat); hipFree(sw.f0);
  hipFree(sw.cnt); hipFree(dargs); hipFree(doff);
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if (e != hipSuccess) { set_error(hipGetErrorString(e)); return -1; }
  for (long long q = 0; q < n; q++) p_out[q] = V[PRS * n + q];
  if (stats) {
    stats[0] = (long long)c[4]; stats[1] = (long long)c[6]; stats[2] = (long long)c[5];
    stats[3] = (long long)wind::surr_calls(c);
  }
  if (ms) *ms = fms;
  return 0;
}

int pmx_mlp_predict(pmx_mlp* m, const float* X, long long n, float* Y) {
  if (!m || !X || !Y || n <= 0) { set_error("bad argument"); return -1; }
  float *dx = nullptr, *dy = nullptr;
  PMX_HIP_CHECK(hipMalloc(&dx, sizeof(float) * 4 * n));
  hipError_t e = hipMalloc(&dy, sizeof(float) * n);
  if (e == hipSuccess) e = hipMemcpyAsync(dx, X, sizeof(float) * 4 * n, hipMemcpyHostToDevice, m->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(mlp_predict_kernel, dim3((unsigned)((n + 15) / 16)), dim3(64), 0,
                       m->stream, m->w, dx, n, dy);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(Y, dy, sizeof(float) * n, hipMemcpyDeviceToHost, m->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  hipFree(dx);
  hipFree(dy);
  if (e != hipSuccess) { set_error(hipGetErrorString(e)); return -1; }
  return 0;
}

}  // extern "C"
