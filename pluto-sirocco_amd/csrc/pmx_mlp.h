// Device helpers of the MLP cooling surrogate: parameter layout, BF16
// conversion, matrix-core fragments and the dense layers of a 16-sample
// tile.  One 64-lane wavefront owns a tile.
//
// v_mfma_f32_16x16x16_bf16 operand layout (per lane l, r = l % 16,
// q = l / 16): A[r][4q .. 4q+3], B[4q .. 4q+3][r], and the accumulator
// holds D[4q + i][r] for i = 0..3.
#pragma once
#include <hip/hip_runtime.h>

namespace pmx {
namespace mlp {

constexpr int NH = 64;
// FP32 parameter vector: W1[64][4] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] w4[64] b4
constexpr int OFF_W1 = 0, OFF_B1 = 256, OFF_W2 = 320, OFF_B2 = 4416, OFF_W3 = 4480,
              OFF_B3 = 8576, OFF_W4 = 8640, OFF_B4 = 8704, NP = 8705;
constexpr int NPP = 8712;   // row stride of the per-workgroup partial gradients

// Training hull (log10): rho [g/cm^3], T [K], xi, dt / t_cool
constexpr double HULL_LO[4] = {-16.0, 4.0, 0.0, -2.0};
constexpr double HULL_HI[4] = {-10.0, 8.0, 6.0, 2.0};
constexpr float YS = 2.f;   // output scale: log10(T_f / T_0) = YS * y

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

struct Weights {
  const float* P;               // FP32 master parameters (biases are read here)
  const unsigned short* W1p;    // [64][16] BF16, inputs 4..15 zero
  const unsigned short* W2;     // [64 out][64 in]
  const unsigned short* W3;
  const unsigned short* W2T;    // [64 in][64 out]
  const unsigned short* W3T;
  const unsigned short* w4;     // [64]
};

__device__ __forceinline__ unsigned short f2bf(float f) {   // round to nearest even
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float(((unsigned)b) << 16); }

__device__ __forceinline__ f4 mma(s4 a, s4 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ s4 frag(const unsigned short* p) {
  return *reinterpret_cast<const s4*>(p);
}

__device__ __forceinline__ float sigm(float z) { return 1.f / (1.f + expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
__device__ __forceinline__ float dsilu(float z) {
  float s = sigm(z);
  return s * (1.f + z * (1.f - s));
}

// raw features (log10 values) -> network inputs in [-1, 1] over the hull
__device__ __forceinline__ void load_features(const float* x, float* xn) {
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const float c = (float)(0.5 * (HULL_LO[q] + HULL_HI[q]));
    const float h = (float)(0.5 * (HULL_HI[q] - HULL_LO[q]));
    xn[q] = (x[q] - c) / h;
  }
}

// acc[nt] (16 samples x 16 outputs nt*16..) = in[16][K] . W[64][K]^T
template <int K>
__device__ __forceinline__ void dense16(const unsigned short* in, const unsigned short* W, f4* acc,
                                        int l) {
  const int r = l & 15, kq = (l >> 4) * 4;
#pragma unroll
  for (int nt = 0; nt < 4; nt++) acc[nt] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < K; ks += 16) {
    s4 a = frag(in + r * K + ks + kq);
#pragma unroll
    for (int nt = 0; nt < 4; nt++) acc[nt] = mma(a, frag(W + (nt * 16 + r) * K + ks + kq), acc[nt]);
  }
}

// bias + SiLU; stores the pre-activation (optional), the BF16 activation
// [16][64] and its transpose [64][16] (optional)
__device__ __forceinline__ void epilogue(const f4* acc, const float* b, float* Z, unsigned short* H,
                                         unsigned short* HT, int l) {
  const int c = l & 15, r0 = (l >> 4) * 4;
#pragma unroll
  for (int nt = 0; nt < 4; nt++) {
    const int n = nt * 16 + c;
    const float bn = b[n];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int s = r0 + i;
      const float z = acc[nt][i] + bn;
      if (Z) Z[s * 64 + n] = z;
      const unsigned short hb = f2bf(silu(z));
      H[s * 64 + n] = hb;
      if (HT) HT[n * 16 + s] = hb;
    }
  }
}

// y[s] = H3[s] . w4 + b4 for s = l % 16 (all four lanes of a sample agree)
__device__ __forceinline__ float output_unit(const unsigned short* H3, const Weights& w, int l) {
  const int s = l & 15, k0 = (l >> 4) * 16;
  float sum = 0.f;
#pragma unroll
  for (int k = 0; k < 16; k++) sum += bf2f(H3[s * 64 + k0 + k]) * bf2f(w.w4[k0 + k]);
  sum += __shfl_xor(sum, 16);
  sum += __shfl_xor(sum, 32);
  return sum + w.P[OFF_B4];
}

}  // namespace mlp
}  // namespace pmx
