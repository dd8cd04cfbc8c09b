// Shared host/device helpers of libplutomx.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdio>
#include <string>

namespace pmx {

// Last error message (per thread), returned by pmx_last_error().
void set_error(const std::string& msg);

#define PMX_HIP_CHECK(call)                                                        \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) {                                                        \
      ::pmx::set_error(std::string(#call) + ": " + hipGetErrorString(e_));         \
      return -1;                                                                   \
    }                                                                              \
  } while (0)

enum { BC_PERIODIC = 0, BC_OUTFLOW = 1, BC_REFLECTIVE = 2 };

// Ghost-index mapping of the reference's Boundary() for one direction
// (Src/boundary.c:617-763): periodic wraps, outflow copies the first/last
// interior cell, reflective mirrors about the face (and the caller flips
// the sign of the normal velocity).  Boundaries are applied one direction
// after another in the reference, so a ghost cell that is out of range in
// several directions takes the composition of the per-direction maps.
__device__ __forceinline__ int map_bc(int g, int n, int bcl, int bcr, bool& flip) {
  if (g < 0) {
    if (bcl == BC_PERIODIC) g += n;
    else if (bcl == BC_OUTFLOW) g = 0;
    else { g = -1 - g; flip = !flip; }
  } else if (g >= n) {
    if (bcr == BC_PERIODIC) g -= n;
    else if (bcr == BC_OUTFLOW) g = n - 1;
    else { g = 2 * n - 1 - g; flip = !flip; }
  }
  return g < 0 ? 0 : (g >= n ? n - 1 : g);
}

// Spread reductions.  Device-scope atomics that every workgroup of a
// launch aims at one word serialise (on MI355X about 10 ns each: some
// 20 us at the end of a 1024-workgroup launch for two such reductions);
// the workgroups of a launch therefore fold their results into RED_SLOTS
// words RED_STRIDE words (128 B, a cache line) apart, chosen by workgroup,
// and the single consumer of the reduction folds the slots.
constexpr int RED_SLOTS = 32, RED_STRIDE = 16;
__host__ __device__ inline int red_index(unsigned wg) { return (int)(wg % RED_SLOTS) * RED_STRIDE; }

// Block id remap so that the tiles one XCD executes are contiguous in the
// logical tile order (block b runs on XCD b % 8 on MI355X): neighbouring
// tiles then share that XCD's L2 for their halo reads.
__device__ __forceinline__ int xcd_remap(int b, int nblocks) {
  const int nxcd = 8;
  int per = nblocks / nxcd, rem = nblocks % nxcd;
  int x = b % nxcd, r = b / nxcd;
  int start = x * per + (x < rem ? x : rem);
  return start + r;
}

}  // namespace pmx
