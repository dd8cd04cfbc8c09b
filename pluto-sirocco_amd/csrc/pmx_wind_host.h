// Host-side helpers and structures shared by the wind and surrogate
// translation units.
#pragma once
#include <hip/hip_runtime.h>
#include "../../include/plutomx.h"
#include "pmx_wind_math.h"

namespace pmx {
namespace wind {

// Time state and step reductions of a wind context, in device memory.
struct Scalars {
  double t, dt, inv_dt_hyp, dt_cool;
  unsigned long long cdt_bits;     // max C_dt of the current step's stage 1
  unsigned long long dtcool_bits;  // running min of the cooling dt candidates
  int err;
  int halt;                        // PMX_FLAG_DT_TOO_SMALL once set (sticky)
};

// Arguments of the cooling kernels.
struct CoolArgs {
  double* V;
  int nr, nt;
  long long sv;
  Const k;
  const double* x1;      // ghost-inclusive radial centres
  const double* xi_s;
  const double* tr_s;
  const double* pre;     // [5][sv] cooling prefactors, or null: all 1.0
  Scalars* s;
  const double* dtc;     // per-cell dt (independent cells), or null: s->dt
};

// Several grids' cooling arguments at once: grid g owns the global cell
// indices [off[g], off[g] + args[g].sv); off[g] are multiples of 128 so a
// 128-thread workgroup never straddles two grids.  `total` = off[G].
struct CoolBatch {
  const CoolArgs* args;   // device array, G entries
  const long long* off;   // device array, G + 1 entries
  int G;
  long long total;
};

__device__ __forceinline__ int find_grid(const long long* off, int G, long long q) {
  int lo = 0, hi = G - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= q) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Work buffers of the surrogate-cooled SplitSource (one entry per cell).
struct SurrWork {
  long long* list;       // stiff cells (global index) that take the network + refinement
  long long* fb;         // cells that take the plain integrator
  float* feat;           // [cell-list][4] network inputs
  float* yhat;
  double* f0;            // initial rate of each listed cell
  unsigned long long* cnt;   // [0] listed, [1] fallback, [2] accepted, [3] Radiat calls;
                             // [4..7] running totals of the same
  int guess;                 // 0: the network's guess; 1: no network (ablation)
};

Const make_const(const pmx_wind_desc* d);

}  // namespace wind

namespace mlp {
// Launches the surrogate-cooled SplitSource on `stream` (see pmx_mlp.hip).
int surrogate_cool(pmx_mlp* m, const wind::CoolBatch& b, const wind::SurrWork& w,
                   hipStream_t stream);
// The plain integrator on every cell of the batch.
int plain_cool(const wind::CoolBatch& b, unsigned long long* calls, hipStream_t stream);
}  // namespace mlp
}  // namespace pmx
