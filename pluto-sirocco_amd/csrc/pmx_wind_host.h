// Host-side helpers and structures shared by the wind and surrogate
// translation units.
#pragma once
#include <hip/hip_runtime.h>
#include "../../include/plutomx.h"
#include "pmx_common.h"
#include "pmx_wind_math.h"

namespace pmx {
namespace wind {

// Time state and step reductions of a wind context, in device memory.
struct Scalars {
  double t, dt, inv_dt_hyp, dt_cool;
  unsigned long long cdt_bits;     // max C_dt of the current step's stage 1
  unsigned long long dtcool_bits;  // running min of the cooling dt candidates
  double tstop;                    // runtime.tstop
  double dt_uncut;                 // the dt a cut to tstop replaced
  long long step;                  // g_stepNumber of the next step
  int err;
  int halt;                        // PMX_FLAG_DT_TOO_SMALL / _COOL_SUBSTEPS / _TSTOP once set
                                   // (sticky): later main-loop steps are no-ops
  int last;                        // the next step is the last one (dt cut to tstop - t)
  // spread slots of the two reductions (words red_index(k)): the stage-1
  // C_dt maxima (0 = none) and the cooling dt minima (~0 = none), folded
  // into cdt_bits / dtcool_bits by take_cdt / fold_dtcool
  unsigned long long cdt_red[RED_SLOTS * RED_STRIDE];
  unsigned long long dtcool_red[RED_SLOTS * RED_STRIDE];
};

// Empty spread slots.
__host__ __device__ inline void red_clear(Scalars& s) {
  for (int k = 0; k < RED_SLOTS; k++) {
    s.cdt_red[k * RED_STRIDE] = 0ull;
    s.dtcool_red[k * RED_STRIDE] = ~0ull;
  }
}

// The stage-1 maximum of C_dt (non-negative doubles order like their bit
// patterns); clears the words for the next step.
__device__ inline double take_cdt(Scalars* s) {
  unsigned long long m = s->cdt_bits;
#pragma unroll
  for (int k = 0; k < RED_SLOTS; k++) {
    const unsigned long long v = s->cdt_red[k * RED_STRIDE];
    m = v > m ? v : m;
  }
  s->cdt_bits = 0ull;
#pragma unroll
  for (int k = 0; k < RED_SLOTS; k++) s->cdt_red[k * RED_STRIDE] = 0ull;
  return __longlong_as_double((long long)m);
}

// Folds the cooling dt candidates of the spread slots into dtcool_bits
// (all slots are read before any is cleared: one memory latency).
__host__ __device__ inline void fold_dtcool(Scalars* s) {
  unsigned long long m = s->dtcool_bits;
#pragma unroll
  for (int k = 0; k < RED_SLOTS; k++) {
    const unsigned long long v = s->dtcool_red[k * RED_STRIDE];
    m = v < m ? v : m;
  }
#pragma unroll
  for (int k = 0; k < RED_SLOTS; k++) s->dtcool_red[k * RED_STRIDE] = ~0ull;
  s->dtcool_bits = m;
}

// Src/main.c:226-229: the step that would reach tstop (to 1e-8 relative)
// is the last one, and its dt is cut to land on tstop.
__host__ __device__ inline void clamp_last_step(Scalars& s) {
  if ((s.t + s.dt) >= s.tstop * (1.0 - 1.e-8)) {
    s.dt_uncut = s.dt;
    s.dt = (s.tstop - s.t);
    s.last = 1;
  }
}

// End of a main-loop step, after NextTimeStep: the run-stopping conditions
// (a cooling integration over its sub-step limit, the step that reached
// tstop) halt the run; otherwise the next step's dt is cut to tstop.
__device__ inline void end_of_step(Scalars* s) {
  s->step++;
  if (s->err & PMX_FLAG_COOL_SUBSTEPS) {
    s->halt |= PMX_FLAG_COOL_SUBSTEPS;
    s->err &= ~PMX_FLAG_COOL_SUBSTEPS;
  }
  if (s->last) {
    s->halt |= PMX_FLAG_TSTOP;
    s->last = 0;
    return;
  }
  Scalars q = *s;
  clamp_last_step(q);
  s->dt = q.dt;
  s->dt_uncut = q.dt_uncut;
  s->last = q.last;
}

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
  int loop;              // a main-loop step: nothing to do once s->halt is set
};

// Several grids' cooling arguments at once: grid g owns the global cell
// indices [off[g], off[g] + args[g].sv); off[g] are multiples of 128 so a
// 128-thread workgroup never straddles two grids.  `total` = off[G].
struct CoolBatch {
  const CoolArgs* args;   // device array, G entries
  const long long* off;   // device array, G + 1 entries
  int G;
  long long total;
  int loop;               // a main-loop step: cells of halted grids are left alone
};

__device__ __forceinline__ int find_grid(const long long* off, int G, long long q) {
  int lo = 0, hi = G - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= q) lo = mid; else hi = mid - 1;
  }
  return lo;
}

constexpr int SURR_CALLS = 8;
#define PMX_SURR_COUNTERS (::pmx::wind::SURR_CALLS + ::pmx::RED_SLOTS * ::pmx::RED_STRIDE)

// Sum of the Radiat-call slots of a host copy of the counters.
inline unsigned long long surr_calls(const unsigned long long* c) {
  unsigned long long n = 0;
  for (int k = 0; k < RED_SLOTS; k++) n += c[SURR_CALLS + red_index(k)];
  return n;
}

// Work buffers of the surrogate-cooled SplitSource (one entry per cell).
struct SurrWork {
  long long* list;       // stiff cells (global index) that take the network + refinement
  long long* fb;         // cells that take the plain integrator
  float* feat;           // [cell-list][4] network inputs
  double* f0;            // initial rate of each listed cell
  unsigned long long* cnt;   // PMX_SURR_COUNTERS entries, zeroed at creation: [0], [1]
                             // the tails of list and fb during a call (zero between
                             // calls); running totals [4] listed, [5] fallback,
                             // [6] accepted; Radiat calls in the spread slots
                             // cnt[SURR_CALLS + red_index(k)] (summed by the reader)
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
