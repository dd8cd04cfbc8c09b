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
  int pad;
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
  Scalars* s;
  const double* dtc;     // per-cell dt (independent cells), or null: s->dt
};

// Work buffers of the surrogate-cooled SplitSource (one entry per cell).
struct SurrWork {
  int* list;             // stiff cells that take the network + refinement
  int* fb;               // cells that take the plain integrator
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
int surrogate_cool(pmx_mlp* m, const wind::CoolArgs& a, const wind::SurrWork& w,
                   hipStream_t stream);
}  // namespace mlp
}  // namespace pmx
