// The Cartesian stage kernel for domains whose sides do not all share one
// boundary rule (BCK_ANY: the run-time map_bc per side).  These
// instantiations are compiled here, apart from the single-rule ones in
// pmx_hydro.hip: that translation unit is built with LLVM's iterative-ILP
// scheduler, under which the register allocator of this toolchain
// (ROCm 7.2) crashes on the PPM / HLLC / weighted-stage BCK_ANY kernel.
#include <hip/hip_runtime.h>

#include "pmx_common.h"
#include "pmx_hydro_kernels.h"

namespace pmx {
namespace {

template <int DIMS, int REC, int SOLVER, int MODE>
void launch(const StageArgs& a, int nblocks, hipStream_t st) {
  hipLaunchKernelGGL((stage_kernel<DIMS, REC, SOLVER, MODE, BCK_ANY, TILE_X, tile_y<DIMS>(), NTHREADS>),
                     dim3(nblocks), dim3(NTHREADS), 0, st, a);
}

template <int DIMS, int REC, int SOLVER>
void by_mode(int mode, const StageArgs& a, int nblocks, hipStream_t st) {
  if (mode == MODE_FIRST) launch<DIMS, REC, SOLVER, MODE_FIRST>(a, nblocks, st);
  else if (mode == MODE_WEIGHTED) launch<DIMS, REC, SOLVER, MODE_WEIGHTED>(a, nblocks, st);
  else launch<DIMS, REC, SOLVER, MODE_THIRD>(a, nblocks, st);
}

template <int DIMS, int REC>
void by_solver(int solver, int mode, const StageArgs& a, int nblocks, hipStream_t st) {
  if (solver == SOL_ROE) by_mode<DIMS, REC, SOL_ROE>(mode, a, nblocks, st);
  else by_mode<DIMS, REC, SOL_HLLC>(mode, a, nblocks, st);
}

template <int DIMS>
void by_rec(int rec, int solver, int mode, const StageArgs& a, int nblocks, hipStream_t st) {
  switch (rec) {
    case REC_PLM_VL: by_solver<DIMS, REC_PLM_VL>(solver, mode, a, nblocks, st); break;
    case REC_PLM_MC: by_solver<DIMS, REC_PLM_MC>(solver, mode, a, nblocks, st); break;
    case REC_PLM_MM: by_solver<DIMS, REC_PLM_MM>(solver, mode, a, nblocks, st); break;
    case REC_PLM_DEFAULT: by_solver<DIMS, REC_PLM_DEFAULT>(solver, mode, a, nblocks, st); break;
    default: by_solver<DIMS, REC_PPM>(solver, mode, a, nblocks, st); break;
  }
}

}  // namespace

void launch_stage_any(int dims, int rec, int solver, int mode, const StageArgs& a, int nblocks,
                      hipStream_t stream) {
  if (dims == 3) by_rec<3>(rec, solver, mode, a, nblocks, stream);
  else by_rec<2>(rec, solver, mode, a, nblocks, stream);
}

}  // namespace pmx
