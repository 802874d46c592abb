// Batched CG: many independent systems D_k^dag D_k x_k = b_k of different lattices
// (simplicial/dual, refinement n, Lt) and masses advanced together, one launch per CG
// phase for the whole batch.  The blocks of a launch are laid out system after system
// (each system's count a multiple of 8, so the XCD-aware tile order holds per system);
// each block finds its system, runs that system's LDS-staged stencil tile (time tile
// from the system's autotuned shape) and reduces into that system's scalars, so systems
// converge independently (a finished system's blocks exit at once).  The iteration is the
// half-storage one of cg_solve (binary16 r, p, w scaled per system, FP32 x updated once
// per group of four iterations, FP64 reliable updates, true-residual convergence).
#pragma once
#include <vector>

#include "cg.hpp"

namespace qed3 {

std::vector<CGResult> batch_cg(const std::vector<Dirac*>& Ds, const std::vector<Spin<double>*>& xs,
                               const std::vector<const Spin<double>*>& bs, const CGOptions& o);

}  // namespace qed3
