// Flexible CG with the learned preconditioner (see fcg.hip).
#pragma once
#include "cg.hpp"
#include "mpnn.hpp"

namespace qed3 {

struct FCGStats {
  int precond_applies = 0;
  int fallback = 0;             // 1 when plain mixed CG finished the solve
  int fallback_iterations = 0;
  int shifted = 0;              // steps whose z was shifted along r to stay a descent direction
};

CGResult fcg_solve(Dirac& D, Net& net, Spin<double>* x, const Spin<double>* b, const CGOptions& o,
                   FCGStats* stats);

}  // namespace qed3
