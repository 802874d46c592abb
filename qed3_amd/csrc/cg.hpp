// Device-driven CG on A = D^dagger D (the role of MatPoly::solve,
// src/both_3d/matpoly.h:204-267).
//
// Each iteration is five kernels (plus the one-block reduction finalisers) whose control
// flow lives on the device:
//   K1  p = r + beta p_old (formed while gathering), w = D p, gam = |w|^2
//   K2  q = D^dag w,       alpha = mu/gam, x += alpha p, r -= alpha q, |r|^2
//   U1  x64 += x_inner                           (only when a reliable update is due)
//   U2  t = D x64                                (FP64, only when due)
//   U3  r64 = b - D^dag t, r_inner = r64, |r64|^2 (FP64, only when due)
// Kernels read the scalars (alpha, beta, flags) from device memory and exit at once
// when the solve is finished or no update is due, so a chunk of iterations is
// captured once in a hipGraph and replayed with one host poll per chunk.
//
// Inner precision float gives FP32 CG with FP64 reliable updates (Sleijpen & van der
// Vorst 1996; Clark et al., CPC 181 (2010) 1517): the FP64 residual is recomputed
// whenever |r| has dropped by `delta` since the last update, and convergence is only
// declared on the recomputed FP64 residual.  Inner precision double is plain FP64 CG
// whose stop test is likewise confirmed on the true residual.  With `half`, r, p and w
// are stored in IEEE binary16 (FP32 arithmetic, FP32 x_inner), scaled to unit rms per
// component at the start and after every reliable update (one extra pass that only runs
// then): half the field bytes of the FP32 iteration.
#pragma once
#include <string>
#include <vector>
#include "dirac.hpp"

namespace qed3 {

// Per-kernel timing of the two stencil kernels of every iteration (HIP event pairs on the
// solver's stream, direct launches instead of graphs).  After the solve, ms[k] is the mean
// duration of kernel class k over the iterations actually run: 0 = K1 (D with the p
// update), 1 = K2 (D^dag with the r update; for the half-storage solver only the
// iterations that leave their x term pending), 2 = K2 with the x update (half-storage
// solver; = class 1 for the others).
struct KernelTimes {
  static constexpr int kClasses = 5;
  double ms[kClasses] = {};
  int count[kClasses] = {};
};

struct CGOptions {
  double tol = 1e-10;
  int maxiter = 100000;
  bool mixed = false;
  bool half = false;    // binary16-stored inner fields (implies the mixed scheme)
  double delta = 0.01;  // reliable update when |r| falls below delta * max |r| since the last one
  int chunk = 0;  // iterations per graph replay (0 = auto)
  bool use_graph = true;  // replay chunks of iterations as a HIP graph (QED3_NO_GRAPH=1 in the
                          // environment launches the same kernels directly, e.g. for profilers)
  int fcg_m = 4;  // directions kept by flexible CG
  KernelTimes* times = nullptr;  // when set: direct launches, K1/K2 timed with HIP events
};

struct CGResult {
  int status = 0;  // 0 ok, 3 no convergence, 4 breakdown
  int iterations = 0;
  int reliable_updates = 0;
  double true_resid = 0;
  double seconds = 0;
};

// Solve D^dag D x = b for FP64 device fields x, b (x overwritten, starting from 0).
CGResult cg_solve(Dirac& D, Spin<double>* x, const Spin<double>* b, const CGOptions& o);

// ||b - D^dag D x|| / ||b|| in FP64
double true_residual(Dirac& D, const Spin<double>* x, const Spin<double>* b);

// The FP64 CG of cg_solve on work memory the caller provides (ws_bytes at ws, at least
// cg_workspace_bytes(points)), allocating nothing: direct launches instead of a captured
// graph, scalars read back into host stack memory, and the LDS stencil only where the
// context already holds its halo tables (the plain stencil otherwise).
size_t cg_workspace_bytes(size_t points);
// Several systems by the half-storage solver (o.half implied), each on its own context's
// stream: every round queues one chunk of iterations of every unfinished system before the
// host waits for any, so the GPU runs the systems' chunks back to back (and concurrently
// where their streams overlap) without idling on host polls.
std::vector<CGResult> cg_solve_half_many(const std::vector<Dirac*>& Ds, const std::vector<Spin<double>*>& xs,
                                         const std::vector<const Spin<double>*>& bs, const CGOptions& o);
CGResult cg_solve_f64_workspace(Dirac& D, Spin<double>* x, const Spin<double>* b, const CGOptions& o, void* ws,
                                size_t ws_bytes);
// That solve in the library's own per-GPU work array (noalloc.hip); returns a QED3_* status
// (argument or lookup errors, with their text in err), the solve's outcome in res.
int cg_noalloc(Dirac& D, void* d_x, const void* d_b, double tol, int maxiter, CGResult* res, std::string* err);

// Refine the FP32 / binary16 launch shape (D and D^dag together) by timing a fixed number
// of half-storage CG iterations per candidate: the CG kernels carry epilogues the plain
// D apply timed by Dirac::autotune does not, and prefer other shapes.  Needs a gauge
// field; runs once per context.
void tune_cg_tiles(Dirac& D);

}  // namespace qed3
