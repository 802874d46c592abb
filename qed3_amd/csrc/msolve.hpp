// Multi-shift CG on D^dag D, spectral bounds, and the Zolotarev overlap operator
// (src/both_3d/overlap.h:137-465).
#pragma once
#include <memory>
#include <vector>

#include "cg.hpp"
#include "dirac.hpp"
#include "zolotarev.hpp"

namespace qed3 {

constexpr int kMaxShifts = 32;

// (D^dag D + s_i) x_i = b for all shifts s_i >= 0 at once on one Krylov space (Jegerlehner,
// hep-lat/9612014): CG on the smallest shift, the others carried by the scalar recursions
// zeta_i; shift i stops when its residual zeta_i |r| reaches tol |b|.  FP64.  Reports the
// FP64 true residual |b - (D^dag D + s_i) x_i| / |b| per shift.
std::vector<CGResult> multishift_cg(Dirac& D, const std::vector<double>& shifts,
                                    const std::vector<Spin<double>*>& x, const Spin<double>* b,
                                    const CGOptions& o);

// Extreme eigenvalues of D^dag D: power iteration for mu_max, inverse iteration (FP64 CG
// solves) for mu_min, each until the Rayleigh quotient changes by less than tol.
void spectral_bounds(Dirac& D, double tol, int maxiter, double* mu_min, double* mu_max);

// D_ov = 1 + X eps(X^dag X), X = D / lambda_max, with the Zolotarev partial fractions:
//   D_ov psi     = psi + (C/lambda_max) D [psi + sum_m A_m (D^dag D/lambda_max^2 + sigma_m)^{-1} psi]
//   D_ov^dag psi = psi + C [y + sum_m A_m (D^dag D/lambda_max^2 + sigma_m)^{-1} y],
//                  y = D^dag psi / lambda_max
// (Overlap::mult/adj_deviceAsyncLaunch, overlap.h:309-361), k = lambda_min/lambda_max.
class Overlap {
 public:
  Overlap(Dirac& D, int n, double lambda_min, double lambda_max, double tol);
  ~Overlap();
  Overlap(const Overlap&) = delete;
  Overlap& operator=(const Overlap&) = delete;
  void apply(bool dag, Spin<double>* out, const Spin<double>* in);
  const Zolotarev& zolotarev() const { return z_; }
  long last_iterations() const { return iters_; }

 private:
  Dirac& D_;
  Zolotarev z_;
  double lmax_, tol_;
  std::vector<Spin<double>*> sol_;
  Spin<double>* y_ = nullptr;
  Spin<double>* t_ = nullptr;
  long iters_ = 0;
};

}  // namespace qed3
