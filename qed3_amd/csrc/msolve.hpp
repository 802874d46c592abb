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
  // Pseudofermion force of the overlap fermion for every link, f_l = -2 Re <eta, dD_ov/dtheta_l eta>
  // with lambda_max and the Zolotarev coefficients held fixed (Overlap::precalc_grad and
  // grad_deviceAsyncLaunch, overlap.h:393-465, for all links at once):
  //   f_l = -(2C/lmax) { Re <eta, dD Z_0> - sum_m A_m [Re <D Z_m / lmax, dD Y_m>
  //                                                  + Re <D Y_m / lmax, dD Z_m>] },
  //   Z_m = R_m eta, Y_m = R_m D^dag eta / lmax, R_m = (D^dag D / lmax^2 + sigma_m)^{-1},
  //   Z_0 = eta + sum_m A_m Z_m.
  // fs [Lt][E], ft [Lt][V] device arrays.  Two multi-shift solves (to the overlap's tol).
  void force(const Spin<double>* eta, double* fs, double* ft);
  // CG on D_ov + D_ov^dag (= D_ov^dag D_ov for an exactly unitary sign function), x from 0,
  // to |b - A x| <= tol |b| on the recursive residual, then the true residual is reported
  // (PseudoFermion<Overlap>::update_eta: MatPoly::solve on DHD_deviceAsyncLaunch,
  // pseudofermion.h:51, overlap.h:364-377, matpoly.h:204-267).  FP64.
  CGResult solve_dhd(Spin<double>* x, const Spin<double>* b, double tol, int maxiter);
  // Spectral bounds of the current D^dag D and the Zolotarev coefficients recomputed
  // (Overlap::update -> compute_lambda_max, Zolotarev::update; overlap.h:200-209, 217-306):
  // lambda_min = sqrt((1 - 100 tol) mu_min), lambda_max = sqrt((1 + 100 tol) mu_max).
  void refresh_bounds(double tol = 1e-8, int maxiter = 2000);
  double lambda_max() const { return lmax_; }
  const Zolotarev& zolotarev() const { return z_; }
  long last_iterations() const { return iters_; }

 private:
  Dirac& D_;
  Zolotarev z_;
  double lmax_, tol_;
  std::vector<Spin<double>*> sol_;
  std::vector<Spin<double>*> ysol_;  // second set of shifted solutions (force only)
  void apply_dhd(Spin<double>* out, const Spin<double>* in, Spin<double>* tmp);
  Spin<double>* y_ = nullptr;
  Spin<double>* t_ = nullptr;
  long iters_ = 0;
};

}  // namespace qed3
