// Zolotarev's optimal rational approximation of sign(x) on k <= |x| <= 1 in the
// partial-fraction form used by the overlap operator (src/both_3d/overlap.h:7-135, after
// A.D. Kennedy, hep-lat/0402038):
//   eps(x) = C x (1 + sum_{m=1}^{size-1} A_m / (x^2 - k^2/cp_m)),   size = n/2 + 1,
// with c_m = -(cn/sn)^2 (2 K' m / n), cp_m = -(cn/sn)^2 (2 K' (m - 1/2) / n) of modulus
// k' = sqrt(1 - k^2) and K' = K(k'), and max_{k<=x<=1} |eps(x) - 1| = Delta.
#pragma once
#include <vector>

namespace qed3 {

// Jacobi elliptic functions sn, cn, dn of u with modulus k (0 <= k < 1) and the complete
// integral K(k), by the arithmetic-geometric mean and descending Landen transformation
// (Abramowitz & Stegun 16.4, 17.6).
void jacobi_sncndn(double u, double k, double& sn, double& cn, double& dn, double& K);

struct Zolotarev {
  int n = 21, size = 11;
  double k = 0.01;
  std::vector<double> c, cp, A;  // index 0 unused (as in the reference)
  double M = 1, lambda_inv = 1, E = 1, C = 1;
  Zolotarev(int n_, double k_) : n(n_) { set(k_); }
  void set(double k_);
  double operator()(double x) const;  // partial-fraction form
  double delta() const { return (lambda_inv - 1.0) / (lambda_inv + 1.0); }
  // the shifts sigma_m = -k^2/cp_m > 0 of (X^dag X + sigma_m)^{-1}, m = 1..size-1
  std::vector<double> shifts() const;
};

}  // namespace qed3
