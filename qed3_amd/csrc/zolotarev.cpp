#include "zolotarev.hpp"

#include <cmath>
#include <stdexcept>

namespace qed3 {

void jacobi_sncndn(double u, double k, double& sn, double& cn, double& dn, double& K) {
  if (!(k >= 0 && k < 1)) throw std::invalid_argument("elliptic modulus must be in [0, 1)");
  constexpr int kMax = 40;
  double a[kMax + 1], c[kMax + 1];
  a[0] = 1.0;
  double b = std::sqrt(1.0 - k * k);
  c[0] = k;
  int N = 0;
  while (std::abs(c[N]) > 1e-17 && N < kMax) {  // AGM: a_{n+1}, b_{n+1}, c_{n+1}
    const double an = a[N];
    a[N + 1] = 0.5 * (an + b);
    c[N + 1] = 0.5 * (an - b);
    b = std::sqrt(an * b);
    N++;
  }
  K = M_PI / (2.0 * a[N]);
  double phi = std::ldexp(a[N] * u, N);  // 2^N a_N u
  double phi_prev = phi;
  for (int i = N; i > 0; i--) {  // phi_{n-1} = (phi_n + asin(c_n/a_n sin phi_n)) / 2
    phi_prev = phi;
    phi = 0.5 * (phi + std::asin(c[i] / a[i] * std::sin(phi)));
  }
  sn = std::sin(phi);
  cn = std::cos(phi);
  dn = N > 0 ? cn / std::cos(phi_prev - phi) : 1.0;
}

void Zolotarev::set(double k_) {
  if (!(k_ > 0 && k_ < 1) || n < 1) throw std::invalid_argument("Zolotarev needs 0 < k < 1, n >= 1");
  k = k_;
  size = n / 2 + 1;
  c.assign(size, 0.0);
  cp.assign(size, 0.0);
  A.assign(size, 0.0);
  const double kp = std::sqrt(1.0 - k * k);
  double sn, cn, dn, Kp;
  jacobi_sncndn(0.0, kp, sn, cn, dn, Kp);  // K(k')
  double xibar = 1.0;
  for (int m = 1; m < size; m++) {
    jacobi_sncndn(2.0 * Kp * m / n, kp, sn, cn, dn, Kp);
    c[m] = -(cn / sn) * (cn / sn);
    jacobi_sncndn(2.0 * Kp * (m - 0.5) / n, kp, sn, cn, dn, Kp);
    cp[m] = -(cn / sn) * (cn / sn);
    if (m == 1) xibar = 1.0 / dn;
  }
  M = 1.0;
  for (int m = 1; m < size; m++) M *= (1.0 - c[m]) / (1.0 - cp[m]);
  lambda_inv = xibar / M;
  for (int m = 1; m < size; m++)
    lambda_inv *= (1.0 - c[m] * xibar * xibar) / (1.0 - cp[m] * xibar * xibar);
  E = 1.0;
  for (int m = 1; m < size; m++) {
    double num = 1.0, den = 1.0;
    for (int l = 1; l < size; l++) {
      num *= k * k / cp[m] - k * k / c[l];
      if (l != m) den *= k * k / cp[m] - k * k / cp[l];
    }
    A[m] = num / den;
    E *= c[m] / cp[m];
  }
  C = E * 2.0 / (1.0 + lambda_inv) / (k * M);
}

double Zolotarev::operator()(double x) const {
  double s = 1.0;
  for (int m = 1; m < size; m++) s += A[m] / (x * x - k * k / cp[m]);
  return C * x * s;
}

std::vector<double> Zolotarev::shifts() const {
  std::vector<double> s(size - 1);
  for (int m = 1; m < size; m++) s[m - 1] = -k * k / cp[m];
  return s;
}

}  // namespace qed3
