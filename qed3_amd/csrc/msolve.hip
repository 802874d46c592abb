// Multi-shift CG, spectral bounds and the overlap operator (see msolve.hpp).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <random>
#include <stdexcept>

#include "msolve.hpp"

namespace qed3 {
namespace {

struct ShiftSet {
  Spin<double>* x[kMaxShifts];
  Spin<double>* p[kMaxShifts];
  double a[kMaxShifts];  // alpha_i (kernel A) or zeta_i (kernel B)
  double b[kMaxShifts];  // beta_i (kernel B)
  int n;
};

struct FinStore {
  double2* out;
  __device__ bool skip() const { return false; }
  __device__ void operator()(double2 t) const { *out = t; }
};

__device__ __forceinline__ void axpy_(Spin<double>& y, double a, const Spin<double>& x) {
  y.a.x += a * x.a.x;
  y.a.y += a * x.a.y;
  y.b.x += a * x.b.x;
  y.b.y += a * x.b.y;
}
__device__ __forceinline__ double nrm2(const Spin<double>& v) {
  return v.a.x * v.a.x + v.a.y * v.a.y + v.b.x * v.b.x + v.b.y * v.b.y;
}

// partials of (|w|^2, |p|^2)
__global__ __launch_bounds__(256) void k_pair_norms(const Spin<double>* __restrict__ w,
                                                    const Spin<double>* __restrict__ p, size_t n,
                                                    double2* partials) {
  double2 acc = {0.0, 0.0};
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    acc.x += nrm2(w[i]);
    acc.y += nrm2(p[i]);
  }
  store_partial(acc, partials, blockIdx.x);
}

// r -= alpha (q + s0 p); x_i += alpha_i p_i; partial |r|^2
__global__ __launch_bounds__(256) void k_ms_a(Spin<double>* r, const Spin<double>* __restrict__ q,
                                              const Spin<double>* __restrict__ p, double alpha,
                                              double s0, ShiftSet S, size_t n, double2* partials) {
  double2 acc = {0.0, 0.0};
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const Spin<double> pp = p[i], qq = q[i];
    Spin<double> rr = r[i];
    axpy_(rr, -alpha, qq);
    axpy_(rr, -alpha * s0, pp);
    r[i] = rr;
    acc.x += nrm2(rr);
    for (int k = 0; k < S.n; k++) {
      Spin<double> xv = S.x[k][i];
      axpy_(xv, S.a[k], S.p[k][i]);
      S.x[k][i] = xv;
    }
  }
  store_partial(acc, partials, blockIdx.x);
}

// p_i = zeta_i r + beta_i p_i; p = r + beta p
__global__ __launch_bounds__(256) void k_ms_b(const Spin<double>* __restrict__ r, Spin<double>* p,
                                              double beta, ShiftSet S, size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const Spin<double> rr = r[i];
    Spin<double> pp = p[i];
    Spin<double> v = rr;
    axpy_(v, beta, pp);
    p[i] = v;
    for (int k = 0; k < S.n; k++) {
      const Spin<double> pk = S.p[k][i];
      Spin<double> o;
      o.a = {S.a[k] * rr.a.x + S.b[k] * pk.a.x, S.a[k] * rr.a.y + S.b[k] * pk.a.y};
      o.b = {S.a[k] * rr.b.x + S.b[k] * pk.b.x, S.a[k] * rr.b.y + S.b[k] * pk.b.y};
      S.p[k][i] = o;
    }
  }
}

// out = c0 in + c1 (v + sum_k a_k y_k)   (y_k in S.x, a_k in S.a)
__global__ __launch_bounds__(256) void k_combine(Spin<double>* out, const Spin<double>* __restrict__ in,
                                                 double c0, const Spin<double>* __restrict__ v,
                                                 double c1, ShiftSet S, size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    Spin<double> s = v[i];
    for (int k = 0; k < S.n; k++) axpy_(s, S.a[k], S.x[k][i]);
    Spin<double> o;
    if (in) {
      o = in[i];
      o.a = {c0 * o.a.x, c0 * o.a.y};
      o.b = {c0 * o.b.x, c0 * o.b.y};
    } else {
      o.a = {0.0, 0.0};
      o.b = {0.0, 0.0};
    }
    axpy_(o, c1, s);
    out[i] = o;
  }
}

// res = b - u - s x; partial |res|^2
__global__ __launch_bounds__(256) void k_shift_resid(const Spin<double>* __restrict__ b,
                                                     const Spin<double>* __restrict__ u,
                                                     const Spin<double>* __restrict__ x, double s,
                                                     size_t n, double2* partials) {
  double2 acc = {0.0, 0.0};
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    Spin<double> v = b[i];
    axpy_(v, -1.0, u[i]);
    axpy_(v, -s, x[i]);
    acc.x += nrm2(v);
  }
  store_partial(acc, partials, blockIdx.x);
}

// v = w / |w| with the given scale; partial <v_old, w> is done by Dirac::dot
__global__ __launch_bounds__(256) void k_scale(Spin<double>* v, const Spin<double>* __restrict__ w,
                                               double s, size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    Spin<double> o = w[i];
    o.a = {s * o.a.x, s * o.a.y};
    o.b = {s * o.b.x, s * o.b.y};
    v[i] = o;
  }
}

template <typename T>
T* dnew(size_t n, hipStream_t st) {
  T* p = nullptr;
  QED3_HIP(hipMallocAsync((void**)&p, std::max<size_t>(n, 1) * sizeof(T), st));
  return p;
}

struct Reducer {
  double2* partials;
  double2* out;
  double2* host;
  unsigned nb;
  hipStream_t st;
  double2 get() {
    hipLaunchKernelGGL(k_finalize<FinStore>, dim3(1), dim3(256), 0, st, partials, nb, FinStore{out});
    QED3_HIP(hipMemcpyAsync(host, out, sizeof(double2), hipMemcpyDeviceToHost, st));
    QED3_HIP(hipStreamSynchronize(st));
    return *host;
  }
};

}  // namespace

std::vector<CGResult> multishift_cg(Dirac& D, const std::vector<double>& shifts,
                                    const std::vector<Spin<double>*>& x, const Spin<double>* b,
                                    const CGOptions& o) {
  const auto t0 = std::chrono::steady_clock::now();
  const int ns = int(shifts.size());
  if (ns < 1 || ns > kMaxShifts || int(x.size()) != ns)
    throw std::invalid_argument("multishift_cg: 1..32 shifts, one solution vector each");
  for (double s : shifts)
    if (!(s >= 0) || !std::isfinite(s)) throw std::invalid_argument("multishift_cg: shifts must be >= 0");
  const size_t n = D.points();
  hipStream_t st = D.stream();
  const unsigned nb = unsigned(std::min<size_t>((n + 255) / 256, 4096));
  const double s0 = *std::min_element(shifts.begin(), shifts.end());
  Spin<double>* r = dnew<Spin<double>>(n, st);
  Spin<double>* p = dnew<Spin<double>>(n, st);
  Spin<double>* w = dnew<Spin<double>>(n, st);
  Spin<double>* q = dnew<Spin<double>>(n, st);
  std::vector<Spin<double>*> ps(ns);
  for (int i = 0; i < ns; i++) ps[i] = dnew<Spin<double>>(n, st);
  double2* partials = dnew<double2>(nb, st);
  double2* out = dnew<double2>(1, st);
  double2* host = nullptr;
  QED3_HIP(hipHostMalloc((void**)&host, sizeof(double2), hipHostMallocDefault));
  Reducer red{partials, out, host, nb, st};
  const size_t bytes = n * sizeof(Spin<double>);
  QED3_HIP(hipMemcpyAsync(r, b, bytes, hipMemcpyDeviceToDevice, st));
  QED3_HIP(hipMemcpyAsync(p, b, bytes, hipMemcpyDeviceToDevice, st));
  for (int i = 0; i < ns; i++) {
    QED3_HIP(hipMemsetAsync(x[i], 0, bytes, st));
    QED3_HIP(hipMemcpyAsync(ps[i], b, bytes, hipMemcpyDeviceToDevice, st));
  }
  hipLaunchKernelGGL(k_pair_norms, dim3(nb), dim3(256), 0, st, b, b, n, partials);
  double rho = red.get().x;
  const double bn2 = rho;
  std::vector<double> z(ns, 1.0), zo(ns, 1.0), zn(ns, 1.0);
  std::vector<int> active(ns, bn2 > 0 ? 1 : 0), iters(ns, 0);
  double a_prev = 1.0, b_prev = 0.0;
  int it = 0;
  auto pack = [&](ShiftSet& S, const std::vector<double>& va, const std::vector<double>* vb) {
    S.n = 0;
    for (int i = 0; i < ns; i++) {
      if (!active[i]) continue;
      S.x[S.n] = x[i];
      S.p[S.n] = ps[i];
      S.a[S.n] = va[i];
      S.b[S.n] = vb ? (*vb)[i] : 0.0;
      S.n++;
    }
  };
  std::vector<double> ai(ns), bi(ns);
  while (it < o.maxiter && std::any_of(active.begin(), active.end(), [](int a) { return a != 0; })) {
    D.apply<double>(false, w, p, false);
    D.apply<double>(true, q, w, false);
    hipLaunchKernelGGL(k_pair_norms, dim3(nb), dim3(256), 0, st, w, p, n, partials);
    const double2 wp = red.get();
    const double pq = wp.x + s0 * wp.y;
    if (!(pq > 0) || !std::isfinite(pq)) break;
    const double alpha = rho / pq;
    for (int i = 0; i < ns; i++) {
      if (!active[i]) continue;
      const double ds = shifts[i] - s0;
      zn[i] = z[i] * zo[i] * a_prev / (alpha * b_prev * (zo[i] - z[i]) + zo[i] * a_prev * (1.0 + ds * alpha));
      ai[i] = alpha * zn[i] / z[i];
    }
    ShiftSet S;
    pack(S, ai, nullptr);
    hipLaunchKernelGGL(k_ms_a, dim3(nb), dim3(256), 0, st, r, q, p, alpha, s0, S, n, partials);
    const double rho_new = red.get().x;
    const double beta = rho_new / rho;
    for (int i = 0; i < ns; i++)
      if (active[i]) bi[i] = beta * (zn[i] / z[i]) * (zn[i] / z[i]);
    pack(S, zn, &bi);
    hipLaunchKernelGGL(k_ms_b, dim3(nb), dim3(256), 0, st, r, p, beta, S, n);
    QED3_HIP(hipGetLastError());
    it++;
    for (int i = 0; i < ns; i++) {
      if (!active[i]) continue;
      zo[i] = z[i];
      z[i] = zn[i];
      iters[i] = it;
      if (std::abs(z[i]) * std::sqrt(rho_new) <= o.tol * std::sqrt(bn2)) active[i] = 0;
    }
    a_prev = alpha;
    b_prev = beta;
    rho = rho_new;
  }
  // FP64 true residual per shift
  std::vector<CGResult> res(ns);
  for (int i = 0; i < ns; i++) {
    D.apply<double>(false, w, x[i], false);
    D.apply<double>(true, q, w, false);
    hipLaunchKernelGGL(k_shift_resid, dim3(nb), dim3(256), 0, st, b, q, x[i], shifts[i], n, partials);
    const double rr = red.get().x;
    res[i].iterations = iters[i];
    res[i].true_resid = bn2 > 0 ? std::sqrt(rr / bn2) : 0.0;
    res[i].status = res[i].true_resid <= o.tol * 1.0000001 ? 0 : 3;
  }
  for (void* ptr : {(void*)r, (void*)p, (void*)w, (void*)q, (void*)partials, (void*)out})
    (void)hipFreeAsync(ptr, st);
  for (auto* ptr : ps) (void)hipFreeAsync(ptr, st);
  QED3_HIP(hipStreamSynchronize(st));
  (void)hipHostFree(host);
  const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (auto& r2 : res) r2.seconds = sec;
  return res;
}

void spectral_bounds(Dirac& D, double tol, int maxiter, double* mu_min, double* mu_max) {
  const size_t n = D.points();
  hipStream_t st = D.stream();
  const unsigned nb = unsigned(std::min<size_t>((n + 255) / 256, 4096));
  Spin<double>* v = dnew<Spin<double>>(n, st);
  Spin<double>* w = dnew<Spin<double>>(n, st);
  Spin<double>* t = dnew<Spin<double>>(n, st);
  double2* partials = dnew<double2>(nb, st);
  double2* out = dnew<double2>(1, st);
  double2* host = nullptr;
  QED3_HIP(hipHostMalloc((void**)&host, sizeof(double2), hipHostMallocDefault));
  Reducer red{partials, out, host, nb, st};
  // seeded start vector with every component nonzero
  std::vector<Spin<double>> h(n);
  std::mt19937_64 g(12345);
  std::normal_distribution<double> nd;
  for (auto& s : h) s = Spin<double>{{nd(g), nd(g)}, {nd(g), nd(g)}};
  QED3_HIP(hipMemcpyAsync(v, h.data(), n * sizeof(Spin<double>), hipMemcpyHostToDevice, st));
  auto normalize = [&](Spin<double>* dst, const Spin<double>* src) {
    hipLaunchKernelGGL(k_pair_norms, dim3(nb), dim3(256), 0, st, src, src, n, partials);
    const double nn = red.get().x;
    hipLaunchKernelGGL(k_scale, dim3(nb), dim3(256), 0, st, dst, src, 1.0 / std::sqrt(nn), n);
  };
  auto rayleigh = [&](const Spin<double>* a, const Spin<double>* c) {
    double re, im;
    D.dot(a, c, &re, &im);
    return re;
  };
  normalize(v, v);
  double mu = 0, mu_old = -1;
  for (int it = 0; it < maxiter; it++) {
    D.apply<double>(false, t, v, false);
    D.apply<double>(true, w, t, false);
    mu = rayleigh(v, w);
    normalize(v, w);
    if (std::abs(mu - mu_old) <= tol * std::abs(mu)) break;
    mu_old = mu;
  }
  *mu_max = mu;
  QED3_HIP(hipMemcpyAsync(v, h.data(), n * sizeof(Spin<double>), hipMemcpyHostToDevice, st));
  normalize(v, v);
  CGOptions o;
  o.tol = std::min(1e-12, tol * 1e-2);
  o.mixed = false;
  double nu = 0, nu_old = -1;
  for (int it = 0; it < maxiter; it++) {
    const CGResult r = cg_solve(D, w, v, o);
    if (r.status != 0) throw std::runtime_error("spectral_bounds: inner CG did not converge");
    nu = rayleigh(v, w);  // <v, A^{-1} v> -> 1/mu_min
    normalize(v, w);
    if (std::abs(nu - nu_old) <= tol * std::abs(nu)) break;
    nu_old = nu;
  }
  *mu_min = 1.0 / nu;
  for (void* ptr : {(void*)v, (void*)w, (void*)t, (void*)partials, (void*)out}) (void)hipFreeAsync(ptr, st);
  QED3_HIP(hipStreamSynchronize(st));
  (void)hipHostFree(host);
}

Overlap::Overlap(Dirac& D, int n, double lambda_min, double lambda_max, double tol)
    : D_(D), z_(n, lambda_min / lambda_max), lmax_(lambda_max), tol_(tol) {
  if (!(lambda_min > 0 && lambda_max > lambda_min)) throw std::invalid_argument("need 0 < lambda_min < lambda_max");
  if (z_.size - 1 > kMaxShifts) throw std::invalid_argument("Zolotarev degree too large (at most 65)");
  const size_t bytes = D.points() * sizeof(Spin<double>);
  sol_.resize(z_.size - 1);
  for (auto& p : sol_) QED3_HIP(hipMalloc(&p, bytes));
  QED3_HIP(hipMalloc(&y_, bytes));
  QED3_HIP(hipMalloc(&t_, bytes));
}

Overlap::~Overlap() {
  for (auto* p : sol_) (void)hipFree(p);
  for (auto* p : ysol_) (void)hipFree(p);
  (void)hipFree(y_);
  (void)hipFree(t_);
}

void Overlap::apply(bool dag, Spin<double>* out, const Spin<double>* in) {
  const size_t n = D_.points();
  hipStream_t st = D_.stream();
  const unsigned nb = unsigned(std::min<size_t>((n + 255) / 256, 4096));
  // (D^dag D / lmax^2 + sigma_m)^{-1} v = lmax^2 (D^dag D + lmax^2 sigma_m)^{-1} v
  std::vector<double> sh = z_.shifts();
  for (double& s : sh) s *= lmax_ * lmax_;
  const Spin<double>* rhs = in;
  if (dag) {  // y = D^dag psi / lmax
    D_.apply<double>(true, t_, in, false);
    hipLaunchKernelGGL(k_scale, dim3(nb), dim3(256), 0, st, y_, t_, 1.0 / lmax_, n);
    rhs = y_;
  }
  CGOptions o;
  o.tol = tol_;
  o.mixed = false;
  const std::vector<CGResult> r = multishift_cg(D_, sh, sol_, rhs, o);
  iters_ = r.empty() ? 0 : r[0].iterations;
  for (const auto& ri : r)
    if (ri.status != 0) throw std::runtime_error("overlap: shifted solve did not converge");
  ShiftSet S;
  S.n = int(sol_.size());
  for (int m = 0; m < S.n; m++) {
    S.x[m] = sol_[m];
    S.p[m] = nullptr;
    S.a[m] = z_.A[m + 1] * lmax_ * lmax_;
    S.b[m] = 0;
  }
  if (!dag) {
    // t = psi + sum A_m Z_m ; out = psi + (C/lmax) D t
    hipLaunchKernelGGL(k_combine, dim3(nb), dim3(256), 0, st, t_, static_cast<const Spin<double>*>(nullptr),
                       0.0, in, 1.0, S, n);
    D_.apply<double>(false, y_, t_, false);
    ShiftSet none;
    none.n = 0;
    hipLaunchKernelGGL(k_combine, dim3(nb), dim3(256), 0, st, out, in, 1.0, y_, z_.C / lmax_, none, n);
  } else {
    // out = psi + C (y + sum A_m Y_m)
    hipLaunchKernelGGL(k_combine, dim3(nb), dim3(256), 0, st, out, in, 1.0, y_, z_.C, S, n);
  }
  QED3_HIP(hipGetLastError());
  QED3_HIP(hipStreamSynchronize(st));
}

void Overlap::force(const Spin<double>* eta, double* fs, double* ft) {
  const size_t n = D_.points();
  hipStream_t st = D_.stream();
  const unsigned nb = unsigned(std::min<size_t>((n + 255) / 256, 4096));
  const int ns = int(sol_.size());
  if (ysol_.empty()) {
    ysol_.resize(ns);
    for (auto& p : ysol_) QED3_HIP(hipMalloc(&p, n * sizeof(Spin<double>)));
  }
  // R_m v = lmax^2 (D^dag D + lmax^2 sigma_m)^{-1} v: the solver's x_m are Z_m / lmax^2, Y_m / lmax^2
  std::vector<double> sh = z_.shifts();
  for (double& s : sh) s *= lmax_ * lmax_;
  D_.apply<double>(true, t_, eta, false);
  hipLaunchKernelGGL(k_scale, dim3(nb), dim3(256), 0, st, y_, t_, 1.0 / lmax_, n);  // D^dag eta / lmax
  CGOptions o;
  o.tol = tol_;
  o.mixed = false;
  auto check = [](const std::vector<CGResult>& r) {
    for (const auto& ri : r)
      if (ri.status != 0) throw std::runtime_error("overlap force: shifted solve did not converge");
  };
  check(multishift_cg(D_, sh, sol_, eta, o));
  check(multishift_cg(D_, sh, ysol_, y_, o));
  const double l2 = lmax_ * lmax_, c0 = -2.0 * z_.C / lmax_;
  // Z_0 = eta + sum_m A_m lmax^2 x_m into t_, then f = c0 Re <eta, dD Z_0>
  ShiftSet S;
  S.n = ns;
  for (int m = 0; m < ns; m++) {
    S.x[m] = sol_[m];
    S.p[m] = nullptr;
    S.a[m] = z_.A[m + 1] * l2;
    S.b[m] = 0;
  }
  hipLaunchKernelGGL(k_combine, dim3(nb), dim3(256), 0, st, t_, static_cast<const Spin<double>*>(nullptr), 0.0,
                     eta, 1.0, S, n);
  D_.force_bilinear(eta, t_, nullptr, nullptr, c0, true, fs, ft);
  // pole terms: c0 (-A_m) lmax^3 [Re <D xZ_m, dD xY_m> + Re <D xY_m, dD xZ_m>]
  for (int m = 0; m < ns; m++) {
    D_.apply<double>(false, t_, sol_[m], false);
    D_.apply<double>(false, y_, ysol_[m], false);
    D_.force_bilinear(t_, ysol_[m], y_, sol_[m], -c0 * z_.A[m + 1] * l2 * lmax_, false, fs, ft);
  }
  QED3_HIP(hipGetLastError());
  QED3_HIP(hipStreamSynchronize(st));
}

void Overlap::apply_dhd(Spin<double>* out, const Spin<double>* in, Spin<double>* tmp) {
  const size_t n = D_.points();
  const unsigned nb = unsigned(std::min<size_t>((n + 255) / 256, 4096));
  apply(false, out, in);
  apply(true, tmp, in);
  ShiftSet none;
  none.n = 0;
  hipLaunchKernelGGL(k_combine, dim3(nb), dim3(256), 0, D_.stream(), out, out, 1.0, tmp, 1.0, none, n);
  QED3_HIP(hipGetLastError());
}

CGResult Overlap::solve_dhd(Spin<double>* x, const Spin<double>* b, double tol, int maxiter) {
  if (!(tol > 0) || maxiter < 0) throw std::invalid_argument("solve_dhd: tol > 0, maxiter >= 0");
  const auto t0 = std::chrono::steady_clock::now();
  const size_t n = D_.points();
  hipStream_t st = D_.stream();
  const unsigned nb = unsigned(std::min<size_t>((n + 255) / 256, 4096));
  Spin<double>* r = dnew<Spin<double>>(n, st);
  Spin<double>* p = dnew<Spin<double>>(n, st);
  Spin<double>* q = dnew<Spin<double>>(n, st);
  Spin<double>* tmp = dnew<Spin<double>>(n, st);
  double2* partials = dnew<double2>(nb, st);
  double2* out = dnew<double2>(1, st);
  double2* host = nullptr;
  QED3_HIP(hipHostMalloc((void**)&host, sizeof(double2), hipHostMallocDefault));
  Reducer red{partials, out, host, nb, st};
  const size_t bytes = n * sizeof(Spin<double>);
  ShiftSet none;
  none.n = 0;
  QED3_HIP(hipMemsetAsync(x, 0, bytes, st));
  QED3_HIP(hipMemcpyAsync(r, b, bytes, hipMemcpyDeviceToDevice, st));
  QED3_HIP(hipMemcpyAsync(p, b, bytes, hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_pair_norms, dim3(nb), dim3(256), 0, st, b, b, n, partials);
  const double bn2 = red.get().x;
  double rr = bn2;
  CGResult res;
  int it = 0;
  while (bn2 > 0 && rr > tol * tol * bn2 && it < maxiter) {
    apply_dhd(q, p, tmp);
    double pq, im;
    D_.dot(p, q, &pq, &im);
    if (!(pq > 0) || !std::isfinite(pq)) {
      res.status = 4;
      break;
    }
    const double alpha = rr / pq;
    hipLaunchKernelGGL(k_combine, dim3(nb), dim3(256), 0, st, x, x, 1.0, p, alpha, none, n);
    hipLaunchKernelGGL(k_combine, dim3(nb), dim3(256), 0, st, r, r, 1.0, q, -alpha, none, n);
    hipLaunchKernelGGL(k_pair_norms, dim3(nb), dim3(256), 0, st, r, r, n, partials);
    const double rr_new = red.get().x;
    hipLaunchKernelGGL(k_combine, dim3(nb), dim3(256), 0, st, p, r, 1.0, p, rr_new / rr, none, n);
    rr = rr_new;
    it++;
  }
  // true residual |b - A x| / |b|
  apply_dhd(q, x, tmp);
  hipLaunchKernelGGL(k_shift_resid, dim3(nb), dim3(256), 0, st, b, q, x, 0.0, n, partials);
  const double tr = red.get().x;
  res.iterations = it;
  res.true_resid = bn2 > 0 ? std::sqrt(tr / bn2) : 0.0;
  if (res.status == 0 && !(rr <= tol * tol * bn2)) res.status = 3;
  for (void* ptr : {(void*)r, (void*)p, (void*)q, (void*)tmp, (void*)partials, (void*)out}) (void)hipFreeAsync(ptr, st);
  QED3_HIP(hipStreamSynchronize(st));
  (void)hipHostFree(host);
  res.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return res;
}

void Overlap::refresh_bounds(double tol, int maxiter) {
  double mu_min = 0, mu_max = 0;
  spectral_bounds(D_, tol, maxiter, &mu_min, &mu_max);
  const double lmin = std::sqrt((1.0 - 100.0 * tol) * mu_min), lmax = std::sqrt((1.0 + 100.0 * tol) * mu_max);
  if (!(lmin > 0 && lmax > lmin)) throw std::runtime_error("overlap: bad spectral bounds");
  z_.set(lmin / lmax);
  lmax_ = lmax;
}

}  // namespace qed3
