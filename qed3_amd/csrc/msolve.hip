// Multi-shift CG, spectral bounds and the overlap operator (see msolve.hpp).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <random>
#include <stdexcept>

#include "msolve.hpp"
#include "stencil.hpp"

namespace qed3 {
namespace {

// fields y_k with coefficients a_k for k_combine
struct ShiftSet {
  Spin<double>* x[kMaxShifts];
  double a[kMaxShifts];
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

// ---- device-driven multi-shift CG: every scalar recursion runs in one-block finalisers, so a
// chunk of iterations is one graph replay and the host polls once per chunk ----------------
struct MSScalars {
  double rho, bn2, alpha, beta, a_prev, b_prev, s0, tol2;
  double shift[kMaxShifts], z[kMaxShifts], zo[kMaxShifts], zn[kMaxShifts], ai[kMaxShifts], bi[kMaxShifts];
  int active[kMaxShifts];  // still iterating
  int now[kMaxShifts];     // active in the current iteration (what K_a and K_b update)
  int iters[kMaxShifts];
  int ns, it, maxiter;
  int done;  // 0 running, 1 every shift converged, 3 iteration cap, 4 breakdown
};

struct ShiftPtrs {
  Spin<double>* x[kMaxShifts];
  Spin<double>* p[kMaxShifts];
  int n;
};

// D or D^dag into out, skipped once the solve is done
struct EpiStoreMS {
  static constexpr bool kReduce = false;
  Spin<double>* out;
  const MSScalars* sc;
  __device__ bool skip() const { return sc->done != 0; }
  __device__ void begin() {}
  __device__ void operator()(size_t i, const Spin<double>& v, double2&) const { out[i] = v; }
  __device__ void operator()(double2) const {}
};

// partials of (|w|^2, |p|^2)
__global__ __launch_bounds__(256) void k_msd_norms(const Spin<double>* __restrict__ w,
                                                   const Spin<double>* __restrict__ p, size_t n,
                                                   double2* partials, const MSScalars* sc) {
  if (sc->done != 0) return;
  double2 acc = {0.0, 0.0};
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    acc.x += nrm2(w[i]);
    acc.y += nrm2(p[i]);
  }
  store_partial(acc, partials, blockIdx.x);
}

// alpha = rho / <p, (A + s0) p>; zeta_i and alpha_i of every active shift (Jegerlehner)
struct FinMSAlpha {
  MSScalars* sc;
  __device__ bool skip() const { return sc->done != 0; }
  __device__ void operator()(double2 t) const {
    const double pq = t.x + sc->s0 * t.y;
    if (!(pq > 0.0) || !(pq < 1e300)) {
      sc->done = 4;
      return;
    }
    const double alpha = sc->rho / pq, ap = sc->a_prev, bp = sc->b_prev;
    sc->alpha = alpha;
    for (int i = 0; i < sc->ns; i++) {
      sc->now[i] = sc->active[i];
      if (!sc->active[i]) continue;
      const double ds = sc->shift[i] - sc->s0, z = sc->z[i], zo = sc->zo[i];
      const double zn = z * zo * ap / (alpha * bp * (zo - z) + zo * ap * (1.0 + ds * alpha));
      sc->zn[i] = zn;
      sc->ai[i] = alpha * zn / z;
    }
  }
};

// r -= alpha (q + s0 p); x_i += alpha_i p_i for the shifts active now; partial |r|^2
__global__ __launch_bounds__(256) void k_msd_a(Spin<double>* r, const Spin<double>* __restrict__ q,
                                               const Spin<double>* __restrict__ p, ShiftPtrs S, size_t n,
                                               double2* partials, const MSScalars* sc) {
  if (sc->done != 0) return;
  const double alpha = sc->alpha, s0 = sc->s0;
  double2 acc = {0.0, 0.0};
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const Spin<double> pp = p[i], qq = q[i];
    Spin<double> rr = r[i];
    axpy_(rr, -alpha, qq);
    axpy_(rr, -alpha * s0, pp);
    r[i] = rr;
    acc.x += nrm2(rr);
    for (int k = 0; k < S.n; k++) {
      if (!sc->now[k]) continue;  // uniform
      Spin<double> xv = S.x[k][i];
      axpy_(xv, sc->ai[k], S.p[k][i]);
      S.x[k][i] = xv;
    }
  }
  store_partial(acc, partials, blockIdx.x);
}

// beta, beta_i; then the shifts' zeta history, iteration counts and convergence
// (|zeta_i| |r| <= tol |b|), the iteration cap
struct FinMSBeta {
  MSScalars* sc;
  __device__ bool skip() const { return sc->done != 0; }
  __device__ void operator()(double2 t) const {
    const double rho_new = t.x, beta = rho_new / sc->rho;
    sc->beta = beta;
    sc->it += 1;
    int any = 0;
    for (int i = 0; i < sc->ns; i++) {
      if (!sc->active[i]) continue;
      const double z = sc->z[i], zn = sc->zn[i];
      sc->bi[i] = beta * (zn / z) * (zn / z);
      sc->zo[i] = z;
      sc->z[i] = zn;
      sc->iters[i] = sc->it;
      if (fabs(zn) * sqrt(rho_new) <= sqrt(sc->tol2 * sc->bn2))
        sc->active[i] = 0;
      else
        any = 1;
    }
    sc->a_prev = sc->alpha;
    sc->b_prev = beta;
    sc->rho = rho_new;
    if (!any)
      sc->done = 1;
    else if (sc->it >= sc->maxiter)
      sc->done = 3;
  }
};

// p = r + beta p; p_i = zeta_i r + beta_i p_i for the shifts active now
__global__ __launch_bounds__(256) void k_msd_b(const Spin<double>* __restrict__ r, Spin<double>* p, ShiftPtrs S,
                                               size_t n, const MSScalars* sc) {
  if (sc->done != 0) return;
  const double beta = sc->beta;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const Spin<double> rr = r[i];
    Spin<double> v = rr;
    axpy_(v, beta, p[i]);
    p[i] = v;
    for (int k = 0; k < S.n; k++) {
      if (!sc->now[k]) continue;  // uniform
      const double a = sc->zn[k], bb = sc->bi[k];
      const Spin<double> pk = S.p[k][i];
      Spin<double> o;
      o.a = {a * rr.a.x + bb * pk.a.x, a * rr.a.y + bb * pk.a.y};
      o.b = {a * rr.b.x + bb * pk.b.x, a * rr.b.y + bb * pk.b.y};
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
  const double bn2 = red.get().x;
  MSScalars h{};
  h.rho = bn2;
  h.bn2 = bn2;
  h.a_prev = 1.0;
  h.b_prev = 0.0;
  h.s0 = s0;
  h.tol2 = o.tol * o.tol;
  h.ns = ns;
  h.maxiter = o.maxiter;
  for (int i = 0; i < ns; i++) {
    h.shift[i] = shifts[i];
    h.z[i] = h.zo[i] = h.zn[i] = 1.0;
    h.active[i] = bn2 > 0 ? 1 : 0;
  }
  h.done = (bn2 > 0 && o.maxiter > 0) ? 0 : (bn2 > 0 ? 3 : 1);
  MSScalars* sc = dnew<MSScalars>(1, st);
  MSScalars* hs = nullptr;
  QED3_HIP(hipHostMalloc((void**)&hs, sizeof(MSScalars), hipHostMallocDefault));
  *hs = h;
  QED3_HIP(hipMemcpyAsync(sc, hs, sizeof(MSScalars), hipMemcpyHostToDevice, st));
  ShiftPtrs S;
  S.n = ns;
  for (int i = 0; i < ns; i++) {
    S.x[i] = x[i];
    S.p[i] = ps[i];
  }
  const DiracView<double> v64 = D.view<double>();
  const int nn = D.tables().max_nn;
  const HaloView hvf = D.halo_for<double>(false), hvt = D.halo_for<double>(true);
  const GridReduce gr = D.reducer();
  auto one_iter = [&] {
    launch_stencil<double, false>(v64, nn, p, EpiStoreMS{w, sc}, D.tile<double>(false), gr, st, hvf);
    launch_stencil<double, true>(v64, nn, w, EpiStoreMS{q, sc}, D.tile<double>(true), gr, st, hvt);
    hipLaunchKernelGGL(k_msd_norms, dim3(nb), dim3(256), 0, st, w, p, n, partials, sc);
    hipLaunchKernelGGL(k_finalize<FinMSAlpha>, dim3(1), dim3(256), 0, st, partials, nb, FinMSAlpha{sc});
    hipLaunchKernelGGL(k_msd_a, dim3(nb), dim3(256), 0, st, r, q, p, S, n, partials, sc);
    hipLaunchKernelGGL(k_finalize<FinMSBeta>, dim3(1), dim3(256), 0, st, partials, nb, FinMSBeta{sc});
    hipLaunchKernelGGL(k_msd_b, dim3(nb), dim3(256), 0, st, r, p, S, n, sc);
  };
  // chunks of iterations replayed as one graph; the kernels of iterations after the end
  // exit at once
  const int chunk = 8;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (std::getenv("QED3_NO_GRAPH") == nullptr) {
    QED3_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    for (int c = 0; c < chunk; c++) one_iter();
    QED3_HIP(hipStreamEndCapture(st, &graph));
    QED3_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  }
  for (int c = 0; c <= o.maxiter / chunk + 1; c++) {
    QED3_HIP(hipMemcpyAsync(hs, sc, sizeof(MSScalars), hipMemcpyDeviceToHost, st));
    QED3_HIP(hipStreamSynchronize(st));
    if (hs->done != 0) break;
    if (exec)
      QED3_HIP(hipGraphLaunch(exec, st));
    else
      for (int k = 0; k < chunk; k++) one_iter();
    QED3_HIP(hipGetLastError());
  }
  QED3_HIP(hipMemcpyAsync(hs, sc, sizeof(MSScalars), hipMemcpyDeviceToHost, st));
  QED3_HIP(hipStreamSynchronize(st));
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  std::vector<int> iters(hs->iters, hs->iters + ns);
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
  for (void* ptr : {(void*)r, (void*)p, (void*)w, (void*)q, (void*)partials, (void*)out, (void*)sc})
    (void)hipFreeAsync(ptr, st);
  for (auto* ptr : ps) (void)hipFreeAsync(ptr, st);
  QED3_HIP(hipStreamSynchronize(st));
  (void)hipHostFree(host);
  (void)hipHostFree(hs);
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
    S.a[m] = z_.A[m + 1] * lmax_ * lmax_;
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
    S.a[m] = z_.A[m + 1] * l2;
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
