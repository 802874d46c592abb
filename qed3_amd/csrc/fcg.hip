// Flexible CG on D^dagger D with the learned preconditioner (Notay, SIAM J. Sci. Comput.
// 22 (2000) 1444, FCG(m): each new direction is A-orthogonalised against the last m),
// FP32 vectors with FP64 reductions and FP64 reliable updates.  The net K is trained as
// an approximate inverse of D, and the preconditioner of D^dag D is M = K K^dag, Hermitian
// positive semi-definite by construction: z = K(K^dag r), so <z,r> = |K^dag r|^2.  Rounding
// (BF16 activations) makes M vary slightly from step to step, which FCG tolerates.  z is
// still kept a descent direction: when Re<z,r> < eps |z||r|, z is shifted along r until
// the angle bound holds.  A non-positive
// curvature, a non-finite value or a stalled preconditioned iteration hands the remaining
// residual to the plain mixed CG; convergence is judged on the FP64 true residual only.
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "cg.hpp"
#include "fcg.hpp"
#include "stencil.hpp"

namespace qed3 {
namespace {

using cd = std::complex<double>;

__global__ __launch_bounds__(256) void k_cdot32(const Spin<float>* __restrict__ a,
                                                const Spin<float>* __restrict__ b, size_t n,
                                                GridReduce gr, double2* out) {
  double2 acc = {0.0, 0.0};
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const Spin<float> x = a[i], y = b[i];
    acc.x += double(x.a.x) * y.a.x + double(x.a.y) * y.a.y + double(x.b.x) * y.b.x + double(x.b.y) * y.b.y;
    acc.y += double(x.a.x) * y.a.y - double(x.a.y) * y.a.x + double(x.b.x) * y.b.y - double(x.b.y) * y.b.x;
  }
  grid_reduce(acc, gr, [out](double2 t) { *out = t; });
}

// y = y + a x  (complex a)
__global__ __launch_bounds__(256) void k_caxpy32(Spin<float>* y, const Spin<float>* __restrict__ x,
                                                 float ar, float ai, size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const Spin<float> u = x[i];
    Spin<float> v = y[i];
    v.a.x += ar * u.a.x - ai * u.a.y;
    v.a.y += ar * u.a.y + ai * u.a.x;
    v.b.x += ar * u.b.x - ai * u.b.y;
    v.b.y += ar * u.b.y + ai * u.b.x;
    y[i] = v;
  }
}

// x64 += x32; x32 = 0
__global__ __launch_bounds__(256) void k_acc64(Spin<double>* x64, Spin<float>* x32, size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const Spin<float> a = x32[i];
    Spin<double> v = x64[i];
    v.a.x += a.a.x;
    v.a.y += a.a.y;
    v.b.x += a.b.x;
    v.b.y += a.b.y;
    x64[i] = v;
    Spin<float> z;
    z.a = {0.f, 0.f};
    z.b = {0.f, 0.f};
    x32[i] = z;
  }
}

// x64 += d
__global__ __launch_bounds__(256) void k_add64(Spin<double>* x64, const Spin<double>* __restrict__ d,
                                               size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const Spin<double> a = d[i];
    Spin<double> v = x64[i];
    v.a.x += a.a.x;
    v.a.y += a.a.y;
    v.b.x += a.b.x;
    v.b.y += a.b.y;
    x64[i] = v;
  }
}

// r64 = b - t; r32 = r64; |r64|^2
__global__ __launch_bounds__(256) void k_resid64(const Spin<double>* __restrict__ b,
                                                 const Spin<double>* __restrict__ t, Spin<double>* r64,
                                                 Spin<float>* r32, size_t n, GridReduce gr, double2* out) {
  double2 acc = {0.0, 0.0};
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const Spin<double> bb = b[i], tt = t[i];
    Spin<double> r;
    r.a = {bb.a.x - tt.a.x, bb.a.y - tt.a.y};
    r.b = {bb.b.x - tt.b.x, bb.b.y - tt.b.y};
    r64[i] = r;
    Spin<float> f;
    f.a = {float(r.a.x), float(r.a.y)};
    f.b = {float(r.b.x), float(r.b.y)};
    r32[i] = f;
    acc.x += r.a.x * r.a.x + r.a.y * r.a.y + r.b.x * r.b.x + r.b.y * r.b.y;
  }
  grid_reduce(acc, gr, [out](double2 v) { *out = v; });
}

struct Buf {
  Spin<float>*r, *z, *w, *x, *y;
  std::vector<Spin<float>*> P, Q;  // stored directions p_j and q_j = A p_j
  Spin<double>*r64, *t64, *t2;
  double2* out;
};

}  // namespace

CGResult fcg_solve(Dirac& D, Net& net, Spin<double>* x, const Spin<double>* b, const CGOptions& o,
                   FCGStats* stats) {
  const auto t0 = std::chrono::steady_clock::now();
  const size_t n = D.points();
  hipStream_t st = D.stream();
  const GridReduce gr = D.reducer();
  const unsigned nbk = unsigned(std::min<size_t>((n + 255) / 256, 8192));
  const int m = std::max(1, std::min(o.fcg_m, 16));  // directions kept (FCG(m))
  const bool trace = std::getenv("QED3_FCG_TRACE") != nullptr;
  Buf B{};
  B.P.resize(m + 1);
  B.Q.resize(m + 1);
  for (Spin<float>** p : {&B.r, &B.z, &B.w, &B.x, &B.y})
    QED3_HIP(hipMallocAsync((void**)p, n * sizeof(Spin<float>), st));
  for (int j = 0; j <= m; j++) {
    QED3_HIP(hipMallocAsync((void**)&B.P[j], n * sizeof(Spin<float>), st));
    QED3_HIP(hipMallocAsync((void**)&B.Q[j], n * sizeof(Spin<float>), st));
  }
  for (Spin<double>** p : {&B.r64, &B.t64, &B.t2})
    QED3_HIP(hipMallocAsync((void**)p, n * sizeof(Spin<double>), st));
  QED3_HIP(hipMallocAsync((void**)&B.out, sizeof(double2), st));
  auto reduce_host = [&]() {
    double2 h;
    QED3_HIP(hipMemcpyAsync(&h, B.out, sizeof(h), hipMemcpyDeviceToHost, st));
    QED3_HIP(hipStreamSynchronize(st));
    return cd(h.x, h.y);
  };
  auto dot32 = [&](const Spin<float>* a, const Spin<float>* c) {
    hipLaunchKernelGGL(k_cdot32, dim3(nbk), dim3(256), 0, st, a, c, n, gr, B.out);
    return reduce_host();
  };
  // r64 = b - A x64 ; r32 = r64 ; returns |r64|^2
  auto true_resid = [&]() {
    D.apply<double>(false, B.t2, x, false);
    D.apply<double>(true, B.t64, B.t2, false);
    hipLaunchKernelGGL(k_resid64, dim3(nbk), dim3(256), 0, st, b, B.t64, B.r64, B.r, n, gr, B.out);
    return reduce_host().real();
  };
  QED3_HIP(hipMemsetAsync(x, 0, n * sizeof(Spin<double>), st));
  QED3_HIP(hipMemsetAsync(B.x, 0, n * sizeof(Spin<float>), st));
  const double bn2 = true_resid();  // x = 0: r = b
  const double tol2 = o.tol * o.tol * bn2;
  CGResult res;
  FCGStats s;
  bool done = !(bn2 > 0);
  double maxr2 = bn2, rr = bn2, prev_true = bn2;
  int since_ru = 0, stall = 0;
  std::vector<double> pq_hist(m + 1, 0.0);
  int nstored = 0;  // directions currently kept (ring of m+1 slots)
  for (int it = 0; !done && it < o.maxiter; it++) {
    const int cur = it % (m + 1);
    Spin<float>* p = B.P[cur];
    Spin<float>* qv = B.Q[cur];
    net.apply_adj(B.y, B.r, 1);
    net.apply(B.z, B.y, 1);
    s.precond_applies++;
    {
      constexpr double kEps = 0.05;
      const double zz = dot32(B.z, B.z).real(), zr = dot32(B.z, B.r).real();
      const double need = kEps * std::sqrt(zz * rr);
      if (!(zr >= need) && std::isfinite(zz) && rr > 0) {
        const double mu = (need - zr) / rr;
        hipLaunchKernelGGL(k_caxpy32, dim3(nbk), dim3(256), 0, st, B.z, B.r, float(mu), 0.f, n);
        s.shifted++;
      }
    }
    // p = z - sum_j <q_j, z>/<q_j, p_j> p_j over the kept directions
    QED3_HIP(hipMemcpyAsync(p, B.z, n * sizeof(Spin<float>), hipMemcpyDeviceToDevice, st));
    for (int jj = 1; jj <= nstored; jj++) {
      const int j = (it - jj + (m + 1) * 4) % (m + 1);
      const cd c = dot32(B.Q[j], B.z) / pq_hist[j];
      hipLaunchKernelGGL(k_caxpy32, dim3(nbk), dim3(256), 0, st, p, B.P[j], float(-c.real()), float(-c.imag()), n);
    }
    D.apply<float>(false, B.w, p, false);
    D.apply<float>(true, qv, B.w, false);
    const double pq = dot32(p, qv).real();
    const cd pr = dot32(p, B.r);
    if (trace)
      fprintf(stderr, "fcg it %d pq %.3e pr %.3e%+.3ei |r|^2 %.3e shifted %d\n", it, pq, pr.real(),
              pr.imag(), rr, s.shifted);
    if (!(pq > 0) || !std::isfinite(pq) || !std::isfinite(pr.real())) {
      s.fallback = 1;
      break;
    }
    pq_hist[cur] = pq;
    nstored = std::min(nstored + 1, m);
    const cd al = pr / pq;
    hipLaunchKernelGGL(k_caxpy32, dim3(nbk), dim3(256), 0, st, B.x, p, float(al.real()), float(al.imag()), n);
    hipLaunchKernelGGL(k_caxpy32, dim3(nbk), dim3(256), 0, st, B.r, qv, float(-al.real()), float(-al.imag()), n);
    rr = dot32(B.r, B.r).real();
    res.iterations++;
    since_ru++;
    if (rr < o.delta * o.delta * maxr2 || rr <= tol2 || since_ru >= 100) {
      hipLaunchKernelGGL(k_acc64, dim3(nbk), dim3(256), 0, st, x, B.x, n);
      rr = true_resid();
      res.reliable_updates++;
      since_ru = 0;
      maxr2 = rr;
      if (trace) fprintf(stderr, "fcg reliable update: true |r|^2/|b|^2 = %.3e\n", rr / bn2);
      if (rr <= tol2) {
        done = true;
        break;
      }
      stall = (rr > 0.5 * prev_true) ? stall + 1 : 0;
      prev_true = rr;
      if (stall >= 5) {
        s.fallback = 1;
        break;
      }
    } else {
      maxr2 = std::max(maxr2, rr);
    }
  }
  if (!done) {
    // fold the FP32 correction in and let plain mixed CG finish (or decide) the solve
    hipLaunchKernelGGL(k_acc64, dim3(nbk), dim3(256), 0, st, x, B.x, n);
    rr = true_resid();
    if (rr > tol2 && rr > 0) {
      s.fallback = 1;
      CGOptions o2 = o;
      o2.mixed = true;
      o2.tol = std::min(0.999, o.tol * std::sqrt(bn2 / rr));
      Spin<double>* dx = B.t64;
      const CGResult r2 = cg_solve(D, dx, B.r64, o2);
      s.fallback_iterations = r2.iterations;
      res.iterations += r2.iterations;
      res.reliable_updates += r2.reliable_updates;
      hipLaunchKernelGGL(k_add64, dim3(nbk), dim3(256), 0, st, x, (const Spin<double>*)dx, n);
    }
  }
  for (Spin<float>* p : {B.r, B.z, B.w, B.x, B.y}) (void)hipFreeAsync(p, st);
  for (int j = 0; j <= m; j++) {
    (void)hipFreeAsync(B.P[j], st);
    (void)hipFreeAsync(B.Q[j], st);
  }
  for (Spin<double>* p : {B.r64, B.t64, B.t2}) (void)hipFreeAsync(p, st);
  (void)hipFreeAsync(B.out, st);
  QED3_HIP(hipStreamSynchronize(st));
  res.true_resid = true_residual(D, x, b);
  res.status = res.true_resid <= o.tol * 1.0000001 ? 0 : 3;
  res.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (stats) *stats = s;
  return res;
}

}  // namespace qed3
