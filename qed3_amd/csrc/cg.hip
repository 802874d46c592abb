// Device-driven FP64 and mixed FP32/FP64 CG on D^dagger D (see cg.hpp).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "cg.hpp"
#include "cg_epi.hpp"
#include "stencil.hpp"

namespace qed3 {

namespace {

using namespace cgepi;

template <typename T>
__global__ __launch_bounds__(256) void k_init(const Spin<double>* __restrict__ b, Spin<double>* x64,
                                              Spin<T>* xin, Spin<T>* r, Spin<T>* p, size_t n,
                                              CGScalars* sc, GridReduce gr, double tol2,
                                              double delta2, int maxiter) {
  double2 acc = {0.0, 0.0};
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x) {
    const Spin<double> bb = b[i];
    Spin<T> t;
    t.a = {T(bb.a.x), T(bb.a.y)};
    t.b = {T(bb.b.x), T(bb.b.y)};
    r[i] = t;
    p[i] = t;
    Spin<double> z;
    z.a = {0.0, 0.0};
    z.b = {0.0, 0.0};
    x64[i] = z;
    if ((void*)xin != (void*)x64) {
      Spin<T> zt;
      zt.a = {T(0), T(0)};
      zt.b = {T(0), T(0)};
      xin[i] = zt;
    }
    acc.x += norm2(bb);
  }
  grid_reduce(acc, gr, [=](double2 t) {
    sc->mu = t.x;
    sc->bnorm2 = t.x;
    sc->maxr2 = t.x;
    sc->true_r2 = t.x;
    sc->gam = 0;
    sc->alpha = 0;
    sc->beta = 0;
    sc->tol2 = tol2;
    sc->delta2 = delta2;
    sc->iter = 0;
    sc->n_ru = 0;
    sc->maxiter = maxiter;
    sc->need_ru = 0;
    sc->done = (t.x == 0.0) ? 1 : (t.x < 1e300 ? 0 : 2);
  });
}

template <typename T>
__global__ __launch_bounds__(256) void k_flush(Spin<double>* x64, Spin<T>* xin, size_t n,
                                               const CGScalars* sc, int always) {
  if (!always && (sc->done != 0 || sc->need_ru == 0)) return;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x) {
    const Spin<T> a = xin[i];
    Spin<double> xx = x64[i];
    xx.a.x += a.a.x;
    xx.a.y += a.a.y;
    xx.b.x += a.b.x;
    xx.b.y += a.b.y;
    x64[i] = xx;
    Spin<T> z;
    z.a = {T(0), T(0)};
    z.b = {T(0), T(0)};
    xin[i] = z;
  }
}

__global__ __launch_bounds__(256) void k_resid_norms(const Spin<double>* __restrict__ b,
                                                     const Spin<double>* __restrict__ ax, size_t n,
                                                     GridReduce gr, double2* out) {
  double2 acc = {0.0, 0.0};
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x) {
    const Spin<double> bb = b[i], a = ax[i];
    const double d0 = bb.a.x - a.a.x, d1 = bb.a.y - a.a.y, d2 = bb.b.x - a.b.x, d3 = bb.b.y - a.b.y;
    acc.x += d0 * d0 + d1 * d1 + d2 * d2 + d3 * d3;
    acc.y += norm2(bb);
  }
  grid_reduce(acc, gr, [out](double2 t) { *out = t; });
}

template <typename T>
struct Work {
  Spin<T>* r = nullptr;
  Spin<T>* p = nullptr;
  Spin<T>* p2 = nullptr;  // the other buffer of the double-buffered search direction
  Spin<T>* w = nullptr;
  Spin<T>* xin = nullptr;
  Spin<double>* t = nullptr;
};

size_t stream_blocks(size_t n) { return std::min<size_t>((n + 255) / 256, 8192); }

template <typename T>
CGResult run_cg(Dirac& D, Spin<double>* x, const Spin<double>* b, const CGOptions& o) {
  constexpr bool kMixed = std::is_same<T, float>::value;
  const size_t n = D.points();
  hipStream_t st = D.stream();
  auto t0 = std::chrono::steady_clock::now();
  Work<T> W;
  QED3_HIP(hipMallocAsync((void**)&W.r, n * sizeof(Spin<T>), st));
  QED3_HIP(hipMallocAsync((void**)&W.p, n * sizeof(Spin<T>), st));
  QED3_HIP(hipMallocAsync((void**)&W.p2, n * sizeof(Spin<T>), st));
  QED3_HIP(hipMallocAsync((void**)&W.w, n * sizeof(Spin<T>), st));
  if (kMixed) {
    QED3_HIP(hipMallocAsync((void**)&W.xin, n * sizeof(Spin<T>), st));
    QED3_HIP(hipMallocAsync((void**)&W.t, n * sizeof(Spin<double>), st));
  } else {
    W.xin = reinterpret_cast<Spin<T>*>(x);
    W.t = reinterpret_cast<Spin<double>*>(W.w);
  }
  CGScalars* sc = nullptr;
  QED3_HIP(hipMallocAsync((void**)&sc, sizeof(CGScalars), st));
  CGScalars* hsc = nullptr;
  QED3_HIP(hipHostMalloc((void**)&hsc, sizeof(CGScalars), hipHostMallocDefault));

  const GridReduce gr = D.reducer();
  const int nbk = int(stream_blocks(n));
  const double tol2 = o.tol * o.tol;
  const double delta2 = kMixed ? o.delta * o.delta : 0.0;
  hipLaunchKernelGGL(k_init<T>, dim3(nbk), dim3(256), 0, st, b, x, W.xin, W.r, W.p, n, sc, gr,
                     tol2, delta2, o.maxiter);
  QED3_HIP(hipGetLastError());

  const DiracView<T> vin = D.view<T>();
  const DiracView<double> v64 = D.view<double>();
  const int nn = D.tables().max_nn;
  const HaloView hvf = D.halo_for<T>(false), hvt = D.halo_for<T>(true);
  const HaloView hv64f = D.halo_for<double>(false), hv64t = D.halo_for<double>(true);
  // One iteration.  K1 forms p = r + beta p_old on the fly (beta = 0 and p_old = r at the
  // start; after a reliable update r and beta are the true residual's), stores that p in
  // the other buffer and applies D to it; K2 reads the new p.  p_old is read at halo points
  // of other blocks while the new p is written, hence two buffers alternating by parity.
  Spin<T>* P[2] = {W.p, W.p2};
  auto one_iter = [&](int par) {
    Spin<T>* pold = P[par];
    Spin<T>* pnew = P[par ^ 1];
    EpiK1P<T> k1;
    k1.w = W.w;
    k1.sc = sc;
    k1.pnew = pnew;
    launch_stencil_src<T, false>(vin, nn, SrcPUpdate<T>{W.r, pold, sc}, k1, D.tile<T>(false), gr, st, hvf);
    launch_stencil<T, true>(vin, nn, W.w, EpiK2<T>{W.xin, W.r, pnew, sc, T(0)}, D.tile<T>(true), gr, st, hvt);
    if (kMixed)
      hipLaunchKernelGGL(k_flush<T>, dim3(nbk), dim3(256), 0, st, x, W.xin, n, sc, 0);
    launch_stencil<double, false>(v64, nn, x, EpiStoreIfRU{W.t, sc}, D.tile<double>(false), gr, st, hv64f);
    launch_stencil<double, true>(v64, nn, W.t, EpiResid<T>{b, W.r, sc}, D.tile<double>(true), gr, st, hv64t);
  };

  int chunk = o.chunk > 0 ? o.chunk : (n > (1u << 20) ? 8 : 16);
  chunk += chunk & 1;  // even, so every replay of a captured chunk starts at parity 0
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  if (o.use_graph && std::getenv("QED3_NO_GRAPH") == nullptr) {
    QED3_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < chunk; i++) one_iter(i & 1);
    QED3_HIP(hipStreamEndCapture(st, &graph));
    QED3_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  }
  const int max_chunks = o.maxiter / chunk + 4;
  for (int c = 0; c < max_chunks; c++) {
    if (exec)
      QED3_HIP(hipGraphLaunch(exec, st));
    else
      for (int i = 0; i < chunk; i++) one_iter(i & 1);
    QED3_HIP(hipGetLastError());
    QED3_HIP(hipMemcpyAsync(hsc, sc, sizeof(CGScalars), hipMemcpyDeviceToHost, st));
    QED3_HIP(hipStreamSynchronize(st));
    if (hsc->done != 0) break;
  }
  if (kMixed) hipLaunchKernelGGL(k_flush<T>, dim3(nbk), dim3(256), 0, st, x, W.xin, n, sc, 1);
  QED3_HIP(hipMemcpyAsync(hsc, sc, sizeof(CGScalars), hipMemcpyDeviceToHost, st));
  QED3_HIP(hipStreamSynchronize(st));
  CGResult res;
  res.iterations = hsc->iter;
  res.reliable_updates = hsc->n_ru;
  res.status = hsc->done == 1 ? 0 : (hsc->done == 2 ? 4 : 3);
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  (void)hipFreeAsync(W.r, st);
  (void)hipFreeAsync(W.p, st);
  (void)hipFreeAsync(W.p2, st);
  (void)hipFreeAsync(W.w, st);
  if (kMixed) {
    (void)hipFreeAsync(W.xin, st);
    (void)hipFreeAsync(W.t, st);
  }
  (void)hipFreeAsync(sc, st);
  QED3_HIP(hipStreamSynchronize(st));
  (void)hipHostFree(hsc);
  res.true_resid = true_residual(D, x, b);
  res.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (res.status == 0 && !(res.true_resid <= o.tol * 1.0000001)) res.status = 3;
  return res;
}

}  // namespace

double true_residual(Dirac& D, const Spin<double>* x, const Spin<double>* b) {
  const size_t n = D.points();
  hipStream_t st = D.stream();
  Spin<double>*t1 = nullptr, *t2 = nullptr;
  double2* out = nullptr;
  QED3_HIP(hipMallocAsync((void**)&t1, n * sizeof(Spin<double>), st));
  QED3_HIP(hipMallocAsync((void**)&t2, n * sizeof(Spin<double>), st));
  QED3_HIP(hipMallocAsync((void**)&out, sizeof(double2), st));
  D.apply_dagd<double>(t2, x, t1, false);
  hipLaunchKernelGGL(k_resid_norms, dim3(int(stream_blocks(n))), dim3(256), 0, st, b, t2, n,
                     D.reducer(), out);
  QED3_HIP(hipGetLastError());
  double2 h;
  QED3_HIP(hipMemcpyAsync(&h, out, sizeof(h), hipMemcpyDeviceToHost, st));
  QED3_HIP(hipStreamSynchronize(st));
  (void)hipFreeAsync(t1, st);
  (void)hipFreeAsync(t2, st);
  (void)hipFreeAsync(out, st);
  QED3_HIP(hipStreamSynchronize(st));
  return h.y > 0 ? std::sqrt(h.x / h.y) : 0.0;
}

CGResult cg_solve(Dirac& D, Spin<double>* x, const Spin<double>* b, const CGOptions& o) {
  if (o.mixed) return run_cg<float>(D, x, b, o);
  return run_cg<double>(D, x, b, o);
}

}  // namespace qed3
