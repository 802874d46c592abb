// Batched mixed-precision CG over systems of different shapes (see batch.hpp).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>

#include "batch.hpp"
#include "cg_epi.hpp"
#include "stencil.hpp"

namespace qed3 {
namespace {

using namespace cgepi;

struct Sys {
  DiracView<float> v32;
  DiracView<double> v64;
  int ns;
  int nxr;                 // site ranges of 256
  int nst32, ts32;         // FP32 time tiling
  int nst64, ts64;         // FP64 time tiling
  int blk32, nblk32;       // first block / blocks in the FP32 stencil grids
  int blk64, nblk64;       // ... in the FP64 stencil grids
  int sblk, nsblk;         // ... in the streaming grids
  size_t n;                // points
  Spin<float>*r, *p, *w, *xin;
  Spin<double>*x, *t;
  const Spin<double>* b;
  CGScalars* sc;
  unsigned int* counter;
};

constexpr int kBlock = 256;

// index of the system owning global block b (first[k] ascending, first[0] = 0)
__device__ __forceinline__ int sys_of(const int* first, int nsys, int b) {
  int lo = 0, hi = nsys - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first[mid] <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

enum Phase { kK1 = 0, kK2 = 1, kU2 = 2, kU3 = 3 };

template <int PH>
__global__ __launch_bounds__(256) void k_batch_stencil(const Sys* __restrict__ sys, const int* first,
                                                       int nsys, double2* partials) {
  const int k = sys_of(first, nsys, blockIdx.x);
  const Sys& S = sys[k];
  constexpr bool F64 = (PH == kU2 || PH == kU3);
  const int blk0 = F64 ? S.blk64 : S.blk32;
  const int nb = F64 ? S.nblk64 : S.nblk32;
  const int ts = F64 ? S.ts64 : S.ts32;
  const int lb = blockIdx.x - blk0;
  const int xr = lb % S.nxr, st = lb / S.nxr;
  const int x = xr * kBlock + threadIdx.x;
  double2 acc = {0.0, 0.0};
  auto run = [&](auto epi, const auto& v, const auto* in) {
    if (epi.skip()) return;
    epi.begin();
    using T = typename std::remove_const<typename std::remove_pointer<decltype(in)>::type>::type;
    if (S.ns == 6)
      stencil_site<decltype(T{}.a.x), 6, PH == kK2 || PH == kU3>(v, in, epi, x, st * ts, ts, acc);
    else
      stencil_site<decltype(T{}.a.x), 3, PH == kK2 || PH == kU3>(v, in, epi, x, st * ts, ts, acc);
    if constexpr (decltype(epi)::kReduce) store_partial(acc, partials + blk0, unsigned(lb));
    (void)nb;
  };
  if constexpr (PH == kK1) run(EpiK1<float>{S.w, S.sc}, S.v32, (const Spin<float>*)S.p);
  if constexpr (PH == kK2) run(EpiK2<float>{S.xin, S.r, S.p, S.sc, 0.f}, S.v32, (const Spin<float>*)S.w);
  if constexpr (PH == kU2) run(EpiStoreIfRU{S.t, S.sc}, S.v64, (const Spin<double>*)S.x);
  if constexpr (PH == kU3) run(EpiResid<float>{S.b, S.r, S.sc}, S.v64, (const Spin<double>*)S.t);
}

// one block per system: sum that system's partials in block order and run its finaliser
template <int PH>
__global__ __launch_bounds__(256) void k_batch_finalize(const Sys* __restrict__ sys, const double2* partials) {
  const Sys& S = sys[blockIdx.x];
  constexpr bool F64 = (PH == kU2 || PH == kU3);
  const int blk0 = F64 ? S.blk64 : S.blk32;
  const int nb = F64 ? S.nblk64 : S.nblk32;
  auto fin = [&](auto epi) {
    if (epi.skip()) return;
    double2 s = {0.0, 0.0};
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
      const double2 p = partials[blk0 + i];
      s.x += p.x;
      s.y += p.y;
    }
    const double2 tot = block_sum2(s);
    if (threadIdx.x == 0) epi(tot);
  };
  if constexpr (PH == kK1) fin(EpiK1<float>{S.w, S.sc});
  if constexpr (PH == kK2) fin(EpiK2<float>{S.xin, S.r, S.p, S.sc, 0.f});
  if constexpr (PH == kU3) fin(EpiResid<float>{S.b, S.r, S.sc});
}

// streaming phases: 0 init, 1 flush (x64 += xin when an update is due), 2 p update,
// 3 final flush
template <int PH>
__global__ __launch_bounds__(256) void k_batch_stream(const Sys* __restrict__ sys, const int* first,
                                                      int nsys, double2* partials, double tol2,
                                                      double delta2, int maxiter) {
  const int k = sys_of(first, nsys, blockIdx.x);
  const Sys& S = sys[k];
  const int lb = blockIdx.x - S.sblk;
  const size_t stride = size_t(S.nsblk) * blockDim.x;
  if (PH == 1 && (S.sc->done != 0 || S.sc->need_ru == 0)) return;
  if (PH == 2 && S.sc->done != 0) return;
  const float beta = PH == 2 ? float(S.sc->beta) : 0.f;
  double2 acc = {0.0, 0.0};
  for (size_t i = size_t(lb) * blockDim.x + threadIdx.x; i < S.n; i += stride) {
    if (PH == 0) {
      const Spin<double> bb = S.b[i];
      Spin<float> t;
      t.a = {float(bb.a.x), float(bb.a.y)};
      t.b = {float(bb.b.x), float(bb.b.y)};
      S.r[i] = t;
      S.p[i] = t;
      Spin<double> z;
      z.a = {0.0, 0.0};
      z.b = {0.0, 0.0};
      S.x[i] = z;
      Spin<float> zf;
      zf.a = {0.f, 0.f};
      zf.b = {0.f, 0.f};
      S.xin[i] = zf;
      acc.x += norm2(bb);
    } else if (PH == 1 || PH == 3) {
      const Spin<float> a = S.xin[i];
      Spin<double> xx = S.x[i];
      xx.a.x += a.a.x;
      xx.a.y += a.a.y;
      xx.b.x += a.b.x;
      xx.b.y += a.b.y;
      S.x[i] = xx;
      Spin<float> z;
      z.a = {0.f, 0.f};
      z.b = {0.f, 0.f};
      S.xin[i] = z;
    } else {
      const Spin<float> rr = S.r[i];
      Spin<float> pp = S.p[i];
      pp.a.x = rr.a.x + beta * pp.a.x;
      pp.a.y = rr.a.y + beta * pp.a.y;
      pp.b.x = rr.b.x + beta * pp.b.x;
      pp.b.y = rr.b.y + beta * pp.b.y;
      S.p[i] = pp;
    }
  }
  if (PH == 0) {
    CGScalars* sc = S.sc;
    grid_reduce_part(acc, partials + S.sblk, S.counter, unsigned(lb), unsigned(S.nsblk),
                     [=](double2 t) {
                       sc->mu = t.x;
                       sc->bnorm2 = t.x;
                       sc->maxr2 = t.x;
                       sc->true_r2 = t.x;
                       sc->gam = sc->alpha = sc->beta = 0;
                       sc->tol2 = tol2;
                       sc->delta2 = delta2;
                       sc->iter = sc->n_ru = sc->need_ru = 0;
                       sc->maxiter = maxiter;
                       sc->done = (t.x == 0.0) ? 1 : (t.x < 1e300 ? 0 : 2);
                     });
  }
}

}  // namespace

std::vector<CGResult> batch_cg(const std::vector<Dirac*>& Ds, const std::vector<Spin<double>*>& xs,
                               const std::vector<const Spin<double>*>& bs, const CGOptions& o) {
  const auto t0 = std::chrono::steady_clock::now();
  const int ns = int(Ds.size());
  if (ns == 0 || xs.size() != Ds.size() || bs.size() != Ds.size())
    throw std::invalid_argument("batch: sizes differ");
  hipStream_t st = Ds[0]->stream();
  std::vector<Sys> sys(ns);
  std::vector<int> f32(ns), f64(ns), fs(ns);
  int tot32 = 0, tot64 = 0, tots = 0;
  std::vector<void*> owned;
  auto dalloc = [&](size_t bytes) {
    void* p = nullptr;
    QED3_HIP(hipMalloc(&p, bytes));
    owned.push_back(p);
    return p;
  };
  CGScalars* scal = static_cast<CGScalars*>(dalloc(ns * sizeof(CGScalars)));
  unsigned int* ctr = static_cast<unsigned int*>(dalloc(ns * sizeof(unsigned int)));
  QED3_HIP(hipMemsetAsync(ctr, 0, ns * sizeof(unsigned int), st));
  for (int k = 0; k < ns; k++) {
    Dirac& D = *Ds[k];
    if (!D.has_gauge()) throw std::invalid_argument("batch: a system has no gauge field");
    Sys& S = sys[k];
    S.v32 = D.view<float>();
    S.v64 = D.view<double>();
    S.ns = D.tables().max_nn;
    S.n = D.points();
    S.nxr = (D.V() + kBlock - 1) / kBlock;
    S.ts32 = std::max(1, std::min(D.tile<float>(false).ts, D.Lt()));
    S.ts64 = std::max(1, std::min(D.tile<double>(false).ts, D.Lt()));
    S.nst32 = (D.Lt() + S.ts32 - 1) / S.ts32;
    S.nst64 = (D.Lt() + S.ts64 - 1) / S.ts64;
    S.blk32 = f32[k] = tot32;
    S.nblk32 = S.nxr * S.nst32;
    tot32 += S.nblk32;
    S.blk64 = f64[k] = tot64;
    S.nblk64 = S.nxr * S.nst64;
    tot64 += S.nblk64;
    S.sblk = fs[k] = tots;
    S.nsblk = int(std::min<size_t>((S.n + kBlock - 1) / kBlock, 2048));
    tots += S.nsblk;
    S.r = static_cast<Spin<float>*>(dalloc(S.n * sizeof(Spin<float>)));
    S.p = static_cast<Spin<float>*>(dalloc(S.n * sizeof(Spin<float>)));
    S.w = static_cast<Spin<float>*>(dalloc(S.n * sizeof(Spin<float>)));
    S.xin = static_cast<Spin<float>*>(dalloc(S.n * sizeof(Spin<float>)));
    S.t = static_cast<Spin<double>*>(dalloc(S.n * sizeof(Spin<double>)));
    S.x = xs[k];
    S.b = bs[k];
    S.sc = scal + k;
    S.counter = ctr + k;
  }
  Sys* dsys = static_cast<Sys*>(dalloc(ns * sizeof(Sys)));
  int* d32 = static_cast<int*>(dalloc(ns * sizeof(int)));
  int* d64 = static_cast<int*>(dalloc(ns * sizeof(int)));
  int* dss = static_cast<int*>(dalloc(ns * sizeof(int)));
  double2* partials = static_cast<double2*>(dalloc(std::max({tot32, tot64, tots}) * sizeof(double2)));
  QED3_HIP(hipMemcpyAsync(dsys, sys.data(), ns * sizeof(Sys), hipMemcpyHostToDevice, st));
  QED3_HIP(hipMemcpyAsync(d32, f32.data(), ns * sizeof(int), hipMemcpyHostToDevice, st));
  QED3_HIP(hipMemcpyAsync(d64, f64.data(), ns * sizeof(int), hipMemcpyHostToDevice, st));
  QED3_HIP(hipMemcpyAsync(dss, fs.data(), ns * sizeof(int), hipMemcpyHostToDevice, st));
  const double tol2 = o.tol * o.tol, delta2 = o.delta * o.delta;
  hipLaunchKernelGGL(k_batch_stream<0>, dim3(tots), dim3(kBlock), 0, st, dsys, dss, ns, partials, tol2,
                     delta2, o.maxiter);
  QED3_HIP(hipGetLastError());
  auto one_iter = [&]() {
    hipLaunchKernelGGL(k_batch_stencil<kK1>, dim3(tot32), dim3(kBlock), 0, st, dsys, d32, ns, partials);
    hipLaunchKernelGGL(k_batch_finalize<kK1>, dim3(ns), dim3(kBlock), 0, st, dsys, partials);
    hipLaunchKernelGGL(k_batch_stencil<kK2>, dim3(tot32), dim3(kBlock), 0, st, dsys, d32, ns, partials);
    hipLaunchKernelGGL(k_batch_finalize<kK2>, dim3(ns), dim3(kBlock), 0, st, dsys, partials);
    hipLaunchKernelGGL(k_batch_stream<1>, dim3(tots), dim3(kBlock), 0, st, dsys, dss, ns, partials, tol2,
                       delta2, o.maxiter);
    hipLaunchKernelGGL(k_batch_stencil<kU2>, dim3(tot64), dim3(kBlock), 0, st, dsys, d64, ns, partials);
    hipLaunchKernelGGL(k_batch_stencil<kU3>, dim3(tot64), dim3(kBlock), 0, st, dsys, d64, ns, partials);
    hipLaunchKernelGGL(k_batch_finalize<kU3>, dim3(ns), dim3(kBlock), 0, st, dsys, partials);
    hipLaunchKernelGGL(k_batch_stream<2>, dim3(tots), dim3(kBlock), 0, st, dsys, dss, ns, partials, tol2,
                       delta2, o.maxiter);
  };
  const int chunk = o.chunk > 0 ? o.chunk : 8;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (o.use_graph && std::getenv("QED3_NO_GRAPH") == nullptr) {
    QED3_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < chunk; i++) one_iter();
    QED3_HIP(hipStreamEndCapture(st, &graph));
    QED3_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  }
  std::vector<CGScalars> hs(ns);
  const int max_chunks = o.maxiter / chunk + 4;
  for (int c = 0; c < max_chunks; c++) {
    if (exec)
      QED3_HIP(hipGraphLaunch(exec, st));
    else
      for (int i = 0; i < chunk; i++) one_iter();
    QED3_HIP(hipMemcpyAsync(hs.data(), scal, ns * sizeof(CGScalars), hipMemcpyDeviceToHost, st));
    QED3_HIP(hipStreamSynchronize(st));
    bool all = true;
    for (auto& s : hs) all = all && s.done != 0;
    if (all) break;
  }
  hipLaunchKernelGGL(k_batch_stream<3>, dim3(tots), dim3(kBlock), 0, st, dsys, dss, ns, partials, tol2,
                     delta2, o.maxiter);
  QED3_HIP(hipMemcpyAsync(hs.data(), scal, ns * sizeof(CGScalars), hipMemcpyDeviceToHost, st));
  QED3_HIP(hipStreamSynchronize(st));
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  for (void* p : owned) (void)hipFree(p);
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::vector<CGResult> out(ns);
  for (int k = 0; k < ns; k++) {
    out[k].iterations = hs[k].iter;
    out[k].reliable_updates = hs[k].n_ru;
    out[k].status = hs[k].done == 1 ? 0 : (hs[k].done == 2 ? 4 : 3);
    out[k].true_resid = true_residual(*Ds[k], xs[k], bs[k]);
    Ds[k]->release_scratch();  // many contexts: do not keep a residual's temporaries on each
    if (out[k].status == 0 && !(out[k].true_resid <= o.tol * 1.0000001)) out[k].status = 3;
    out[k].seconds = secs;
  }
  return out;
}

}  // namespace qed3
