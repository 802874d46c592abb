// Device-driven FP64 and mixed FP32/FP64 CG on D^dagger D (see cg.hpp).
#include <algorithm>
#include <array>
#include <chrono>
#include <functional>
#include <memory>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "cg.hpp"
#include "cg_epi.hpp"
#include "cg_drive.hpp"
#include "stencil.hpp"

namespace qed3 {

namespace {

using namespace cgepi;
using namespace cgdrive;

template <typename T>
__global__ __launch_bounds__(256) void k_init(const Spin<double>* __restrict__ b, Spin<double>* x64,
                                              Spin<T>* xin, Spin<T>* r, Spin<T>* p, size_t n,
                                              GridReduce gr) {
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
  store_partial(acc, gr.partials, blockIdx.x);
}

// finaliser of k_init: solver scalars from |b|^2
struct FinStart {
  CGScalars* sc;
  double tol2, delta2;
  int maxiter;
  __device__ bool skip() const { return false; }
  __device__ void operator()(double2 t) const {
    sc->mu = t.x;
    sc->bnorm2 = t.x;
    sc->maxr2 = t.x;
    sc->true_r2 = t.x;
    sc->scale = 1.0;
    sc->rescale = 0;
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
  }
};

template <typename T>
__global__ __launch_bounds__(256) void k_flush(Spin<double>* x64, Spin<T>* xin, size_t n,
                                               const CGScalars* sc, int always) {
  if (!always && (sc->done != 0 || sc->need_ru == 0)) return;
  const double s = sc->scale;  // 1 unless the inner fields are stored scaled
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x) {
    const Spin<T> a = xin[i];
    Spin<double> xx = x64[i];
    xx.a.x += s * a.a.x;
    xx.a.y += s * a.a.y;
    xx.b.x += s * a.b.x;
    xx.b.y += s * a.b.y;
    x64[i] = xx;
    Spin<T> z;
    z.a = {T(0), T(0)};
    z.b = {T(0), T(0)};
    xin[i] = z;
  }
}

// finaliser of k_init_h
struct FinStartH {
  CGScalars* sc;
  double ncomp, tol2, delta2;
  int maxiter;
  __device__ bool skip() const { return false; }
  __device__ void operator()(double2 t) const { half_start(sc, t.x, ncomp, tol2, delta2, maxiter); }
};

// r = b / scale in half (FP64 arithmetic, so any finite b is representable)
__global__ __launch_bounds__(256) void k_store_b0(const Spin<double>* __restrict__ b, SpinH* r, size_t n,
                                                  const CGScalars* sc, const phase16_t* rot) {
  if (sc->done != 0) return;
  const double inv = 1.0 / sc->scale;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x)
    half_store_b(b, r, inv, i, rot);
}

struct RingAll {
  SpinH* p[kXGroup];
};

// r = rfac * rt in half after a reliable update (rt holds the true residual at the
// outgoing scale, rfac = old scale / new scale).  With copy_last the update ran at the end
// of a chunk, so the next iteration starts a group (position 0, reading its p_old from the
// last slot): the direction of the last iteration run is copied there.
__global__ __launch_bounds__(256) void k_rescale(const Spin<float>* __restrict__ rt, SpinH* r, RingAll ring,
                                                 int copy_last, size_t n, const CGScalars* sc,
                                                 const phase16_t* rot) {
  if (sc->done != 0 || sc->rescale == 0) return;
  const float inv = float(sc->rfac);
  const int last = sc->last_pos;
  const SpinH* src = (copy_last && last != kXGroup - 1) ? ring.p[last] : nullptr;
  SpinH* dst = ring.p[kXGroup - 1];
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x) {
    half_rescale_point(rt, r, inv, i, rot);
    if (src) dst[i] = src[i];
  }
}

__global__ __launch_bounds__(256) void k_resid_norms(const Spin<double>* __restrict__ b,
                                                     const Spin<double>* __restrict__ ax, size_t n,
                                                     GridReduce gr) {
  double2 acc = {0.0, 0.0};
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x) {
    const Spin<double> bb = b[i], a = ax[i];
    const double d0 = bb.a.x - a.a.x, d1 = bb.a.y - a.a.y, d2 = bb.b.x - a.b.x, d3 = bb.b.y - a.b.y;
    acc.x += d0 * d0 + d1 * d1 + d2 * d2 + d3 * d3;
    acc.y += norm2(bb);
  }
  store_partial(acc, gr.partials, blockIdx.x);
}

struct FinOut {
  double2* out;
  __device__ bool skip() const { return false; }
  __device__ void operator()(double2 t) const { *out = t; }
};

template <typename T>
struct Work {
  Spin<T>* r = nullptr;
  Spin<T>* p = nullptr;
  Spin<T>* p2 = nullptr;  // the other buffer of the double-buffered search direction
  Spin<T>* w = nullptr;
  Spin<T>* xin = nullptr;
  Spin<double>* t = nullptr;
};


template <typename T>
CGResult run_cg(Dirac& D, Spin<double>* x, const Spin<double>* b, const CGOptions& o) {
  constexpr bool kMixed = std::is_same<T, float>::value;
  const size_t n = D.points();
  hipStream_t st = D.stream();
  auto t0 = std::chrono::steady_clock::now();
  Work<T> W;
  const size_t fb = n * sizeof(Spin<T>);
  W.r = static_cast<Spin<T>*>(D.scratch(0, fb));
  W.p = static_cast<Spin<T>*>(D.scratch(1, fb));
  W.p2 = static_cast<Spin<T>*>(D.scratch(2, fb));
  W.w = static_cast<Spin<T>*>(D.scratch(3, fb));
  if (kMixed) {
    W.xin = static_cast<Spin<T>*>(D.scratch(4, fb));
    W.t = static_cast<Spin<double>*>(D.scratch(5, n * sizeof(Spin<double>)));
  } else {
    W.xin = reinterpret_cast<Spin<T>*>(x);
    W.t = reinterpret_cast<Spin<double>*>(W.w);
  }
  CGScalars* sc = static_cast<CGScalars*>(D.scratch(7, sizeof(CGScalars)));
  CGScalars* hsc = D.host_scalars();

  const GridReduce gr = D.reducer();
  const int nbk = int(stream_blocks(n));
  const int ngate = int(gated_blocks(n));
  const double tol2 = o.tol * o.tol;
  const double delta2 = kMixed ? o.delta * o.delta : 0.0;
  hipLaunchKernelGGL(k_init<T>, dim3(nbk), dim3(256), 0, st, b, x, W.xin, W.r, W.p, n, gr);
  hipLaunchKernelGGL(k_finalize<FinStart>, dim3(1), dim3(256), 0, st, gr.partials, unsigned(nbk),
                     FinStart{sc, tol2, delta2, o.maxiter});
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
  EventTimer timer(o.times);
  auto one_iter = [&](int i, int k) {
    const int par = i & 1;
    Spin<T>* pold = P[par];
    Spin<T>* pnew = P[par ^ 1];
    EpiK1P<T> k1;
    k1.w = W.w;
    k1.sc = sc;
    k1.pnew = pnew;
    launch_stencil_src<T, false>(vin, nn, SrcPUpdate<T>{W.r, pold, sc}, k1, D.tile<T>(false), gr, st, hvf,
                                 timer.pair(0, k));
    launch_stencil_src<T, true>(vin, nn, SrcPlain<T>{W.w}, EpiK2<T>{W.xin, W.r, pnew, sc, T(0)},
                                D.tile<T>(true), gr, st, hvt, timer.pair(1, k));
    if (kMixed)
      hipLaunchKernelGGL(k_flush<T>, dim3(ngate), dim3(256), 0, st, x, W.xin, n, sc, 0);
    launch_stencil<double, false>(v64, nn, x, EpiStoreIfRU{W.t, sc}, D.tile<double>(false), gr, st, hv64f);
    launch_stencil<double, true>(v64, nn, W.t, EpiResid<T>{b, W.r, sc}, D.tile<double>(true), gr, st, hv64t);
  };

  drive_chunks(one_iter, [] {}, o, n, st, sc, hsc);
  timer.finish(hsc->iter);
  if (o.times) o.times->ms[2] = o.times->ms[1], o.times->count[2] = o.times->count[1];
  if (kMixed) hipLaunchKernelGGL(k_flush<T>, dim3(nbk), dim3(256), 0, st, x, W.xin, n, sc, 1);
  QED3_HIP(hipMemcpyAsync(hsc, sc, sizeof(CGScalars), hipMemcpyDeviceToHost, st));
  QED3_HIP(hipStreamSynchronize(st));
  CGResult res;
  res.iterations = hsc->iter;
  res.reliable_updates = hsc->n_ru;
  res.status = hsc->done == 1 ? 0 : (hsc->done == 2 ? 4 : 3);
  res.true_resid = final_residual(D, x, b, *hsc);
  res.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (res.status == 0 && !(res.true_resid <= o.tol * 1.0000001)) res.status = 3;
  return res;
}

// FP32 arithmetic on binary16-stored r, p, w (scaled to unit rms at the start and at
// every reliable update), FP32 x_inner, FP64 reliable updates.  Same kernel sequence as
// the mixed solver plus the rescale pass, which only runs after an update.
// The half-storage solve as an object: the constructor queues the start kernels and prepares
// the chunk loop (graph capture), loop() runs it, finish() flushes x and reports; so several
// systems' chunks can be queued on their streams before any host wait.
class HalfCG {
 public:
  HalfCG(Dirac& D, Spin<double>* x, const Spin<double>* b, const CGOptions& o)
      : D_(D), x_(x), b_(b), o_(o), n_(D.points()), st_(D.stream()), t0_(std::chrono::steady_clock::now()),
        timer_(o.times) {
    const size_t hb = n_ * sizeof(SpinH), fb = n_ * sizeof(Spin<float>);
    r_ = static_cast<SpinH*>(D.scratch(0, hb));
    // ring of search directions: iteration k writes ring[k % kXGroup] and reads its
    // predecessor's slot; ring[kXGroup-1] is the zero p_old of the first iteration
    ring_[0] = static_cast<SpinH*>(D.scratch(1, hb));
    ring_[1] = static_cast<SpinH*>(D.scratch(2, hb));
    ring_[2] = static_cast<SpinH*>(D.scratch(12, hb));
    ring_[3] = static_cast<SpinH*>(D.scratch(13, hb));
    w_ = static_cast<SpinH*>(D.scratch(3, hb));
    xin_ = static_cast<Spin<float>*>(D.scratch(4, fb));
    t_ = static_cast<Spin<double>*>(D.scratch(5, n_ * sizeof(Spin<double>)));
    rt_ = static_cast<Spin<float>*>(D.scratch(6, fb));
    sc_ = static_cast<CGScalars*>(D.scratch(7, sizeof(CGScalars)));
    hsc_ = D.host_scalars();
    gr_ = D.reducer();
    nbk_ = int(stream_blocks(n_));
    ngate_ = int(gated_blocks(n_));
    hipLaunchKernelGGL(k_init_h, dim3(nbk_), dim3(256), 0, st_, b, x, xin_, ring_[kXGroup - 1], n_, gr_);
    hipLaunchKernelGGL(k_finalize<FinStartH>, dim3(1), dim3(256), 0, st_, gr_.partials, unsigned(nbk_),
                       FinStartH{sc_, 4.0 * double(n_), o.tol * o.tol, o.delta * o.delta, o.maxiter});
    // the inner fields live in the temporal gauge (exp(i phi) at every point): b is rotated
    // into it here, the true residual at every reliable update, and x out of it at each flush
    vin_ = D.view_half_temporal_gauge();
    rot_ = D.half_gauge_rotation();
    hipLaunchKernelGGL(k_store_b0, dim3(nbk_), dim3(256), 0, st_, b, r_, n_, sc_, rot_);
    QED3_HIP(hipGetLastError());
    v64_ = D.view<double>();
    nn_ = D.tables().max_nn;
    hvf_ = D.halo_for<float>(false);
    hvt_ = D.halo_for<float>(true);
    hv64f_ = D.halo_for<double>(false);
    hv64t_ = D.halo_for<double>(true);
    inv_ncomp_ = 1.0 / (4.0 * double(n_));
    for (int j = 0; j < kXGroup - 1; j++) pend_.p[j] = ring_[j];
    for (int j = 0; j < kXGroup; j++) ring_all_.p[j] = ring_[j];
    // The reliable update (flush the pending x terms, FP64 r = b - D^dag D x, re-store r at
    // the new scale); its kernels exit at once unless an update is due.  It runs once per
    // chunk of iterations: an iteration that asks for it makes the rest of the chunk's
    // iterations exit at once (EpiK1PH/EpiK2H::skip), and the next chunk starts a group with
    // the last direction copied into the group's p_old slot.  With kernel timing (direct
    // launches whose events must each bracket a real iteration) it runs after every
    // iteration instead.
    ru_each_ = o.times != nullptr;
    std::function<void()> tail = [] {};
    if (!ru_each_) tail = [this] { reliable_update(1); };
    loop_ = std::make_unique<ChunkLoop>([this](int i, int k) { one_iter(i, k); }, tail, o, n_, st_, sc_, hsc_);
  }
  ChunkLoop& loop() { return *loop_; }
  CGResult finish() {
    timer_.finish(hsc_->iter);
    hipLaunchKernelGGL(k_flush_h, dim3(nbk_), dim3(256), 0, st_, x_, xin_, pend_, n_, sc_, 1, rot_);
    QED3_HIP(hipMemcpyAsync(hsc_, sc_, sizeof(CGScalars), hipMemcpyDeviceToHost, st_));
    QED3_HIP(hipStreamSynchronize(st_));
    CGResult res;
    res.iterations = hsc_->iter;
    res.reliable_updates = hsc_->n_ru;
    res.status = hsc_->done == 1 ? 0 : (hsc_->done == 2 ? 4 : 3);
    res.true_resid = final_residual(D_, x_, b_, *hsc_);
    res.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
    if (res.status == 0 && !(res.true_resid <= o_.tol * 1.0000001)) res.status = 3;
    return res;
  }

 private:
  template <int POS>
  void k2(int k) {
    EpiK2H<POS> e{};
    e.x = xin_;
    e.r = r_;
    for (int j = 0; j < kXGroup; j++) e.ring[j] = ring_[j];
    e.sc = sc_;
    launch_stencil_src<float, true>(vin_, nn_, SrcHalf{w_}, e, D_.tile<float>(true), gr_, st_, hvt_,
                                    timer_.pair(EpiK2H<POS>::kUpdateX ? 2 : 1, k));
  }
  void reliable_update(int copy_last) {
    hipLaunchKernelGGL(k_flush_h, dim3(ngate_), dim3(256), 0, st_, x_, xin_, pend_, n_, sc_, 0, rot_);
    launch_stencil<double, false>(v64_, nn_, x_, EpiStoreIfRU{t_, sc_}, D_.tile<double>(false), gr_, st_, hv64f_);
    launch_stencil<double, true>(v64_, nn_, t_, EpiResidH{b_, rt_, sc_, inv_ncomp_, 1.0}, D_.tile<double>(true), gr_,
                                 st_, hv64t_);
    hipLaunchKernelGGL(k_rescale, dim3(ngate_), dim3(256), 0, st_, rt_, r_, ring_all_, copy_last, n_, sc_, rot_);
  }
  void one_iter(int i, int k) {
    const int pos = i % kXGroup;
    SpinH* pold = ring_[(pos + kXGroup - 1) % kXGroup];
    SpinH* pnew = ring_[pos];
    launch_stencil_src<float, false>(vin_, nn_, SrcPUpdateH{r_, pold, sc_}, EpiK1PH{w_, pnew, sc_},
                                     D_.tile<float>(false), gr_, st_, hvf_, timer_.pair(0, k));
    static_assert(kXGroup == 4, "one K2 instantiation per group position");
    switch (pos) {
      case 0: k2<0>(k); break;
      case 1: k2<1>(k); break;
      case 2: k2<2>(k); break;
      default: k2<3>(k); break;
    }
    if (ru_each_) reliable_update(0);
  }

  Dirac& D_;
  Spin<double>* x_;
  const Spin<double>* b_;
  CGOptions o_;
  size_t n_;
  hipStream_t st_;
  std::chrono::steady_clock::time_point t0_;
  EventTimer timer_;
  SpinH* r_ = nullptr;
  SpinH* ring_[kXGroup] = {};
  SpinH* w_ = nullptr;
  Spin<float>* xin_ = nullptr;
  Spin<double>* t_ = nullptr;
  Spin<float>* rt_ = nullptr;
  CGScalars* sc_ = nullptr;
  CGScalars* hsc_ = nullptr;
  GridReduce gr_{};
  int nbk_ = 0, ngate_ = 0, nn_ = 0;
  DiracView<float, phase16_t> vin_{};
  const phase16_t* rot_ = nullptr;
  DiracView<double> v64_{};
  HaloView hvf_{}, hvt_{}, hv64f_{}, hv64t_{};
  double inv_ncomp_ = 0;
  PendRing pend_{};
  RingAll ring_all_{};
  bool ru_each_ = false;
  std::unique_ptr<ChunkLoop> loop_;
};

CGResult run_cg_half(Dirac& D, Spin<double>* x, const Spin<double>* b, const CGOptions& o) {
  HalfCG s(D, x, b, o);
  ChunkLoop& loop = s.loop();
  do {
    loop.launch();
    QED3_HIP(hipStreamSynchronize(loop.stream()));
  } while (!loop.finished());
  return s.finish();
}

}  // namespace

double true_residual(Dirac& D, const Spin<double>* x, const Spin<double>* b) {
  const size_t n = D.points();
  hipStream_t st = D.stream();
  Spin<double>* t1 = static_cast<Spin<double>*>(D.scratch(8, n * sizeof(Spin<double>)));
  Spin<double>* t2 = static_cast<Spin<double>*>(D.scratch(9, n * sizeof(Spin<double>)));
  double2* out = static_cast<double2*>(D.scratch(10, sizeof(double2)));
  D.apply_dagd<double>(t2, x, t1, false);
  const unsigned nbk = unsigned(stream_blocks(n));
  hipLaunchKernelGGL(k_resid_norms, dim3(nbk), dim3(256), 0, st, b, t2, n, D.reducer());
  hipLaunchKernelGGL(k_finalize<FinOut>, dim3(1), dim3(256), 0, st, D.reducer().partials, nbk, FinOut{out});
  QED3_HIP(hipGetLastError());
  double2 h;
  QED3_HIP(hipMemcpyAsync(&h, out, sizeof(h), hipMemcpyDeviceToHost, st));
  QED3_HIP(hipStreamSynchronize(st));
  return h.y > 0 ? std::sqrt(h.x / h.y) : 0.0;
}

size_t cg_workspace_bytes(size_t points) {
  // r, p, p2, w, and two fields for a recomputed true residual (FP64), then the scalars and
  // the residual norms, each region 256-byte aligned
  const size_t f = (points * sizeof(Spin<double>) + 255) / 256 * 256;
  return 6 * f + 512 + 256;
}

CGResult cg_solve_f64_workspace(Dirac& D, Spin<double>* x, const Spin<double>* b, const CGOptions& o, void* ws,
                                size_t ws_bytes) {
  const size_t n = D.points();
  if (cg_workspace_bytes(n) > ws_bytes) throw std::invalid_argument("work memory too small for this lattice");
  hipStream_t st = D.stream();
  const auto t0 = std::chrono::steady_clock::now();
  const size_t f = (n * sizeof(Spin<double>) + 255) / 256 * 256;
  unsigned char* base = static_cast<unsigned char*>(ws);
  Spin<double>* r = reinterpret_cast<Spin<double>*>(base);
  Spin<double>* P[2] = {reinterpret_cast<Spin<double>*>(base + f), reinterpret_cast<Spin<double>*>(base + 2 * f)};
  Spin<double>* w = reinterpret_cast<Spin<double>*>(base + 3 * f);
  Spin<double>* t1 = reinterpret_cast<Spin<double>*>(base + 4 * f);
  Spin<double>* t2 = reinterpret_cast<Spin<double>*>(base + 5 * f);
  CGScalars* sc = reinterpret_cast<CGScalars*>(base + 6 * f);
  double2* out = reinterpret_cast<double2*>(base + 6 * f + 512);
  static_assert(sizeof(CGScalars) <= 512, "scalars region");
  CGScalars hsc{};

  // launch shapes: the context's, without LDS staging where its halo tables are not built
  StencilTile tf = D.tile<double>(false), tt = D.tile<double>(true);
  if (tf.lds && !D.has_halo(tf.block)) tf.lds = 0;
  if (tt.lds && !D.has_halo(tt.block)) tt.lds = 0;
  const HaloView hvf = tf.lds ? D.halo(tf.block) : HaloView{}, hvt = tt.lds ? D.halo(tt.block) : HaloView{};
  const GridReduce gr = D.reducer();
  const int nbk = int(stream_blocks(n));
  hipLaunchKernelGGL(k_init<double>, dim3(nbk), dim3(256), 0, st, b, x, x, r, P[0], n, gr);
  hipLaunchKernelGGL(k_finalize<FinStart>, dim3(1), dim3(256), 0, st, gr.partials, unsigned(nbk),
                     FinStart{sc, o.tol * o.tol, 0.0, o.maxiter});
  QED3_HIP(hipGetLastError());
  const DiracView<double> v64 = D.view<double>();
  const int nn = D.tables().max_nn;
  // the iteration of run_cg<double>
  auto one_iter = [&](int i, int) {
    const int par = i & 1;
    EpiK1P<double> k1;
    k1.w = w;
    k1.sc = sc;
    k1.pnew = P[par ^ 1];
    launch_stencil_src<double, false>(v64, nn, SrcPUpdate<double>{r, P[par], sc}, k1, tf, gr, st, hvf);
    launch_stencil_src<double, true>(v64, nn, SrcPlain<double>{w}, EpiK2<double>{x, r, P[par ^ 1], sc, 0.0}, tt, gr,
                                     st, hvt);
    launch_stencil<double, false>(v64, nn, x, EpiStoreIfRU{w, sc}, tf, gr, st, hvf);
    launch_stencil<double, true>(v64, nn, w, EpiResid<double>{b, r, sc}, tt, gr, st, hvt);
  };
  CGOptions od = o;
  od.use_graph = false;
  od.times = nullptr;
  drive_chunks(one_iter, [] {}, od, n, st, sc, &hsc);
  QED3_HIP(hipMemcpyAsync(&hsc, sc, sizeof(CGScalars), hipMemcpyDeviceToHost, st));
  QED3_HIP(hipStreamSynchronize(st));
  CGResult res;
  res.iterations = hsc.iter;
  res.reliable_updates = hsc.n_ru;
  res.status = hsc.done == 1 ? 0 : (hsc.done == 2 ? 4 : 3);
  if (hsc.done == 1 && hsc.n_ru > 0 && hsc.bnorm2 > 0) {
    res.true_resid = std::sqrt(hsc.true_r2 / hsc.bnorm2);
  } else {  // recompute |b - D^dag D x| / |b| in the work memory
    launch_stencil<double, false>(v64, nn, x, EpiStore<double>{t1}, tf, gr, st, hvf);
    launch_stencil<double, true>(v64, nn, t1, EpiStore<double>{t2}, tt, gr, st, hvt);
    hipLaunchKernelGGL(k_resid_norms, dim3(nbk), dim3(256), 0, st, b, t2, n, gr);
    hipLaunchKernelGGL(k_finalize<FinOut>, dim3(1), dim3(256), 0, st, gr.partials, unsigned(nbk), FinOut{out});
    QED3_HIP(hipGetLastError());
    double2 h;
    QED3_HIP(hipMemcpyAsync(&h, out, sizeof(h), hipMemcpyDeviceToHost, st));
    QED3_HIP(hipStreamSynchronize(st));
    res.true_resid = h.y > 0 ? std::sqrt(h.x / h.y) : 0.0;
  }
  res.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (res.status == 0 && !(res.true_resid <= o.tol * 1.0000001)) res.status = 3;
  return res;
}

namespace {
// Pseudo-random source for the tuning solves (a hash of the component index mapped to
// [-1, 1)), so their residuals fall as a Gaussian source's do.
__device__ __forceinline__ double hash_unit(unsigned long long k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdULL;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ULL;
  k ^= k >> 33;
  return double(k >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}
__global__ void k_tune_source(Spin<double>* b, size_t n) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) {
    Spin<double> v;
    v.a = {hash_unit(4 * i), hash_unit(4 * i + 1)};
    v.b = {hash_unit(4 * i + 2), hash_unit(4 * i + 3)};
    b[i] = v;
  }
}
}  // namespace

void tune_cg_tiles(Dirac& D) {
  if (D.cg_tuned || !D.has_gauge()) return;
  const size_t n = D.points();
  hipStream_t st = D.stream();
  Spin<double>*x = nullptr, *b = nullptr;
  QED3_HIP(hipMalloc((void**)&x, n * sizeof(Spin<double>)));
  QED3_HIP(hipMalloc((void**)&b, n * sizeof(Spin<double>)));
  hipLaunchKernelGGL(k_tune_source, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, b, n);
  QED3_HIP(hipGetLastError());
  CGOptions o;
  o.tol = 1e-10;
  o.maxiter = 400;  // bounds the tuning time on badly conditioned operators
  o.mixed = o.half = true;
  // The first candidate, which the others must beat by 1%, is the whole time extent per
  // block (up to 64 slices): then every block of an XCD steps through the same slice and the
  // halos it gathers stay in that XCD's L2.  The bare-D pick and the other shapes follow.
  const StencilTile bare = D.tile<float>(false);
  const StencilTile start{256, std::min(64, D.Lt()), 1, 1};
  std::vector<StencilTile> cands = {start};
  auto same = [](const StencilTile& a, const StencilTile& b) {
    return a.block == b.block && a.ts == b.ts && a.remap == b.remap && a.lds == b.lds;
  };
  if (!same(bare, start)) cands.push_back(bare);
  for (int block : {128, 256})
    for (int ts : {8, 16, 32, 64}) {
      const StencilTile t{block, ts, 1, 1};
      if (ts <= D.Lt() && !same(t, start) && !same(t, bare)) cands.push_back(t);
    }
  // Cost of a candidate: the wall time of a complete half-storage solve to 1e-10 of a
  // pseudo-random source (graph replays, reliable updates and all), after a warm-up solve
  // (halo tables of a new range size are built there).  Short fixed-length solves timed per
  // kernel ranked the shapes differently from complete solves (dual n=128 Lt=64: 16 slices
  // per block ahead there, 64 slices 5% faster per complete solve), so complete solves are
  // what is timed.  Two passes over the candidates, the second in reverse order, each
  // candidate keeping its lower time; a later candidate replaces the best only when it is
  // more than 1% faster, so timing noise does not flip near-ties.
  const int nc = int(cands.size());
  std::vector<double> cost(nc, 1e30);
  for (int pass = 0; pass < 2; pass++)
    for (int c = 0; c < nc; c++) {
      const int k = pass == 0 ? c : nc - 1 - c;
      const StencilTile& t = cands[k];
      D.set_tile(true, false, t);
      D.set_tile(true, true, t);
      (void)cg_solve(D, x, b, o);
      QED3_HIP(hipStreamSynchronize(st));
      const auto t0 = std::chrono::steady_clock::now();
      const CGResult r = cg_solve(D, x, b, o);
      const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      cost[k] = std::min(cost[k], s);
      if (std::getenv("QED3_TUNE_TRACE"))
        std::fprintf(stderr, "tune_cg_tiles: pass %d block %d ts %d: %.3f ms per solve (%d iterations)\n",
                     pass, t.block, t.ts, 1e3 * s, r.iterations);
    }
  StencilTile best = start;
  double best_s = 1e30;
  for (int k = 0; k < nc; k++)
    if (cost[k] < 0.99 * best_s) {
      best_s = cost[k];
      best = cands[k];
    }
  D.set_tile(true, false, best);
  D.set_tile(true, true, best);
  (void)hipFree(x);
  (void)hipFree(b);
  D.cg_tuned = true;
}

CGResult cg_solve(Dirac& D, Spin<double>* x, const Spin<double>* b, const CGOptions& o) {
  if (o.half) return run_cg_half(D, x, b, o);
  if (o.mixed) return run_cg<float>(D, x, b, o);
  return run_cg<double>(D, x, b, o);
}

std::vector<CGResult> cg_solve_half_many(const std::vector<Dirac*>& Ds, const std::vector<Spin<double>*>& xs,
                                         const std::vector<const Spin<double>*>& bs, const CGOptions& o) {
  const size_t ns = Ds.size();
  if (xs.size() != ns || bs.size() != ns) throw std::invalid_argument("systems: sizes differ");
  std::vector<std::unique_ptr<HalfCG>> solves;
  solves.reserve(ns);
  for (size_t k = 0; k < ns; k++) solves.push_back(std::make_unique<HalfCG>(*Ds[k], xs[k], bs[k], o));
  // rounds: one chunk of every unfinished system queued on its own stream, then the polls
  std::vector<size_t> active(ns);
  for (size_t k = 0; k < ns; k++) active[k] = k;
  while (!active.empty()) {
    for (size_t k : active) solves[k]->loop().launch();
    std::vector<size_t> next;
    for (size_t k : active) {
      QED3_HIP(hipStreamSynchronize(solves[k]->loop().stream()));
      if (!solves[k]->loop().finished()) next.push_back(k);
    }
    active.swap(next);
  }
  std::vector<CGResult> out(ns);
  for (size_t k = 0; k < ns; k++) out[k] = solves[k]->finish();
  return out;
}

}  // namespace qed3
