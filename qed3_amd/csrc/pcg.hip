// Flexible CG on D^dagger D preconditioned by M = K K^dagger, K the time-exact Horner
// network (horner.hpp), on the binary16-stored fields of the half-storage CG.
//
// FCG(1) (Notay, SIAM J. Sci. Comput. 22 (2000) 1444): z = M r, p = z + beta p_old with
//   beta = -<A p_old, z> / <p_old, A p_old> = (<r, z> - <r_old, z>) / <r_old, z_old>,
// the second form because A p_old = (r_old - r) / alpha_old; alpha = <r, z> / |D p|^2.
// The correction <r_old, z> (zero for an exactly symmetric fixed M) keeps the directions
// A-conjugate although the binary16 fields make M vary slightly between applications.
//
// One iteration is the network's 2L - 1 layer launches (K^dagger, whose last layer also
// runs K's first, then K, whose last layer computes the two products below and the scalar
// update) and the two stencil kernels of the half-storage CG: K1 forms p = z + beta p_old
// while gathering, stores it and applies D, K2 applies D^dagger, updates r in place and x.
// With H = K^dagger r, <r, z> = |H|^2 and <r_old, z> = <K^dagger r_old, H> = <H_old, H>, so
// the products need neither residual again: H is double-buffered (by the iteration's
// parity, fixed within a chunk of an even number of iterations) and K's layers, which read
// H as their injection anyway, take the previous one beside it in the last layer.  Reliable
// updates, scales and the temporal gauge are those of the half-storage CG (cg.hip); after
// an update beta carries the change of scale (H_old and p_old are kept at the old scale).
#include <chrono>
#include <cmath>

#include "cg_drive.hpp"
#include "cg_epi.hpp"
#include "horner.hpp"
#include "horner_kernels.hpp"
#include "stencil.hpp"

namespace qed3 {
namespace {

using namespace cgepi;
using namespace cgdrive;


// The directions of the last kFGroup iterations stay in a ring of buffers and their x terms
// pending (alpha_pend[slot]) until a reliable update flushes them; terms still pending when
// a group of kFGroup iterations starts to reuse the ring (no update for up to kFGroup
// iterations) are first added to x_inner by a streaming pass of their own (k_fold_f).  With
// an update every three or four iterations few are, and K2 never touches x.
constexpr int kFGroup = 8;
static_assert(kFGroup <= kMaxPend, "pending alphas of a group");

// K2: q = D^dag w; r = rpend r - alpha q in place, |r|^2 and the reliable-update decision;
// the iteration's x term is left pending in its ring slot `pos`
struct EpiK2F {
  static constexpr bool kReduce = true;
  SpinH* r;
  CGScalars* sc;
  int pos;
  float alpha;
  float f;  // the factor the stored r still owes after a reliable update (else 1)
  __device__ bool skip() const { return sc->done != 0 || sc->need_ru != 0; }
  __device__ void begin() {
    alpha = float(sc->alpha);
    f = float(sc->rpend);
  }
  // r is loaded with the slice's other loads (as EpiK2H)
  __device__ SpinH prefetch(size_t i) const { return r[i]; }
  __device__ void operator()(size_t i, const Spin<float>& q, double2& acc) const { (*this)(i, q, acc, r[i]); }
  __device__ void operator()(size_t i, const Spin<float>& q, double2& acc, const SpinH& rpre) const {
    Spin<float> rr = from_half(rpre);
    rr.a.x = f * rr.a.x - alpha * q.a.x;
    rr.a.y = f * rr.a.y - alpha * q.a.y;
    rr.b.x = f * rr.b.x - alpha * q.b.x;
    rr.b.y = f * rr.b.y - alpha * q.b.y;
    const SpinH rh = to_half(rr);
    sth_nt(r + i, rh);
    acc.x += norm2(from_half(rh));
  }
  __device__ void operator()(double2 t) const {
    const double m = t.x;
    sc->alpha_pend[pos] = sc->alpha;
    sc->last_pos = pos;
    sc->iter += 1;
    sc->rpend = 1.0;  // r is re-stored at the current scale
    if (!(m >= 0.0 && m < 1e300)) {
      sc->done = 2;
      return;
    }
    if (m < sc->delta2 * sc->maxr2 || m * sc->scale * sc->scale <= sc->tol2 * sc->bnorm2)
      sc->need_ru = 1;
    else
      sc->maxr2 = fmax(sc->maxr2, m);
    if (sc->iter >= sc->maxiter && !sc->need_ru) sc->done = 3;
  }
};

struct PendRing8 {
  const SpinH* p[kFGroup];
};

// x64 += scale exp(-i phi) (x_inner + sum_j alpha_pend[j] ring[j]) (before a reliable
// update, or always at the end).  Neither x64 nor x_inner is cleared at the start: each is
// read only once it holds a value (sc->x_set, sc->xin_set), and x_inner counts as cleared
// again after the update (its flag is reset by the update's last kernel).
__global__ __launch_bounds__(256) void k_flush_f(Spin<double>* x64, Spin<float>* xin, PendRing8 ring, size_t n,
                                                 const CGScalars* sc, int always, const phase16_t* rot) {
  if (!always && (sc->done != 0 || sc->need_ru == 0)) return;
  float ap[kFGroup];
#pragma unroll
  for (int j = 0; j < kFGroup; j++) ap[j] = float(sc->alpha_pend[j]);
  const double s = sc->scale;
  const bool xset = sc->x_set != 0, iset = sc->xin_set != 0;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    Spin<float> a;
    if (iset) {
      a = xin[i];
    } else {
      a.a = {0.f, 0.f};
      a.b = {0.f, 0.f};
    }
#pragma unroll
    for (int j = 0; j < kFGroup; j++) {
      if (ap[j] != 0.f) {  // uniform
        const Spin<float> pp = ldh(ring.p[j] + i);
        a.a.x += ap[j] * pp.a.x;
        a.a.y += ap[j] * pp.a.y;
        a.b.x += ap[j] * pp.b.x;
        a.b.y += ap[j] * pp.b.y;
      }
    }
    rotate(a, rot, i, true);
    Spin<double> xx;
    if (xset) {
      xx = x64[i];
    } else {
      xx.a = {0.0, 0.0};
      xx.b = {0.0, 0.0};
    }
    xx.a.x += s * a.a.x;
    xx.a.y += s * a.a.y;
    xx.b.x += s * a.b.x;
    xx.b.y += s * a.b.y;
    x64[i] = xx;
  }
}

// Before a group reuses the ring: x_inner += sum_j alpha_pend[j] ring[j] over the slots still
// pending (no reliable update since their directions were made); k_fold_done_f then marks
// them folded (one thread, after every block of the fold has read the flags).  Both exit at
// once when no slot is pending, or the solve is done (the final flush takes the terms).
__device__ __forceinline__ bool fold_due(const CGScalars* sc) {
  bool any = false;
#pragma unroll
  for (int j = 0; j < kFGroup; j++) any |= sc->alpha_pend[j] != 0.0;
  return sc->done == 0 && any;
}
__global__ __launch_bounds__(256) void k_fold_f(Spin<float>* xin, PendRing8 ring, size_t n, const CGScalars* sc) {
  if (!fold_due(sc)) return;
  float ap[kFGroup];
#pragma unroll
  for (int j = 0; j < kFGroup; j++) ap[j] = float(sc->alpha_pend[j]);
  const bool iset = sc->xin_set != 0;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    Spin<float> xx;
    if (iset) {
      xx = xin[i];
    } else {
      xx.a = {0.f, 0.f};
      xx.b = {0.f, 0.f};
    }
#pragma unroll
    for (int j = 0; j < kFGroup; j++) {
      if (ap[j] != 0.f) {  // uniform
        const Spin<float> pp = ldh(ring.p[j] + i);
        xx.a.x += ap[j] * pp.a.x;
        xx.a.y += ap[j] * pp.a.y;
        xx.b.x += ap[j] * pp.b.x;
        xx.b.y += ap[j] * pp.b.y;
      }
    }
    xin[i] = xx;
  }
}
__global__ void k_fold_done_f(CGScalars* sc) {
  if (!fold_due(sc)) return;
  for (int j = 0; j < kFGroup; j++) sc->alpha_pend[j] = 0.0;
  sc->xin_set = 1;
}

// Start: |b|^2 and p_old = 0 (x and x_inner are not cleared: see k_flush_f)
__global__ __launch_bounds__(256) void k_init_f(const Spin<double>* __restrict__ b, SpinH* p, size_t n,
                                                GridReduce gr) {
  double2 acc = {0.0, 0.0};
  Spin<float> z;
  z.a = {0.f, 0.f};
  z.b = {0.f, 0.f};
  const SpinH zh = to_half(z);
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    acc.x += norm2(ld(b + i));
    p[i] = zh;
  }
  store_partial(acc, gr.partials, blockIdx.x);
}

// the network's layers wait, like K1 and K2, for a pending reliable update
struct EpiSkip {
  static constexpr bool kReduce = false;
  const CGScalars* sc;
  __device__ bool skip() const { return sc->done != 0 || sc->need_ru != 0; }
  __device__ void operator()(size_t, const Spin<float>&, double2&) const {}
  __device__ void operator()(double2) const {}
};

// Last layer of K, injecting H = K^dagger r as stored beside H_old (hk::SrcDots), which
// reduces t = (|H|^2, <H_old, H>) = (<r, z>, <r_old, z>): the FCG scalars.  After a reliable
// update r is the true residual at the new scale while H_old and p_old are at the old one
// (rfac = old / new): beta = (|H|^2 - rfac <H_old, H>) / (rfac rho_old).
struct EpiFcg {
  static constexpr bool kReduce = true;
  CGScalars* sc;
  __device__ bool skip() const { return sc->done != 0 || sc->need_ru != 0; }
  __device__ void operator()(size_t, const Spin<float>&, double2&) const {}
  __device__ void operator()(double2 t) const {
    const double A = t.x, Bv = t.y;
    if (!(A > 0.0 && A < 1e300) || (!sc->fresh && !(fabs(Bv) < 1e300))) {
      sc->done = 2;  // M r must be a descent direction: <r, M r> > 0
      return;
    }
    double beta = 0.0;
    if (!sc->fresh) {
      const double f = sc->rescale ? sc->rfac : 1.0;
      beta = (A - f * Bv) / (f * sc->rho);
    }
    sc->beta = beta;
    sc->rho = A;
    sc->mu = A;  // K1: alpha = mu / |D p|^2
    sc->fresh = 0;
  }
};

struct FinStartF {
  CGScalars* sc;
  double ncomp, tol2, delta2;
  int maxiter;
  __device__ bool skip() const { return false; }
  __device__ void operator()(double2 t) const {
    half_start(sc, t.x, ncomp, tol2, delta2, maxiter);
    sc->fresh = 1;
    sc->rpend = 1.0;
    sc->x_set = 0;
    sc->xin_set = 0;
    sc->rho = 0.0;
  }
};

// r = exp(i phi) b / scale in binary16; H = 0 (no H_old yet)
__global__ __launch_bounds__(256) void k_store_b0_f(const Spin<double>* __restrict__ b, SpinH* r, SpinH* h,
                                                    size_t n, const CGScalars* sc, const phase16_t* rot) {
  if (sc->done != 0) return;
  const double inv = 1.0 / sc->scale;
  Spin<float> z;
  z.a = {0.f, 0.f};
  z.b = {0.f, 0.f};
  const SpinH zh = to_half(z);
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    half_store_b(b, r, inv, i, rot);
    h[i] = zh;
  }
}

// U3 of the preconditioned solve: r = b - D^dag t in FP64, stored in binary16 in the temporal
// gauge at the outgoing scale (its values near delta, so a pass that re-stores it at the new
// scale is not needed: its readers, K^dagger's layers and K2, apply rfac until K2 has
// re-stored it); |r|^2 decides convergence, the new scale and beta as for the half-storage CG
struct EpiResidF {
  static constexpr bool kReduce = true;
  const Spin<double>* b;
  SpinH* r;
  CGScalars* sc;
  const phase16_t* rot;
  double inv_ncomp;
  double inv_so;
  __device__ bool skip() const { return sc->done != 0 || sc->need_ru == 0; }
  __device__ void begin() { inv_so = 1.0 / sc->scale; }
  __device__ void operator()(size_t i, const Spin<double>& v, double2& acc) const {
    const Spin<double> bb = ld(b + i);
    Spin<double> rr;
    rr.a = {bb.a.x - v.a.x, bb.a.y - v.a.y};
    rr.b = {bb.b.x - v.b.x, bb.b.y - v.b.y};
    acc.x += norm2(rr);
    Spin<float> rf;
    rf.a = {float(rr.a.x * inv_so), float(rr.a.y * inv_so)};
    rf.b = {float(rr.b.x * inv_so), float(rr.b.y * inv_so)};
    rotate(rf, rot, i, false);
    sth_nt(r + i, to_half(rf));
  }
  __device__ void operator()(double2 t) const {
    sc->x_set = 1;    // the update's flush has written x
    sc->xin_set = 0;  // ... and taken x_inner
    EpiResidH{b, nullptr, sc, inv_ncomp, 1.0}(t);
    if (sc->done == 0) sc->rpend = sc->rfac;
  }
};

}  // namespace

namespace {
// FP64 field (original gauge) -> binary16 in the temporal gauge, and back
__global__ __launch_bounds__(256) void k_to_tg16(const Spin<double>* __restrict__ in, const phase16_t* rot,
                                                 SpinH* out, size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const Spin<double> d = in[i];
    Spin<float> v;
    v.a = {float(d.a.x), float(d.a.y)};
    v.b = {float(d.b.x), float(d.b.y)};
    cgepi::rotate(v, rot, i, false);
    out[i] = to_half(v);
  }
}
__global__ __launch_bounds__(256) void k_from_tg16(const SpinH* __restrict__ in, const phase16_t* rot,
                                                   Spin<double>* out, size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    Spin<float> v = from_half(in[i]);
    cgepi::rotate(v, rot, i, true);
    Spin<double> d;
    d.a = {double(v.a.x), double(v.a.y)};
    d.b = {double(v.b.x), double(v.b.y)};
    out[i] = d;
  }
}
}  // namespace

namespace {
// M v = K (K^dagger v) by the parallel-in-time layers: K^dagger's layers inject src and
// end in H; its last layer also runs K's first layer (L >= 2) into a G slot; K's layers
// inject H and end in Z, the last one injecting src_last (H, possibly with a side field)
// with epi_last.  G[2] alternate as the layers' outputs; the other layers carry epi.
template <class Src, class Epi, class SrcLast, class EpiLast>
void precond_layers(Horner& K, const HornerView& hv, const DiracView<float, phasepair_t>& vpairs, const HaloView& halo,
                    const Src& src, SpinH* H, SpinH* const* G, SpinH* Z, const Epi& epi, const SrcLast& src_last,
                    const EpiLast& epi_last, GridReduce gr, hipStream_t st, EventTimer* timer, int k) {
  const int L = K.layers();
  const hk::SrcF<SpinH> src_h{H};
  auto ev = [&](int cls) { return timer ? timer->pair(cls, k) : nullptr; };
  const bool fuse = L >= 2;
  const std::complex<double> c0 = K.params().c[L - 1];
  const float2 ck0 = {float(c0.real()), float(c0.imag())};  // K's first layer
  // timing classes: 0 a first layer, 1 a transport layer of K^dagger, 2 one of K, 3 K^dagger's
  // last layer carrying K's first, 4 K's last layer with epi_last
  const SpinH* prev = nullptr;
  for (int l = 0; l < L; l++) {
    const std::complex<double> c = K.params().c[L - 1 - l];
    const float2 cj = {float(c.real()), float(-c.imag())};
    const bool last = l == L - 1;
    SpinH* out = last ? H : G[l & 1];
    SpinH* out2 = (last && fuse) ? G[(L - 1) & 1] : nullptr;
    if (l == 0)
      hk::launch_horner_layer_tc<true, true>(hv, vpairs, halo, src, prev, out, cj, epi, gr, st, ev(out2 ? 3 : 0), out2,
                                             ck0);
    else
      hk::launch_horner_layer_tc<true, false>(hv, vpairs, halo, src, prev, out, cj, epi, gr, st, ev(out2 ? 3 : 1),
                                              out2, ck0);
    prev = out;
  }
  prev = fuse ? G[(L - 1) & 1] : nullptr;
  for (int l = fuse ? 1 : 0; l < L; l++) {
    const std::complex<double> c = K.params().c[L - 1 - l];
    const float2 cj = {float(c.real()), float(c.imag())};
    const bool last = l == L - 1;
    SpinH* out = last ? Z : G[(l + L - 1) & 1];
    if (l == 0 && last)
      hk::launch_horner_layer_tc<false, true>(hv, vpairs, halo, src_last, prev, out, cj, epi_last, gr, st, ev(4));
    else if (l == 0)
      hk::launch_horner_layer_tc<false, true>(hv, vpairs, halo, src_h, prev, out, cj, epi, gr, st, ev(0));
    else if (last)
      hk::launch_horner_layer_tc<false, false>(hv, vpairs, halo, src_last, prev, out, cj, epi_last, gr, st, ev(4));
    else
      hk::launch_horner_layer_tc<false, false>(hv, vpairs, halo, src_h, prev, out, cj, epi, gr, st, ev(2));
    prev = out;
  }
}
}  // namespace

void horner_apply_solver(Horner& K, int which, Spin<double>* out, const Spin<double>* in) {
  if (which < 0 || which > 2) throw std::invalid_argument("which: 0 = K, 1 = K^dagger, 2 = K K^dagger");
  Dirac& D = K.dirac();
  const size_t n = D.points();
  hipStream_t st = D.stream();
  const size_t hb = n * sizeof(SpinH);
  (void)D.view_half_temporal_gauge();
  const DiracView<float, phasepair_t> vpairs = D.view_half_tg_pairs();
  const phase16_t* rot = D.half_gauge_rotation();
  const HornerView hv = K.view();
  const HaloView halo = D.halo(hk::tc_sites(hv.Lt));
  SpinH* V16 = static_cast<SpinH*>(K.work(6, hb));
  SpinH* G[2] = {static_cast<SpinH*>(K.work(7, hb)), static_cast<SpinH*>(K.work(8, hb))};
  const int nbk = int(stream_blocks(n));
  hipLaunchKernelGGL(k_to_tg16, dim3(nbk), dim3(256), 0, st, in, rot, V16, n);
  const hk::SrcF<SpinH> src{V16};
  const hk::EpiNone none{};
  const int L = K.layers();
  const SpinH* prev = nullptr;
  if (which == 2) {  // the solver's own sequence, K's first layer fused into K^dagger's last
    SpinH* H = static_cast<SpinH*>(K.work(9, hb));
    SpinH* Z = static_cast<SpinH*>(K.work(10, hb));
    precond_layers(K, hv, vpairs, halo, src, H, G, Z, none, hk::SrcF<SpinH>{H}, none, D.reducer(), st, nullptr, 0);
    prev = Z;
  } else {
    const bool adj = which == 1;
    for (int l = 0; l < L; l++) {
      const std::complex<double> c = K.params().c[L - 1 - l];
      const float2 cj = {float(c.real()), float(adj ? -c.imag() : c.imag())};
      SpinH* dst = G[l & 1];
      if (adj) {
        if (l == 0)
          hk::launch_horner_layer_tc<true, true>(hv, vpairs, halo, src, prev, dst, cj, none, D.reducer(), st);
        else
          hk::launch_horner_layer_tc<true, false>(hv, vpairs, halo, src, prev, dst, cj, none, D.reducer(), st);
      } else {
        if (l == 0)
          hk::launch_horner_layer_tc<false, true>(hv, vpairs, halo, src, prev, dst, cj, none, D.reducer(), st);
        else
          hk::launch_horner_layer_tc<false, false>(hv, vpairs, halo, src, prev, dst, cj, none, D.reducer(), st);
      }
      prev = dst;
    }
  }
  hipLaunchKernelGGL(k_from_tg16, dim3(nbk), dim3(256), 0, st, prev, rot, out, n);
  QED3_HIP(hipGetLastError());
  QED3_HIP(hipStreamSynchronize(st));
}

CGResult pcg_solve(Horner& K, Spin<double>* x, const Spin<double>* b, const CGOptions& o,
                   KernelTimes* layer_times) {
  Dirac& D = K.dirac();
  const size_t n = D.points();
  hipStream_t st = D.stream();
  const auto t0 = std::chrono::steady_clock::now();
  const size_t hb = n * sizeof(SpinH), fb = n * sizeof(Spin<float>);
  // the half-storage CG's slots for the stencil-side fields (and half of the ring); the
  // network's own fields and the rest of the ring in K's
  SpinH* ring[kFGroup] = {static_cast<SpinH*>(D.scratch(1, hb)),  static_cast<SpinH*>(D.scratch(2, hb)),
                          static_cast<SpinH*>(D.scratch(12, hb)), static_cast<SpinH*>(D.scratch(13, hb)),
                          static_cast<SpinH*>(K.work(6, hb)),     static_cast<SpinH*>(K.work(7, hb)),
                          static_cast<SpinH*>(K.work(8, hb)),     static_cast<SpinH*>(K.work(9, hb))};
  SpinH* w = static_cast<SpinH*>(D.scratch(3, hb));
  Spin<float>* xin = static_cast<Spin<float>*>(D.scratch(4, fb));
  Spin<double>* t = static_cast<Spin<double>*>(D.scratch(5, n * sizeof(Spin<double>)));
  CGScalars* sc = static_cast<CGScalars*>(D.scratch(7, sizeof(CGScalars)));
  CGScalars* hsc = D.host_scalars();
  SpinH* R = static_cast<SpinH*>(K.work(0, hb));
  SpinH* G[2] = {static_cast<SpinH*>(K.work(2, hb)), static_cast<SpinH*>(K.work(3, hb))};
  // H of the even and the odd iterations
  SpinH* H[2] = {static_cast<SpinH*>(K.work(4, hb)), static_cast<SpinH*>(K.work(1, hb))};
  SpinH* Z = static_cast<SpinH*>(K.work(5, hb));

  const GridReduce gr = D.reducer();
  const int nbk = int(stream_blocks(n));
  const int ngate = int(gated_blocks(n));
  hipLaunchKernelGGL(k_init_f, dim3(nbk), dim3(256), 0, st, b, ring[kFGroup - 1], n, gr);
  hipLaunchKernelGGL(k_finalize<FinStartF>, dim3(1), dim3(256), 0, st, gr.partials, unsigned(nbk),
                     FinStartF{sc, 4.0 * double(n), o.tol * o.tol, o.delta * o.delta, o.maxiter});
  const DiracView<float, phase16_t> vin = D.view_half_temporal_gauge();
  const DiracView<float, phasepair_t> vpairs = D.view_half_tg_pairs();
  const phase16_t* rot = D.half_gauge_rotation();
  hipLaunchKernelGGL(k_store_b0_f, dim3(nbk), dim3(256), 0, st, b, R, H[1], n, sc, rot);
  QED3_HIP(hipGetLastError());

  const DiracView<double> v64 = D.view<double>();
  const int nn = D.tables().max_nn;
  const HaloView hvf = D.halo_for<float>(false), hvt = D.halo_for<float>(true);
  const HaloView hv64f = D.halo_for<double>(false), hv64t = D.halo_for<double>(true);
  const HornerView hv = K.view();
  const HaloView halo = D.halo(hk::tc_sites(hv.Lt));
  const int L = K.layers();
  const double inv_ncomp = 1.0 / (4.0 * double(n));
  PendRing8 pend;
  for (int j = 0; j < kFGroup; j++) pend.p[j] = ring[j];
  EventTimer timer(o.times), ltimer(layer_times);
  const hk::SrcFS src_r{R, &sc->rpend};
  const EpiSkip eskip{sc};
  const EpiFcg efcg{sc};

  // one network application M r = K (K^dagger r): K^dagger's output in H[i & 1], K's in Z
  auto precond = [&](int i, int k) {
    SpinH* h = H[i & 1];
    precond_layers(K, hv, vpairs, halo, src_r, h, G, Z, eskip, hk::SrcDots{h, H[(i & 1) ^ 1]}, efcg, gr, st,
                   &ltimer, k);
  };
  auto k2 = [&](int pos, int k) {
    EpiK2F e{};
    e.r = R;
    e.sc = sc;
    e.pos = pos;
    launch_stencil_src<float, true>(vin, nn, SrcHalf{w}, e, D.tile<float>(true), gr, st, hvt, timer.pair(1, k));
  };
  // The reliable update's kernels follow every iteration and exit at once unless it flagged
  // one: with an update every three to four iterations (delta = 0.001), waiting for a
  // chunk's end would leave the rest of the chunk's iterations idle.
  auto reliable_update = [&] {
    hipLaunchKernelGGL(k_flush_f, dim3(ngate), dim3(256), 0, st, x, xin, pend, n, sc, 0, rot);
    launch_stencil<double, false>(v64, nn, x, EpiStoreIfRU{t, sc}, D.tile<double>(false), gr, st, hv64f);
    launch_stencil<double, true>(v64, nn, t, EpiResidF{b, R, sc, rot, inv_ncomp, 1.0}, D.tile<double>(true), gr, st,
                                 hv64t);
  };
  auto one_iter = [&](int i, int k) {
    const int pos = i % kFGroup;
    SpinH* pold = ring[(pos + kFGroup - 1) % kFGroup];
    SpinH* pnew = ring[pos];
    if (pos == 0) {  // the group starts to reuse the ring
      hipLaunchKernelGGL(k_fold_f, dim3(ngate), dim3(256), 0, st, xin, pend, n, sc);
      hipLaunchKernelGGL(k_fold_done_f, dim3(1), dim3(1), 0, st, sc);
    }
    precond(i, k);
    launch_stencil_src<float, false>(vin, nn, SrcPUpdateH{Z, pold, sc}, EpiK1PH{w, pnew, sc},
                                     D.tile<float>(false), gr, st, hvf, timer.pair(0, k));
    k2(pos, k);
    reliable_update();
  };
  // ring positions run 0 .. kFGroup-1 within every chunk
  CGOptions oc = o;
  oc.chunk = ((o.chunk > 0 ? o.chunk : kFGroup) + kFGroup - 1) / kFGroup * kFGroup;
  drive_chunks(one_iter, [] {}, oc, n, st, sc, hsc);
  timer.finish(hsc->iter);
  ltimer.finish(hsc->iter);
  hipLaunchKernelGGL(k_flush_f, dim3(nbk), dim3(256), 0, st, x, xin, pend, n, sc, 1, rot);
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

}  // namespace qed3
