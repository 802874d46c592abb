// Epilogues of the device-driven CG kernels (shared by the single-system solver and the
// batched solver).  Each reads/writes the solver scalars of its own system.
#pragma once
#include "dirac.hpp"
#include "stencil.hpp"

namespace qed3 {
namespace cgepi {

__device__ __forceinline__ bool finite_pos(double v) { return v > 0.0 && v < 1e300; }

__device__ __forceinline__ void k1_finish(CGScalars* sc, double gam) {
  sc->gam = gam;
  if (!finite_pos(gam)) {
    sc->done = 2;
    return;
  }
  sc->alpha = sc->mu / gam;
}

// K1: w = D p, gam = |w|^2, alpha = mu / gam
template <typename T>
struct EpiK1 {
  static constexpr bool kReduce = true;
  Spin<T>* w;
  CGScalars* sc;
  __device__ bool skip() const { return sc->done != 0; }
  __device__ void begin() {}
  __device__ void operator()(size_t i, const Spin<T>& v, double2& acc) const {
    st(w + i, v);
    acc.x += norm2(v);
  }
  __device__ void operator()(double2 t) const { k1_finish(sc, t.x); }
};

// K1 reading p = r + beta p_old (SrcPUpdate) and storing that p at every point it owns for
// K2 and the next iteration (the search-direction update fused into the matvec)
template <typename T>
struct EpiK1P : EpiK1<T> {
  Spin<T>* pnew;
  __device__ void center(size_t i, const Spin<T>& v) const { st(pnew + i, v); }
};

// K2: q = D^dag w; x += alpha p; r -= alpha q; |r|^2 and the reliable-update decision
template <typename T>
struct EpiK2 {
  static constexpr bool kReduce = true;
  Spin<T>* x;
  Spin<T>* r;
  const Spin<T>* p;
  CGScalars* sc;
  T alpha;
  __device__ bool skip() const { return sc->done != 0; }
  __device__ void begin() { alpha = T(sc->alpha); }
  __device__ void operator()(size_t i, const Spin<T>& q, double2& acc) const {
    const Spin<T> pp = ld(p + i);
    Spin<T> xx = ld(x + i), rr = ld(r + i);
    xx.a.x += alpha * pp.a.x;
    xx.a.y += alpha * pp.a.y;
    xx.b.x += alpha * pp.b.x;
    xx.b.y += alpha * pp.b.y;
    rr.a.x -= alpha * q.a.x;
    rr.a.y -= alpha * q.a.y;
    rr.b.x -= alpha * q.b.x;
    rr.b.y -= alpha * q.b.y;
    st(x + i, xx);
    st(r + i, rr);
    acc.x += norm2(rr);
  }
  __device__ void operator()(double2 t) const {
    const double m = t.x;
    sc->iter += 1;
    if (!(m >= 0.0 && m < 1e300)) {
      sc->done = 2;
      return;
    }
    if (m < sc->delta2 * sc->maxr2 || m <= sc->tol2 * sc->bnorm2) {
      sc->need_ru = 1;
    } else {
      sc->beta = m / sc->mu;
      sc->mu = m;
      sc->maxr2 = fmax(sc->maxr2, m);
    }
    if (sc->iter >= sc->maxiter && !sc->need_ru) sc->done = 3;
  }
};

// U2: t = D x64 (only when an update is due)
struct EpiStoreIfRU {
  static constexpr bool kReduce = false;
  Spin<double>* out;
  const CGScalars* sc;
  __device__ bool skip() const { return sc->done != 0 || sc->need_ru == 0; }
  __device__ void begin() {}
  __device__ void operator()(size_t i, const Spin<double>& v, double2&) const { st(out + i, v); }
  __device__ void operator()(double2) const {}
};

// U3: r64 = b - D^dag t; r_inner = r64; |r64|^2 decides convergence and beta
template <typename T>
struct EpiResid {
  static constexpr bool kReduce = true;
  const Spin<double>* b;
  Spin<T>* r;
  CGScalars* sc;
  __device__ bool skip() const { return sc->done != 0 || sc->need_ru == 0; }
  __device__ void begin() {}
  __device__ void operator()(size_t i, const Spin<double>& v, double2& acc) const {
    const Spin<double> bb = ld(b + i);
    Spin<double> rr;
    rr.a = {bb.a.x - v.a.x, bb.a.y - v.a.y};
    rr.b = {bb.b.x - v.b.x, bb.b.y - v.b.y};
    Spin<T> ri;
    ri.a = {T(rr.a.x), T(rr.a.y)};
    ri.b = {T(rr.b.x), T(rr.b.y)};
    st(r + i, ri);
    acc.x += norm2(rr);
  }
  __device__ void operator()(double2 t) const {
    const double m = t.x;
    sc->need_ru = 0;
    sc->n_ru += 1;
    sc->true_r2 = m;
    if (!(m >= 0.0 && m < 1e300)) {
      sc->done = 2;
      return;
    }
    if (m <= sc->tol2 * sc->bnorm2) {
      sc->done = 1;
      return;
    }
    sc->beta = m / sc->mu;
    sc->mu = m;
    sc->maxr2 = m;
    if (sc->iter >= sc->maxiter) sc->done = 3;
  }
};

// ---- half-storage CG (r, p, w in binary16 scaled by sc->scale; x_inner in FP32) -----
// Norms and decisions are taken on the rounded values that are stored, so the scalars
// describe exactly the fields the next kernel reads.

// K1: w = D p stored in half, gam = |w|^2; p (already rounded by SrcPUpdateH) stored at
// every point the block owns; clears the rescale flag of the previous reliable update
struct EpiK1PH {
  static constexpr bool kReduce = true;
  SpinH* w;
  SpinH* pnew;
  CGScalars* sc;
  // an iteration after one that asked for a reliable update waits for it (the single-system
  // solver runs the update once per chunk of iterations)
  __device__ bool skip() const { return sc->done != 0 || sc->need_ru != 0; }
  __device__ void begin() {}
  __device__ void operator()(size_t i, const Spin<float>& v, double2& acc) const {
    const SpinH h = to_half(v);
    sth_nt(w + i, h);
    acc.x += norm2(from_half(h));
  }
  __device__ void center(size_t i, const Spin<float>& v) const { sth_nt(pnew + i, to_half(v)); }
  __device__ void operator()(double2 t) const {
    sc->rescale = 0;
    k1_finish(sc, t.x);
  }
};

// K2: q = D^dag w; r -= alpha q (stored in half); |r|^2 in stored units, compared with the
// tolerance in true units.  x_inner (FP32) is updated once per group of kXGroup
// iterations: the iterations at positions 0 .. kXGroup-2 of a group leave alpha p pending
// (sc->alpha_pend[pos]; each p stays in its slot of the ring of search-direction buffers
// until K1 of the next group overwrites it), the last one adds sum_j alpha_pend[j] p_j +
// alpha p in one read-modify-write of x.  A reliable update or the end of the solve
// flushes the pending terms (k_flush_h).
constexpr int kXGroup = 4;
constexpr int kMaxPend = 8;  // entries of CGScalars::alpha_pend
template <int POS>
struct EpiK2H {
  static constexpr bool kReduce = true;
  static constexpr bool kUpdateX = POS == kXGroup - 1;
  Spin<float>* x;
  SpinH* r;
  const SpinH* ring[kXGroup];  // ring[POS] is this iteration's p
  CGScalars* sc;
  float alpha;
  float ap[kXGroup - 1];
  __device__ bool skip() const { return sc->done != 0 || sc->need_ru != 0; }
  __device__ void begin() {
    alpha = float(sc->alpha);
#pragma unroll
    for (int j = 0; j < kXGroup - 1; j++) ap[j] = kUpdateX ? float(sc->alpha_pend[j]) : 0.f;
  }
  // Positions 0..kXGroup-2 load r ahead of the stencil (2 registers held across the slice).
  // The x-updating position loads everything in the epilogue: holding r there too took the
  // kernel from 94 to 98 VGPRs and from 5 to 4 waves per SIMD, which with one block per
  // 256-site range and 64 slices leaves a second round of blocks (0.55 vs 0.37-0.39 ms).
  template <int P = POS, std::enable_if_t<P != kXGroup - 1, int> = 0>
  __device__ SpinH prefetch(size_t i) const {
    return r[i];
  }
  __device__ void operator()(size_t i, const Spin<float>& q, double2& acc) const {
    (*this)(i, q, acc, r[i]);
  }
  __device__ void operator()(size_t i, const Spin<float>& q, double2& acc, const SpinH& rpre) const {
    Spin<float> rr = from_half(rpre);
    if constexpr (kUpdateX) {
      Spin<float> xx = ld(x + i);
#pragma unroll
      for (int j = 0; j < kXGroup; j++) {
        const float a = j == kXGroup - 1 ? alpha : ap[j];
        if (a != 0.f) {  // uniform: a pending term flushed by a reliable update is 0
          const Spin<float> pp = ldh(ring[j] + i);
          xx.a.x += a * pp.a.x;
          xx.a.y += a * pp.a.y;
          xx.b.x += a * pp.b.x;
          xx.b.y += a * pp.b.y;
        }
      }
      st(x + i, xx);
    }
    rr.a.x -= alpha * q.a.x;
    rr.a.y -= alpha * q.a.y;
    rr.b.x -= alpha * q.b.x;
    rr.b.y -= alpha * q.b.y;
    const SpinH rh = to_half(rr);
    sth_nt(r + i, rh);
    acc.x += norm2(from_half(rh));
  }
  __device__ void operator()(double2 t) const {
    const double m = t.x;
    if constexpr (kUpdateX) {
      for (int j = 0; j < kXGroup - 1; j++) sc->alpha_pend[j] = 0.0;
    } else {
      sc->alpha_pend[POS] = sc->alpha;
    }
    sc->last_pos = POS;
    sc->iter += 1;
    if (!(m >= 0.0 && m < 1e300)) {
      sc->done = 2;
      return;
    }
    if (m < sc->delta2 * sc->maxr2 || m * sc->scale * sc->scale <= sc->tol2 * sc->bnorm2) {
      sc->need_ru = 1;
    } else {
      sc->beta = m / sc->mu;
      sc->mu = m;
      sc->maxr2 = fmax(sc->maxr2, m);
    }
    if (sc->iter >= sc->maxiter && !sc->need_ru) sc->done = 3;
  }
};

// U3: r64 = b - D^dag t in FP64, kept in FP32 in rt at the outgoing scale (so its values
// stay near unit size whatever the magnitude of b) for the rescale pass; |r64|^2 decides
// convergence, the new scale (unit rms per component) and beta, which also carries the
// change of scale of the stored p
struct EpiResidH {
  static constexpr bool kReduce = true;
  const Spin<double>* b;
  Spin<float>* rt;
  CGScalars* sc;
  double inv_ncomp;  // 1 / (4 * points)
  double inv_so;
  __device__ bool skip() const { return sc->done != 0 || sc->need_ru == 0; }
  __device__ void begin() { inv_so = 1.0 / sc->scale; }
  __device__ void operator()(size_t i, const Spin<double>& v, double2& acc) const {
    const Spin<double> bb = ld(b + i);
    Spin<double> rr;
    rr.a = {bb.a.x - v.a.x, bb.a.y - v.a.y};
    rr.b = {bb.b.x - v.b.x, bb.b.y - v.b.y};
    Spin<float> rf;
    rf.a = {float(rr.a.x * inv_so), float(rr.a.y * inv_so)};
    rf.b = {float(rr.b.x * inv_so), float(rr.b.y * inv_so)};
    st(rt + i, rf);
    acc.x += norm2(rr);
  }
  __device__ void operator()(double2 t) const {
    const double m = t.x;
    sc->need_ru = 0;
    sc->n_ru += 1;
    sc->true_r2 = m;
    for (int j = 0; j < kMaxPend; j++) sc->alpha_pend[j] = 0.0;  // flushed before this update
    if (!(m >= 0.0 && m < 1e300)) {
      sc->done = 2;
      return;
    }
    if (m <= sc->tol2 * sc->bnorm2) {
      sc->done = 1;
      return;
    }
    const double so = sc->scale, sn = sqrt(m * inv_ncomp);
    sc->beta = m / (so * so * sc->mu) * (so / sn);
    sc->mu = m / (sn * sn);
    sc->maxr2 = sc->mu;
    sc->scale = sn;
    sc->rfac = so / sn;
    sc->rescale = 1;
    if (sc->iter >= sc->maxiter) sc->done = 3;
  }
};

// ---- per-point and scalar steps of the half-storage solver, shared by the single-system
// and the batched kernels ----------------------------------------------------------------

// solver scalars at the start from |b|^2 (ncomp = 4 * points): the scale makes the stored
// residual unit-rms per component
__device__ __forceinline__ void half_start(CGScalars* sc, double b2, double ncomp, double tol2,
                                           double delta2, int maxiter) {
  const double s = b2 > 0.0 ? sqrt(b2 / ncomp) : 1.0;
  sc->scale = s;
  sc->mu = b2 / (s * s);
  sc->bnorm2 = b2;
  sc->maxr2 = sc->mu;
  sc->true_r2 = b2;
  sc->rfac = 1.0;
  for (int j = 0; j < kMaxPend; j++) sc->alpha_pend[j] = 0.0;
  sc->rescale = 0;
  sc->last_pos = kXGroup - 1;
  sc->gam = 0;
  sc->alpha = 0;
  sc->beta = 0;
  sc->tol2 = tol2;
  sc->delta2 = delta2;
  sc->iter = 0;
  sc->n_ru = 0;
  sc->maxiter = maxiter;
  sc->need_ru = 0;
  sc->done = (b2 == 0.0) ? 1 : (b2 < 1e300 ? 0 : 2);
}

// the pending directions of the current group (ring slots 0 .. kXGroup-2)
struct PendRing {
  const SpinH* p[kXGroup - 1];
};

// The inner fields of the single-system half-storage solver live in the temporal gauge:
// field' = exp(i phi) field at every point (rot = phi as 16-bit angles; null: no rotation).
__device__ __forceinline__ void rotate(Spin<float>& v, const phase16_t* rot, size_t i, bool inverse) {
  if (!rot) return;
  float2 u = ldu(rot + i);
  if (inverse) u.y = -u.y;
  v.a = cmul<float>(u, v.a);
  v.b = cmul<float>(u, v.b);
}

// x64 += scale exp(-i phi) (x_inner + sum_j alpha_pend[j] ring[j]) at point i, x_inner = 0
__device__ __forceinline__ void half_flush_point(Spin<double>* x64, Spin<float>* xin, const PendRing& ring,
                                                 const float* ap, double s, size_t i,
                                                 const phase16_t* rot = nullptr) {
  Spin<float> a = xin[i];
#pragma unroll
  for (int j = 0; j < kXGroup - 1; j++) {
    if (ap[j] != 0.f) {  // uniform
      const Spin<float> pp = ldh(ring.p[j] + i);
      a.a.x += ap[j] * pp.a.x;
      a.a.y += ap[j] * pp.a.y;
      a.b.x += ap[j] * pp.b.x;
      a.b.y += ap[j] * pp.b.y;
    }
  }
  rotate(a, rot, i, true);
  Spin<double> xx = x64[i];
  xx.a.x += s * a.a.x;
  xx.a.y += s * a.a.y;
  xx.b.x += s * a.b.x;
  xx.b.y += s * a.b.y;
  x64[i] = xx;
  Spin<float> z;
  z.a = {0.f, 0.f};
  z.b = {0.f, 0.f};
  xin[i] = z;
}

// r = exp(i phi) b / scale in binary16 at point i (FP64 scaling, so any finite b fits)
__device__ __forceinline__ void half_store_b(const Spin<double>* b, SpinH* r, double inv, size_t i,
                                             const phase16_t* rot = nullptr) {
  const Spin<double> bb = b[i];
  Spin<float> v;
  v.a = {float(bb.a.x * inv), float(bb.a.y * inv)};
  v.b = {float(bb.b.x * inv), float(bb.b.y * inv)};
  rotate(v, rot, i, false);
  r[i] = to_half(v);
}

// r = exp(i phi) rfac rt in binary16 at point i (after a reliable update)
__device__ __forceinline__ void half_rescale_point(const Spin<float>* rt, SpinH* r, float f, size_t i,
                                                   const phase16_t* rot = nullptr) {
  Spin<float> v = rt[i];
  v.a.x *= f;
  v.a.y *= f;
  v.b.x *= f;
  v.b.y *= f;
  rotate(v, rot, i, false);
  r[i] = to_half(v);
}

}  // namespace cgepi
}  // namespace qed3
