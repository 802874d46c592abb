// Epilogues of the device-driven CG kernels (shared by the single-system solver and the
// batched solver).  Each reads/writes the solver scalars of its own system.
#pragma once
#include "dirac.hpp"
#include "stencil.hpp"

namespace qed3 {
namespace cgepi {

__device__ __forceinline__ bool finite_pos(double v) { return v > 0.0 && v < 1e300; }

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
  __device__ void operator()(double2 t) const {
    sc->gam = t.x;
    if (!finite_pos(t.x)) {
      sc->done = 2;
      return;
    }
    sc->alpha = sc->mu / t.x;
  }
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

}  // namespace cgepi
}  // namespace qed3
