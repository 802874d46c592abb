// Matrix-free Wilson-Dirac stencil on S^2 x time for gfx950 (wave64).
//
// One thread owns one site x for a tile of `ts` consecutive time slices: the site's
// neighbour ids, link ids and 2x2 hop parameters are loaded once into registers and
// reused for every slice of the tile, and the temporal neighbours psi(s+-1,x) roll
// through registers.  Sites are numbered along a Hilbert curve, so the spatial gathers
// psi(s,y) of a 256-thread block mostly hit lines its own block just loaded.
//
// Operator (src/both_3d/dirac_ext.h:121-213):
//   (D psi)(s,x) = c_x psi(s,x) + sum_y H_xy e^{i theta_s(x,y)} psi(s,y)
//                + b_- 1/2 kt_x (-r - sigma3) e^{-i theta^t_{s-1}(x)} psi(s-1,x)
//                + b_+ 1/2 kt_x (-r + sigma3) e^{+i theta^t_s(x)}   psi(s+1,x)
// with b_- = bc at s = 0, b_+ = bc at s = Lt-1 (bc = -1 antiperiodic), 1 otherwise.
// D^dagger (the explicit conjugate transpose of sparse_dirac.h:132-143) is the same
// stencil with H_xy -> H_yx^dagger and the two temporal projectors swapped.
//
// The epilogue functor receives each output point and may fuse BLAS-1 work and a
// FP64 partial sum; kernels with a reducing epilogue store one partial per block, and
// launch_stencil follows them with a one-block finaliser kernel that sums the partials
// in block order and runs the epilogue's finaliser (no device-scope fences).
#pragma once
#include <type_traits>
#include <utility>

#include "device_common.hpp"
#include "dirac.hpp"

namespace qed3 {

// Block -> (site range, time tile).  Workgroups are dispatched round-robin over the 8
// XCDs; with `remap` the blocks landing on one XCD cover one contiguous eighth of the
// (Hilbert-ordered) site ranges for every time tile, so that XCD's L2 holds that eighth
// of the hop tables and the spatial halos stay inside it except at its two ends.
struct TileMap {
  int nxr;      // site ranges of blockDim.x sites
  int nxr_pad;  // nxr rounded up to a multiple of 8
  int nst;      // time tiles
  int remap;
};

__device__ __forceinline__ void tile_of(const TileMap& m, int b, int& xr, int& st) {
  if (m.remap) {
    const int per = m.nxr_pad >> 3;
    const int j = b >> 3;
    st = j / per;
    xr = (b & 7) * per + (j - st * per);
  } else {
    st = b / m.nxr_pad;
    xr = b - st * m.nxr_pad;
  }
}

// Input sources of the stencil.  bind() runs once per thread at kernel start and returns
// the per-point reader.  SrcPlain reads a field; SrcPUpdate forms the CG search direction
// p = r + beta p_old on the fly (beta from the solver scalars), so the p update needs no
// pass of its own.
template <typename T>
struct SrcPlain {
  const Spin<T>* __restrict__ p;
  __device__ SrcPlain bind() const { return *this; }
  __device__ Spin<T> operator()(size_t i) const { return ld(p + i); }
};

template <typename T>
struct SrcPUpdate {
  const Spin<T>* r;
  const Spin<T>* p;
  const CGScalars* sc;
  struct Bound {
    const Spin<T>* __restrict__ r;
    const Spin<T>* __restrict__ p;
    T beta;
    __device__ Spin<T> operator()(size_t i) const {
      Spin<T> a = ld(r + i);
      const Spin<T> c = ld(p + i);
      a.a.x += beta * c.a.x;
      a.a.y += beta * c.a.y;
      a.b.x += beta * c.b.x;
      a.b.y += beta * c.b.y;
      return a;
    }
  };
  __device__ Bound bind() const { return Bound{r, p, T(sc->beta)}; }
};

// The same two sources over half-storage fields (FP32 arithmetic).  SrcPUpdateH rounds
// the new p to binary16 before the stencil uses it, so D acts on exactly the p that is
// stored for the x update.
struct SrcHalf {
  const SpinH* __restrict__ p;
  __device__ SrcHalf bind() const { return *this; }
  __device__ Spin<float> operator()(size_t i) const { return ldh(p + i); }
};

struct SrcPUpdateH {
  const SpinH* r;
  const SpinH* p;
  const CGScalars* sc;
  struct Bound {
    const SpinH* __restrict__ r;
    const SpinH* __restrict__ p;
    float beta;
    __device__ Spin<float> operator()(size_t i) const {
      Spin<float> a = ldh(r + i);
      const Spin<float> c = ldh(p + i);
      a.a.x += beta * c.a.x;
      a.a.y += beta * c.a.y;
      a.b.x += beta * c.b.x;
      a.b.y += beta * c.b.y;
      return from_half(to_half(a));
    }
  };
  __device__ Bound bind() const { return Bound{r, p, float(sc->beta)}; }
};

// Phase of the temporal link (s, x) -> (s+1, x): from the table, or, for a view in the
// temporal gauge (TG, d.uPoly set), 1 except at the wrap (s = Lt-1).  TG is a template
// argument so that neither form carries the other's branch and loads inside the slice loop.
template <bool TG, typename T, typename P>
__device__ __forceinline__ cplx_t<T> temporal_phase(const DiracView<T, P>& d, int s, int x) {
  if constexpr (TG)
    return s == d.Lt - 1 ? ldu(d.uPoly + x) : cplx_t<T>{T(1), T(0)};
  else
    return ldu(d.uT + size_t(s) * d.V + x);
}

// Phase of the spatial link in slot j of site x at slice s (l = the slot's link id or ~id):
// in the temporal gauge from the view's per-site table (consecutive sites, consecutive
// entries; orientation applied), otherwise through the link table.
template <bool TG, int NS, typename T, typename P>
__device__ __forceinline__ cplx_t<T> slot_phase(const DiracView<T, P>& d, int s, int x, int j, int l) {
  if constexpr (TG) {
    return ldu(d.uSite + (size_t(s) * NS + j) * d.V + x);
  } else {
    cplx_t<T> u = ldu(d.uS + size_t(s) * d.E + (l >= 0 ? l : ~l));
    if (l < 0) u.y = -u.y;
    return u;
  }
}

// Calls body(std::bool_constant<TG>) with TG = whether the view is in the temporal gauge,
// decided once per block (a uniform branch on a kernel argument).  Only the 16-bit-angle
// views of the half-storage solver have a temporal-gauge form.
template <typename T, typename P, class Body>
__device__ __forceinline__ void with_gauge(const DiracView<T, P>& d, const Body& body) {
  if constexpr (std::is_same<P, phase16_t>::value) {
    if (d.uPoly)
      body(std::true_type{});
    else
      body(std::false_type{});
  } else {
    body(std::false_type{});
  }
}

// Epilogues may load what they need at an output point early through pre = prefetch(i),
// which the stencil issues with the slice's other loads (before the LDS barrier, so the
// latency overlaps the halo staging) and hands back as operator()(i, v, acc, pre).
template <class E, class = void>
struct EpiPre {
  struct type {};
  static constexpr bool value = false;
};
template <class E>
struct EpiPre<E, std::void_t<decltype(std::declval<const E&>().prefetch(size_t(0)))>> {
  using type = decltype(std::declval<const E&>().prefetch(size_t(0)));
  static constexpr bool value = true;
};
template <class Epi, typename T>
__device__ __forceinline__ void call_epi(const Epi& e, size_t i, const Spin<T>& v, double2& acc,
                                         const typename EpiPre<Epi>::type& pre) {
  if constexpr (EpiPre<Epi>::value)
    e(i, v, acc, pre);
  else
    e(i, v, acc);
}

// Epilogues may take the stencil's input value at each output point (its centre) through
// center(i, v); the others ignore it.
template <class Epi, typename T>
__device__ __forceinline__ auto call_center(const Epi& e, size_t i, const Spin<T>& v, int)
    -> decltype(e.center(i, v), void()) {
  e.center(i, v);
}
template <class Epi, typename T>
__device__ __forceinline__ void call_center(const Epi&, size_t, const Spin<T>&, long) {}

// The stencil for site x over the time slices [s0, s0+ts), handing every output point
// to the epilogue (shared by the single-system and the batched kernels).
template <typename T, int NS, bool DAG, bool TG, class Epi, class In, typename P>
__device__ __forceinline__ void stencil_site_src(const DiracView<T, P>& d, const In& in,
                                                 const Epi& epi, int x, int s0, int ts, double2& acc) {
  using C = cplx_t<T>;
  const int V = d.V, Lt = d.Lt;
  if (x < V && s0 < Lt) {
    int nb[NS], lk[NS];
    real4_t<T> hp[NS];
    const real4_t<T>* hop = DAG ? d.hopd : d.hop;
#pragma unroll
    for (int j = 0; j < NS; j++) {
      nb[j] = d.nbr[j * V + x];
      lk[j] = d.lnk[j * V + x];
      hp[j] = hop[j * V + x];
    }
    const T c = d.diag[x];
    const T kh = d.kth[x];
    const T m0 = kh * (-d.r - T(1)), m1 = kh * (-d.r + T(1));
    // T- = diag(m0, m1), T+ = diag(m1, m0); D uses (T- back, T+ fwd), D^dag swaps them.
    const T bk0 = DAG ? m1 : m0, bk1 = DAG ? m0 : m1;
    const T fw0 = DAG ? m0 : m1, fw1 = DAG ? m1 : m0;
    const int s1 = min(s0 + ts, Lt);
    const int sm = (s0 == 0) ? Lt - 1 : s0 - 1;
    Spin<T> prev = in(size_t(sm) * V + x);
    Spin<T> cur = in(size_t(s0) * V + x);
    C uprev = temporal_phase<TG>(d, sm, x);
    for (int s = s0; s < s1; s++) {
      const int sp = (s == Lt - 1) ? 0 : s + 1;
      const Spin<T> next = in(size_t(sp) * V + x);
      const C ucur = temporal_phase<TG>(d, s, x);
      typename EpiPre<Epi>::type pre{};
      if constexpr (EpiPre<Epi>::value) pre = epi.prefetch(size_t(s) * V + x);
      C o0 = cscale<T>(c, cur.a), o1 = cscale<T>(c, cur.b);
      const size_t sV = size_t(s) * V;
#pragma unroll
      for (int j = 0; j < NS; j++) {
        const C u = slot_phase<TG, NS>(d, s, x, j, lk[j]);
        const Spin<T> py = in(sV + nb[j]);
        const C p0 = cmul<T>(u, py.a), p1 = cmul<T>(u, py.b);
        const T Cc = hp[j].x, Ss = hp[j].y, A = hp[j].z, B = hp[j].w;
        // H = [[-r(C - iS), a - ib], [a + ib, -r(C + iS)]]
        const T rC = d.r * Cc, rS = d.r * Ss;
        o0.x += -rC * p0.x - rS * p0.y + A * p1.x + B * p1.y;
        o0.y += -rC * p0.y + rS * p0.x + A * p1.y - B * p1.x;
        o1.x += A * p0.x - B * p0.y - rC * p1.x + rS * p1.y;
        o1.y += A * p0.y + B * p0.x - rC * p1.y - rS * p1.x;
      }
      {  // backward temporal hop from s-1 with phase exp(-i theta^t_{s-1})
        const T sg = (s == 0) ? d.bc : T(1);
        const C q0 = cconjmul<T>(uprev, prev.a), q1 = cconjmul<T>(uprev, prev.b);
        o0.x += sg * bk0 * q0.x;
        o0.y += sg * bk0 * q0.y;
        o1.x += sg * bk1 * q1.x;
        o1.y += sg * bk1 * q1.y;
      }
      {  // forward temporal hop from s+1 with phase exp(+i theta^t_s)
        const T sg = (s == Lt - 1) ? d.bc : T(1);
        const C q0 = cmul<T>(ucur, next.a), q1 = cmul<T>(ucur, next.b);
        o0.x += sg * fw0 * q0.x;
        o0.y += sg * fw0 * q0.y;
        o1.x += sg * fw1 * q1.x;
        o1.y += sg * fw1 * q1.y;
      }
      Spin<T> o;
      o.a = o0;
      o.b = o1;
      call_epi(epi, size_t(s) * V + x, o, acc, pre);
      call_center(epi, size_t(s) * V + x, cur, 0);
      prev = cur;
      cur = next;
      uprev = ucur;
    }
  }
}

template <typename T, int NS, bool DAG, class Epi, class Src, typename P = cplx_t<T>>
__global__ __launch_bounds__(256) void k_stencil(DiracView<T, P> d, Src in, Epi epi, int ts, TileMap tm,
                                                  GridReduce gr) {
  if (epi.skip()) return;  // uniform: solver already finished
  int xr, st;
  tile_of(tm, blockIdx.x, xr, st);
  const int x = (xr < tm.nxr) ? xr * blockDim.x + threadIdx.x : d.V;
  double2 acc = {0.0, 0.0};
  epi.begin();
  const auto bound = in.bind();
  with_gauge(d, [&](auto tg) {
    stencil_site_src<T, NS, DAG, decltype(tg)::value>(d, bound, epi, x, st * ts, ts, acc);
  });
  if constexpr (Epi::kReduce) store_partial(acc, gr.partials, blockIdx.x);
}

// LDS-staged variant.  A block owns the site range [x0, x0+R) of one tile of time
// slices; per slice it stages the range's spinors plus its precomputed spatial halo
// (the outside neighbours of the range) in LDS with one coalesced load of the range
// and one gather of the halo, then reads all spatial neighbours from LDS through a
// 16-bit local index table.  Every spinor is fetched once per block and slice.
// The work of one block: site range xr (of nxr; ranges beyond are padding) and time tile
// st of ts slices, staging through the LDS tiles tileS (kLdsBuffers tiles of
// blockDim.x + hv.hmax spinors each).  Slices alternate between two tiles, so one barrier
// per slice suffices: slice s writes the tile slice s-2 read, and every thread finished
// slice s-2's reads before it reached slice s-1's barrier.
// Shared by the single-system kernel below and the batched solver.
constexpr int kLdsBuffers = 2;
template <typename T>
__host__ __device__ constexpr size_t lds_bytes(int block, int hmax) {
  return size_t(kLdsBuffers) * size_t(block + hmax) * sizeof(Spin<T>);
}

template <typename T, int NS, bool DAG, bool TG, class Epi, class In, typename P>
__device__ __forceinline__ void stencil_lds_block(const DiracView<T, P>& d, const In& in, const Epi& epi,
                                                  int ts, int xr, int nxr, int st, const HaloView& hv,
                                                  Spin<T>* tileS, double2& acc) {
  using C = cplx_t<T>;
  const int V = d.V, Lt = d.Lt, R = blockDim.x;
  const bool live = xr < nxr;
  const int x0 = xr * R;
  const int x = live ? x0 + threadIdx.x : V;
  const bool act = x < V;
  const int s0 = st * ts;
  const int s1 = min(s0 + ts, Lt);
  const int h0 = live ? hv.halo_off[xr] : 0;
  const int nh = live ? hv.halo_off[xr + 1] - h0 : 0;
  // The halo's site ids are the same for every slice: each thread keeps the first two of
  // its share in registers (entries threadIdx.x and threadIdx.x + R), so the per-slice
  // halo gathers do not wait on an index load; any further entries are read in the loop.
  const int hid0 = (int(threadIdx.x) < nh) ? hv.halo[h0 + threadIdx.x] : -1;
  const int hid1 = (int(threadIdx.x) + R < nh) ? hv.halo[h0 + threadIdx.x + R] : -1;
  int nb[NS], lk[NS];
  real4_t<T> hp[NS];
  T c = 0, bk0 = 0, bk1 = 0, fw0 = 0, fw1 = 0;
  Spin<T> prev, cur;
  C uprev;
  if (act && s0 < Lt) {
    const real4_t<T>* hop = DAG ? d.hopd : d.hop;
#pragma unroll
    for (int j = 0; j < NS; j++) {
      nb[j] = hv.nbl[j * V + x];
      lk[j] = d.lnk[j * V + x];
      hp[j] = hop[j * V + x];
    }
    c = d.diag[x];
    const T kh = d.kth[x];
    const T m0 = kh * (-d.r - T(1)), m1 = kh * (-d.r + T(1));
    bk0 = DAG ? m1 : m0;
    bk1 = DAG ? m0 : m1;
    fw0 = DAG ? m0 : m1;
    fw1 = DAG ? m1 : m0;
    const int sm = (s0 == 0) ? Lt - 1 : s0 - 1;
    prev = in(size_t(sm) * V + x);
    cur = in(size_t(s0) * V + x);
    uprev = temporal_phase<TG>(d, sm, x);
  }
  for (int s = s0; s < s1; s++) {
    const size_t sV = size_t(s) * V;
    // This slice's loads that do not go through LDS (the forward temporal neighbour and
    // the link phases) are issued before the barriers, so they are in flight together with
    // the halo loads instead of after them.
    Spin<T> next;
    C ucur;
    C uu[NS];
    typename EpiPre<Epi>::type pre{};
    if (act) {
      const int sp = (s == Lt - 1) ? 0 : s + 1;
      next = in(size_t(sp) * V + x);
      if constexpr (EpiPre<Epi>::value) pre = epi.prefetch(sV + x);
      ucur = temporal_phase<TG>(d, s, x);
#pragma unroll
      for (int j = 0; j < NS; j++) uu[j] = slot_phase<TG, NS>(d, s, x, j, lk[j]);
    }
    Spin<T>* tb = tileS + (s & 1) * (R + hv.hmax);
    if (act) tb[threadIdx.x] = cur;
    if (hid0 >= 0) tb[R + threadIdx.x] = in(sV + hid0);
    if (hid1 >= 0) tb[2 * R + threadIdx.x] = in(sV + hid1);
    for (int k = threadIdx.x + 2 * R; k < nh; k += R) tb[R + k] = in(sV + hv.halo[h0 + k]);
    __syncthreads();
    if (act) {
      C o0 = cscale<T>(c, cur.a), o1 = cscale<T>(c, cur.b);
#pragma unroll
      for (int j = 0; j < NS; j++) {
        const C u = uu[j];
        const Spin<T> py = tb[nb[j]];
        const C p0 = cmul<T>(u, py.a), p1 = cmul<T>(u, py.b);
        const T Cc = hp[j].x, Ss = hp[j].y, A = hp[j].z, B = hp[j].w;
        const T rC = d.r * Cc, rS = d.r * Ss;
        o0.x += -rC * p0.x - rS * p0.y + A * p1.x + B * p1.y;
        o0.y += -rC * p0.y + rS * p0.x + A * p1.y - B * p1.x;
        o1.x += A * p0.x - B * p0.y - rC * p1.x + rS * p1.y;
        o1.y += A * p0.y + B * p0.x - rC * p1.y - rS * p1.x;
      }
      {
        const T sg = (s == 0) ? d.bc : T(1);
        const C q0 = cconjmul<T>(uprev, prev.a), q1 = cconjmul<T>(uprev, prev.b);
        o0.x += sg * bk0 * q0.x;
        o0.y += sg * bk0 * q0.y;
        o1.x += sg * bk1 * q1.x;
        o1.y += sg * bk1 * q1.y;
      }
      {
        const T sg = (s == Lt - 1) ? d.bc : T(1);
        const C q0 = cmul<T>(ucur, next.a), q1 = cmul<T>(ucur, next.b);
        o0.x += sg * fw0 * q0.x;
        o0.y += sg * fw0 * q0.y;
        o1.x += sg * fw1 * q1.x;
        o1.y += sg * fw1 * q1.y;
      }
      Spin<T> o;
      o.a = o0;
      o.b = o1;
      call_epi(epi, size_t(s) * V + x, o, acc, pre);
      call_center(epi, size_t(s) * V + x, cur, 0);
      prev = cur;
      cur = next;
      uprev = ucur;
    }
  }
  __syncthreads();  // every read of the tiles is done: the caller may stage its next block
}

template <typename T, int NS, bool DAG, class Epi, class Src, typename P = cplx_t<T>>
__global__ __launch_bounds__(256) void k_stencil_lds(DiracView<T, P> d, Src src, Epi epi, int ts, TileMap tm,
                                                      HaloView hv, GridReduce gr) {
  if (epi.skip()) return;
  const auto in = src.bind();
  extern __shared__ unsigned char smem_raw[];
  Spin<T>* tileS = reinterpret_cast<Spin<T>*>(smem_raw);
  int xr, st;
  tile_of(tm, blockIdx.x, xr, st);
  double2 acc = {0.0, 0.0};
  epi.begin();
  with_gauge(d, [&](auto tg) {
    stencil_lds_block<T, NS, DAG, decltype(tg)::value>(d, in, epi, ts, xr, tm.nxr, st, hv, tileS, acc);
  });
  if constexpr (Epi::kReduce) store_partial(acc, gr.partials, blockIdx.x);
}

// Epilogue: plain store.
template <typename T>
struct EpiStore {
  static constexpr bool kReduce = false;
  Spin<T>* out;
  __device__ bool skip() const { return false; }
  __device__ void begin() {}
  __device__ void operator()(size_t i, const Spin<T>& v, double2&) const { st(out + i, v); }
  __device__ void operator()(double2) const {}
};

// Epilogue: store in binary16 (half-storage fields)
struct EpiStoreH {
  static constexpr bool kReduce = false;
  SpinH* out;
  __device__ bool skip() const { return false; }
  __device__ void begin() {}
  __device__ void operator()(size_t i, const Spin<float>& v, double2&) const { out[i] = to_half(v); }
  __device__ void operator()(double2) const {}
};

template <typename T>
__device__ __forceinline__ double norm2(const Spin<T>& v) {
  return double(v.a.x) * v.a.x + double(v.a.y) * v.a.y + double(v.b.x) * v.b.x +
         double(v.b.y) * v.b.y;
}

// `bracket` (optional, a pair of events) is recorded just before and just after the
// stencil kernel itself (not its reduction finaliser), for per-kernel timing.
template <typename T, bool DAG, class Epi, class Src, typename P>
inline void launch_stencil_src(const DiracView<T, P>& v, int max_nn, const Src& in, const Epi& epi,
                               StencilTile t, GridReduce gr, hipStream_t st, const HaloView& hv,
                               const hipEvent_t* bracket = nullptr) {
  TileMap tm;
  tm.nxr = (v.V + t.block - 1) / t.block;
  tm.nxr_pad = (tm.nxr + 7) / 8 * 8;
  tm.nst = (v.Lt + t.ts - 1) / t.ts;
  tm.remap = t.remap;
  const size_t nblk = size_t(tm.nxr_pad) * tm.nst;
  if (Epi::kReduce && nblk > kReduceCapacity)
    throw std::invalid_argument("stencil grid exceeds the reduction scratch");
  const dim3 grid{static_cast<unsigned>(nblk), 1u, 1u};
  if (bracket) QED3_HIP(hipEventRecord(bracket[0], st));
  if (t.lds) {
    if (!hv.nbl) throw std::logic_error("LDS stencil launched without halo tables");
    const size_t smem = lds_bytes<T>(t.block, hv.hmax);
    if (max_nn == 6)
      hipLaunchKernelGGL((k_stencil_lds<T, 6, DAG, Epi, Src, P>), grid, dim3(t.block), smem, st, v, in, epi,
                         t.ts, tm, hv, gr);
    else
      hipLaunchKernelGGL((k_stencil_lds<T, 3, DAG, Epi, Src, P>), grid, dim3(t.block), smem, st, v, in, epi,
                         t.ts, tm, hv, gr);
  } else if (max_nn == 6) {
    hipLaunchKernelGGL((k_stencil<T, 6, DAG, Epi, Src, P>), grid, dim3(t.block), 0, st, v, in, epi, t.ts, tm, gr);
  } else {
    hipLaunchKernelGGL((k_stencil<T, 3, DAG, Epi, Src, P>), grid, dim3(t.block), 0, st, v, in, epi, t.ts, tm, gr);
  }
  if (bracket) QED3_HIP(hipEventRecord(bracket[1], st));
  if constexpr (Epi::kReduce)
    hipLaunchKernelGGL((k_finalize<Epi>), dim3(1), dim3(256), 0, st, gr.partials, unsigned(nblk), epi);
}

template <typename T, bool DAG, class Epi, typename P>
inline void launch_stencil(const DiracView<T, P>& v, int max_nn, const Spin<T>* in, const Epi& epi,
                           StencilTile t, GridReduce gr, hipStream_t st, const HaloView& hv) {
  launch_stencil_src<T, DAG>(v, max_nn, SrcPlain<T>{in}, epi, t, gr, st, hv);
}

}  // namespace qed3
