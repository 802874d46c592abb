// Device side of the time-exact Horner preconditioner (horner.hpp): one layer
// g_out = T^{-1}(c v + S_th g_in) (or its adjoint form) per launch.
//
// Thread = site x; a block owns a range of R Hilbert-consecutive sites.  Under T the spin-up
// row's recurrence runs forward in time and the spin-down row's backward (the other way
// round under T^dagger).  A layer makes two passes over each site's time column:
//   pass 1, s = 0 .. Lt-1: per slice the block stages g_in of its range and spatial halo in
//     LDS (two alternating tiles, one barrier per slice), every thread forms the spatial
//     transports of both rows from the same staged values and link phases, finishes the
//     forward row's recurrence state = (y + kappa_t u^t state) / c_x and stores, at its
//     output point, that row's result beside the other row's y (its recurrence input);
//   pass 2, s = Lt-1 .. 0: every thread reads its own column back, runs the backward row's
//     recurrence and stores both rows' results (no neighbours, no barriers).
// So every neighbour, link phase and input load serves both rows.  Fields are FP32
// spin-interleaved points (the checked apply and the training gradient, in any gauge); the
// solver's binary16 layers are the parallel-in-time form further down.
// The first layer of a network (FIRST) has no input field: no staging, no barrier.
#pragma once
#include <type_traits>

#include "horner.hpp"
#include "stencil.hpp"

namespace qed3 {
namespace hk {

// binary16 complex arithmetic on (re, im) pairs (packed FMAs)
__device__ __forceinline__ half2_t cmul16(half2_t x, half2_t y) {
  const half2_t ys = {-y.y, y.x};
  return x.xx * y + x.yy * ys;
}

// element idx of base through a 32-bit byte offset (idx * sizeof(T) < 2^32), so that the
// access takes a uniform base from scalar registers and a 32-bit vector offset
template <class T>
__device__ __forceinline__ const T& at32(const T* base, unsigned idx) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + idx * unsigned(sizeof(T)));
}
template <class T>
__device__ __forceinline__ T* ptr32(T* base, unsigned idx) {
  return reinterpret_cast<T*>(reinterpret_cast<char*>(base) + idx * unsigned(sizeof(T)));
}
template <class T>
__device__ __forceinline__ const T* cptr32(const T* base, unsigned idx) {
  return reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + idx * unsigned(sizeof(T)));
}

// Loads through a buffer resource (base, nbytes): byte offset voff (per thread) + soff (a
// scalar register: uniform across the wave); loads beyond nbytes return zeros.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, unsigned nbytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), short(0), int(nbytes), 0x00020000);
}
__device__ __forceinline__ SpinH bload_h(const void* base, unsigned nbytes, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(SpinH, __builtin_amdgcn_raw_buffer_load_b64(buf_rsrc(base, nbytes), voff, soff, 0));
}
__device__ __forceinline__ unsigned bload_w(const void* base, unsigned nbytes, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(unsigned, __builtin_amdgcn_raw_buffer_load_b32(buf_rsrc(base, nbytes), voff, soff, 0));
}
// e^{i theta} of the 10-bit angle in bits 10k..10k+9 of a packed word (theta = 2 pi a / 1024)
// as a binary16 pair; the hardware sine and cosine take revolutions
__device__ __forceinline__ half2_t angle_pair10(unsigned w, int k) {
  const float rev = float((w >> (10 * k)) & 1023u) * (1.0f / 1024.0f);
  return half2_t{(_Float16)__builtin_amdgcn_cosf(rev), (_Float16)__builtin_amdgcn_sinf(rev)};
}

// component a of a point as FP32 complex
__device__ __forceinline__ float2 comp(const Spin<float>& v, int a) { return a == 0 ? v.a : v.b; }
__device__ __forceinline__ float2 comp(const SpinH& v, int a) {
  return a == 0 ? float2{float(v.v.x), float(v.v.y)} : float2{float(v.v.z), float(v.v.w)};
}
__device__ __forceinline__ half2_t comp16(const SpinH& v, int a) {
  return a == 0 ? half2_t{v.v.x, v.v.y} : half2_t{v.v.z, v.v.w};
}
__device__ __forceinline__ Spin<float> to_f(const Spin<float>& v) { return v; }
__device__ __forceinline__ Spin<float> to_f(const SpinH& v) { return from_half(v); }
// a point from its two components, stored (binary16 stores are non-temporal: written once,
// read by the next kernel)
__device__ __forceinline__ void put(Spin<float>* p, float2 c0, float2 c1) {
  Spin<float> v;
  v.a = c0;
  v.b = c1;
  *p = v;
}
__device__ __forceinline__ void put(SpinH* p, float2 c0, float2 c1) {
  Spin<float> v;
  v.a = c0;
  v.b = c1;
  sth_nt(p, to_half(v));
}
// the value a load of what put() stored would return
__device__ __forceinline__ Spin<float> stored(const Spin<float>*, float2 c0, float2 c1) {
  Spin<float> v;
  v.a = c0;
  v.b = c1;
  return v;
}
__device__ __forceinline__ Spin<float> stored(const SpinH*, float2 c0, float2 c1) {
  Spin<float> v;
  v.a = c0;
  v.b = c1;
  return from_half(to_half(v));
}

// Injection sources: the point i as stored.
template <typename F>
struct SrcF {
  const F* p;
  __device__ SrcF bind() const { return *this; }
  __device__ F raw(size_t i) const { return p[i]; }
  __device__ F bat(unsigned nbytes, unsigned voff, unsigned soff) const { return bload_h(p, nbytes, voff, soff); }
  __device__ float2 scale(float2 c) const { return c; }
};
// binary16 field whose stored values still owe the factor *f (parallel-in-time layer only):
// the layer folds it into its injection coefficient
struct SrcFS {
  const SpinH* p;
  const double* f;
  struct Bound {
    const SpinH* p;
    float f;
    __device__ SpinH bat(unsigned nbytes, unsigned voff, unsigned soff) const { return bload_h(p, nbytes, voff, soff); }
    __device__ float2 scale(float2 c) const { return {c.x * f, c.y * f}; }
  };
  __device__ Bound bind() const { return {p, float(*f)}; }
};

// Injection with a side field (parallel-in-time layer only): the layer also accumulates
// |v|^2 and Re<o, v> over its points (v the injection as stored, o the side field) into the
// reduction of its epilogue, which must have kReduce.
struct SrcDots {
  static constexpr bool kDots = true;
  const SpinH* p;
  const SpinH* o;
  __device__ SrcDots bind() const { return *this; }
  __device__ SpinH raw(size_t i) const { return p[i]; }
  __device__ SpinH bat(unsigned nbytes, unsigned voff, unsigned soff) const { return bload_h(p, nbytes, voff, soff); }
  __device__ SpinH bside(unsigned nbytes, unsigned voff, unsigned soff) const {
    return bload_h(o, nbytes, voff, soff);
  }
  __device__ float2 scale(float2 c) const { return c; }
};
template <class S, class = void>
struct SrcHasDots : std::false_type {};
template <class S>
struct SrcHasDots<S, std::void_t<decltype(S::kDots)>> : std::bool_constant<S::kDots> {};

struct EpiNone {
  static constexpr bool kReduce = false;
  __device__ bool skip() const { return false; }
  __device__ void operator()(size_t, const Spin<float>&, double2&) const {}
  __device__ void operator()(double2) const {}
};

// LDS tile entry of the two-pass layer: an FP32 point
__device__ __forceinline__ float4 to_tile(const Spin<float>& v) { return {v.a.x, v.a.y, v.b.x, v.b.y}; }

template <int NS, typename F, typename P>
struct SliceLoads {
  F v;
  P u[NS];
  P ut;
  F gc, gh0, gh1;
};

template <int NS, bool ADJ, bool FIRST, bool TG, typename F, class Src, class Epi, typename P>
__global__ __launch_bounds__(256) void k_horner_layer(HornerView hv, DiracView<float, P> d, HaloView halo, Src src,
                                                      const F* __restrict__ gin, F* __restrict__ gout, float2 cj,
                                                      TileMap tm, Epi epi, GridReduce gr) {
  static_assert(std::is_same<F, Spin<float>>::value, "the two-pass layer runs on FP32 fields");
  using Tile = float4;
  if (epi.skip()) return;  // uniform
  extern __shared__ float4 hk_tile_raw[];
  Tile* hk_tile = reinterpret_cast<Tile*>(hk_tile_raw);
  int xr, st;
  tile_of(tm, blockIdx.x, xr, st);
  (void)st;
  const int V = hv.V, Lt = hv.Lt, R = blockDim.x;
  const bool live = xr < tm.nxr;
  const int x0 = xr * R;
  const int x = live ? x0 + int(threadIdx.x) : V;
  const bool act = x < V;
  constexpr int rF = ADJ ? 1 : 0, rB = 1 - rF;  // rows whose recurrence runs forward / backward
  const int h0 = (live && !FIRST) ? halo.halo_off[xr] : 0;
  const int nh = (live && !FIRST) ? halo.halo_off[xr + 1] - h0 : 0;
  const int hid0 = (int(threadIdx.x) < nh) ? halo.halo[h0 + threadIdx.x] : -1;
  const int hid1 = (int(threadIdx.x) + R < nh) ? halo.halo[h0 + threadIdx.x + R] : -1;
  int nb[NS], lk[NS];
  float4 hrF[NS], hrB[NS];
  float invc = 0.f, kt = 0.f;
  if (act) {
    if constexpr (!FIRST) {
      const size_t mF = size_t((ADJ ? 2 : 0) + rF) * NS * V, mB = size_t((ADJ ? 2 : 0) + rB) * NS * V;
#pragma unroll
      for (int j = 0; j < NS; j++) {
        nb[j] = halo.nbl[j * V + x];
        lk[j] = d.lnk[j * V + x];
        hrF[j] = hv.hrow[mF + size_t(j) * V + x];
        hrB[j] = hv.hrow[mB + size_t(j) * V + x];
      }
    }
    invc = hv.invc[x];
    kt = hv.kt[x];
  }
  const auto in = src.bind();
  // Loads of slice s: the injection, (original gauge) the temporal link s-1 -> s that the
  // forward recurrence crosses, the link phases and the input field at the site and at its
  // share of the halo.
  auto load = [&](int s, SliceLoads<NS, F, P>& L) {
    const size_t sV = size_t(s) * V;
    if (act) {
      L.v = in.raw(sV + x);
      if constexpr (!TG) {
        if (s > 0) L.ut = d.uT[size_t(s - 1) * V + x];
      }
      if constexpr (!FIRST) {
        const P* uS = d.uS + size_t(s) * d.E;
#pragma unroll
        for (int j = 0; j < NS; j++) L.u[j] = uS[lk[j] >= 0 ? lk[j] : ~lk[j]];
        L.gc = gin[sV + x];
      }
    }
    if constexpr (!FIRST) {
      if (hid0 >= 0) L.gh0 = gin[sV + hid0];
      if (hid1 >= 0) L.gh1 = gin[sV + hid1];
    }
  };
  // sum_j u_j (m_same g_a + m_flip g_abar)(nb_j) for row a from the staged tile
  auto transport = [&](const SliceLoads<NS, F, P>& L, const Tile* tb, const float4* hr, int a) -> float2 {
    float2 t = {0.f, 0.f};
#pragma unroll
    for (int j = 0; j < NS; j++) {
      float2 u = ldu(&L.u[j]);
      if (lk[j] < 0) u.y = -u.y;
      const float4 g = tb[nb[j]];
      const float2 ga = a == 0 ? float2{g.x, g.y} : float2{g.z, g.w};
      const float2 gb = a == 0 ? float2{g.z, g.w} : float2{g.x, g.y};
      const float2 cm = cadd<float>(cmul<float>({hr[j].x, hr[j].y}, ga), cmul<float>({hr[j].z, hr[j].w}, gb));
      t = cadd<float>(t, cmul<float>(u, cm));
    }
    return t;
  };
  float2 stF = {0.f, 0.f}, stB = {0.f, 0.f};
  // S holds the time wrap (the link from Lt-1 to 0, which T leaves out): each recurrence
  // starts from bc kappa_t w g(the far end of its walk), w the wrap link's phase,
  // conjugated for the forward walk.
  if constexpr (!FIRST) {
    if (act) {
      const cplx_t<float> w = TG ? ldu(d.uPoly + x) : ldu(d.uT + size_t(Lt - 1) * V + x);
      const float2 gF = comp(gin[size_t(Lt - 1) * V + x], rF), gB = comp(gin[x], rB);
      const float2 tF = cmul<float>({w.x, -w.y}, gF), tB = cmul<float>(w, gB);
      stF = {d.bc * kt * tF.x, d.bc * kt * tF.y};
      stB = {d.bc * kt * tB.x, d.bc * kt * tB.y};
    }
  }
  // ---- pass 1: forward in time --------------------------------------------------------
  SliceLoads<NS, F, P> cur, nxt;
  load(0, cur);
  for (int s = 0; s < Lt; s++) {
    if (s + 1 < Lt) load(s + 1, nxt);  // a slice ahead: in flight during this slice's work
    Tile* tb = hk_tile + (s & 1) * (R + halo.hmax);
    if constexpr (!FIRST) {
      if (act) tb[threadIdx.x] = to_tile(cur.gc);
      if (hid0 >= 0) tb[R + threadIdx.x] = to_tile(cur.gh0);
      if (hid1 >= 0) tb[2 * R + threadIdx.x] = to_tile(cur.gh1);
      for (int q = int(threadIdx.x) + 2 * R; q < nh; q += R)
        tb[R + q] = to_tile(gin[size_t(s) * V + halo.halo[h0 + q]]);
      __syncthreads();
    }
    if (act) {
      float2 yF = cmul<float>(cj, comp(cur.v, rF)), yB = cmul<float>(cj, comp(cur.v, rB));
      if constexpr (!FIRST) {
        const float2 tF = transport(cur, tb, hrF, rF), tB = transport(cur, tb, hrB, rB);
        yF.x -= tF.x;
        yF.y -= tF.y;
        yB.x -= tB.x;
        yB.y -= tB.y;
      }
      if (s > 0) {
        float2 t = stF;
        if constexpr (!TG) {
          const float2 ut = ldu(&cur.ut);
          t = cmul<float>({ut.x, -ut.y}, stF);  // the link s-1 -> s, conjugated
        }
        yF.x += kt * t.x;
        yF.y += kt * t.y;
      } else {
        yF.x += stF.x;  // the wrap term, already weighted (0 in a first layer)
        yF.y += stF.y;
      }
      stF = {invc * yF.x, invc * yF.y};
      const size_t i = size_t(s) * V + x;
      if constexpr (rF == 0)
        put(gout + i, stF, yB);
      else
        put(gout + i, yB, stF);
    }
    cur = nxt;
  }
  // ---- pass 2: backward in time over the thread's own column ------------------------
  double2 acc = {0.0, 0.0};
  if (act) {
    F tcur = gout[size_t(Lt - 1) * V + x], tnxt = tcur;
    P ucur{};
    for (int s = Lt - 1; s >= 0; s--) {
      const size_t i = size_t(s) * V + x;
      if (s > 0) tnxt = gout[i - V];
      if constexpr (!TG) {
        if (s < Lt - 1) ucur = d.uT[i];  // the link s -> s+1 the backward recurrence crosses
      }
      const float2 gF = comp(tcur, rF);
      float2 yB = comp(tcur, rB);
      if (s < Lt - 1) {
        float2 t = stB;
        if constexpr (!TG) t = cmul<float>(ldu(&ucur), stB);
        yB.x += kt * t.x;
        yB.y += kt * t.y;
      } else {
        yB.x += stB.x;  // the wrap term
        yB.y += stB.y;
      }
      stB = {invc * yB.x, invc * yB.y};
      if constexpr (rF == 0) {
        put(gout + i, gF, stB);
        epi(i, stored(gout, gF, stB), acc);
      } else {
        put(gout + i, stB, gF);
        epi(i, stored(gout, stB, gF), acc);
      }
      tcur = tnxt;
    }
  }
  if constexpr (Epi::kReduce) store_partial(acc, gr.partials, blockIdx.x);
}

// Launch one layer over the lattice.  `tables` supplies the spatial link phases (and, without
// TG, the temporal ones) of the gauge the fields live in.
template <bool ADJ, bool FIRST, bool TG, typename F, class Src, class Epi, typename P>
inline void launch_horner_layer(const HornerView& hv, const DiracView<float, P>& tables, const HaloView& halo,
                                const Src& src, const F* gin, F* gout, float2 cj, const Epi& epi, GridReduce gr,
                                hipStream_t st, const hipEvent_t* bracket = nullptr) {
  constexpr int R = 256;
  TileMap tm;
  tm.nxr = (hv.V + R - 1) / R;
  tm.nxr_pad = (tm.nxr + 7) / 8 * 8;
  tm.nst = 1;
  tm.remap = 1;
  const unsigned nblk = unsigned(tm.nxr_pad);
  if (Epi::kReduce && nblk > kReduceCapacity) throw std::invalid_argument("layer grid exceeds the reduction scratch");
  const size_t smem = FIRST ? 0 : size_t(2) * (R + halo.hmax) * sizeof(float4);
  if (bracket) QED3_HIP(hipEventRecord(bracket[0], st));
  if (hv.nslot == 6)
    hipLaunchKernelGGL((k_horner_layer<6, ADJ, FIRST, TG, F, Src, Epi, P>), dim3(nblk), dim3(R), smem, st, hv,
                       tables, halo, src, gin, gout, cj, tm, epi, gr);
  else
    hipLaunchKernelGGL((k_horner_layer<3, ADJ, FIRST, TG, F, Src, Epi, P>), dim3(nblk), dim3(R), smem, st, hv,
                       tables, halo, src, gin, gout, cj, tm, epi, gr);
  if (bracket) QED3_HIP(hipEventRecord(bracket[1], st));
  if constexpr (Epi::kReduce)
    hipLaunchKernelGGL((k_finalize<Epi>), dim3(1), dim3(256), 0, st, gr.partials, nblk, epi);
}

// ---- parallel-in-time layer (the solver's path: binary16 fields, temporal gauge) ------
// In the temporal gauge every temporal link inside a column is 1, so both recurrences have
// the constant real coefficient a = kappa_t / c_x:
//   forward  st(s) = a st(s-1) + y(s) / c_x,   backward  st(s) = a st(s+1) + y(s) / c_x,
// with the wrap term (bc kappa_t w g_in at the column's far end) added to y at the first
// (forward) or last (backward) slice.  A linear recurrence splits over time chunks: lane k
// of a site owns slices [kT, kT + T) and runs both recurrences over its chunk from a zero
// start, holding its chunk in registers; the chunks' end values are exchanged in LDS and
// every lane adds carry a^m to its values (m the distance from its chunk's start, for the
// backward row from its end), with carry_k = E_{k-1} + a^T carry_{k-1} (forward) and
// carry_k = E_{k+1} + a^{len_{k+1}} carry_{k+1} (backward), E the lanes' end values.  So a
// layer reads its input and injection once and writes its output once, and a site's time
// column runs on C = ceil(Lt / T) threads (C * S <= kTcBlock, S sites per block; the threads
// of a lane are consecutive, so a wave reads and writes consecutive sites of one slice).  Per
// step the block stages the C slices its lanes are at, for its S sites and their halo (halo
// tables of ranges of S sites).  The spatial link phases come per site as the view's uSite10:
// the slots' angles (orientation applied) at 10 bits, three to a 32-bit word, so a thread
// reads one word per slice on the trivalent lattice (a wave's loads consecutive), turned into
// binary16 (cos, sin) pairs with the hardware sine and cosine.  The rounding (at most
// pi/1024 rad) is of the order of the binary16 arithmetic's and only enters the
// preconditioner.
constexpr int kTcChunk = 16;   // slices per lane
constexpr int kTcDepth = 2;    // steps whose loads are in flight ahead of the one computed
constexpr int kTcBlock = 256;  // most threads per block (S sites x C lanes)
constexpr int kTcSites = 256;  // most sites per block

// LDS written by threads of a wave and read back by other threads of the same wave: the wave
// executes its LDS instructions in order, so only the compiler must keep them in order
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// FUSE (K^dagger's last layer only): the kernel then also applies the first layer of K to
// the output it has just stored, from the same registers: gout2 = T^{-1}(cj2 gout) with
// gout as stored (binary16), which saves K's first launch and its read of gout.
// WS (S = the wave size): a lane is one wave, which stages its tile (its sites and their
// whole halo) and alone reads it, so the per-step staging needs no block barrier.
template <int NS, bool ADJ, bool FIRST, int T, class Src, class Epi, bool FUSE = false, bool UL = false,
          bool H2 = false, bool WS = false>
__global__ __launch_bounds__(kTcBlock, (NS == 3 ? 4 : 2)) void k_horner_layer_tc(HornerView hv, DiracView<float, phasepair_t> d,
                                                         HaloView halo, Src src, const SpinH* __restrict__ gin,
                                                         SpinH* __restrict__ gout, float2 cj, TileMap tm, Epi epi,
                                                         GridReduce gr, int C, SpinH* __restrict__ gout2,
                                                         float2 cj2) {
  static_assert(!FUSE || ADJ, "K's first layer fuses into a layer of K^dagger");
  static_assert(!WS || UL, "wave-synchronous staging needs one wave per lane");
  constexpr bool kDots = SrcHasDots<Src>::value;
  static_assert(!kDots || Epi::kReduce, "the injection's products need the epilogue's reduction");
  if (epi.skip()) return;  // uniform
  extern __shared__ float4 hk_tile_raw[];
  const int S = int(blockDim.x) / C;
  // a wave: consecutive sites of one lane (UL: S a multiple of the wave size, so the lane, its
  // slices and their offsets are wave-uniform and live in scalar registers)
  int lane = int(threadIdx.x) / S;
  if constexpr (UL) lane = __builtin_amdgcn_readfirstlane(lane);
  const int sl = int(threadIdx.x) - lane * S;
  SpinH* tile = reinterpret_cast<SpinH*>(hk_tile_raw);         // [2][C][S + hmax]
  const int row = S + halo.hmax;
  float2* carry = reinterpret_cast<float2*>(tile + (FIRST ? 0 : 2 * C * row));  // [2][kTcBlock]
  HopRow16* hqs = reinterpret_cast<HopRow16*>(carry + 2 * kTcBlock);  // [S][2 rows][NS], the sites' hop rows
  int xr, st_;
  tile_of(tm, blockIdx.x, xr, st_);
  (void)st_;
  const int V = hv.V, Lt = hv.Lt;
  const bool live = xr < tm.nxr;
  const int x0 = xr * S;
  const int x = live ? x0 + sl : V;
  const bool act = x < V;
  const unsigned xc = act ? unsigned(x) : 0u;  // idle threads load site 0 and discard it
  const int s0 = lane * T;
  const int lenU = max(0, min(T, Lt - s0));  // slices of this lane
  const int len = act ? lenU : 0;            // ... that this thread stores
  // slices this thread computes: UL computes them on idle threads too (discarded), so that the
  // step loop branches on the wave-uniform count
  const int lenC = UL ? lenU : len;
  constexpr int rF = ADJ ? 1 : 0, rB = 1 - rF;
  // halo of this block's range: lane c stages slice cT + j of it, its S threads share it
  const int h0 = (live && !FIRST) ? halo.halo_off[xr] : 0;
  const int nh = (live && !FIRST) ? halo.halo_off[xr + 1] - h0 : 0;
  const int hid0 = (sl < nh) ? halo.halo[h0 + sl] : -1;
  const int hid1 = (sl + S < nh) ? halo.halo[h0 + sl + S] : -1;
  const unsigned hc0 = hid0 >= 0 ? unsigned(hid0) : xc, hc1 = hid1 >= 0 ? unsigned(hid1) : xc;
  // neighbour slots in the tile
  int nb[NS];
#pragma unroll
  for (int j = 0; j < NS; j++) nb[j] = 0;
  float invc = 0.f, kt = 0.f;
  if (act) {
    if constexpr (!FIRST) {
#pragma unroll
      for (int j = 0; j < NS; j++) nb[j] = halo.nbl[j * V + x];
      // the site's hop rows go to LDS once (lane 0), read per step by all its lanes
      if (lane == 0) {
        const size_t mF = size_t((ADJ ? 2 : 0) + rF) * NS * V, mB = size_t((ADJ ? 2 : 0) + rB) * NS * V;
#pragma unroll
        for (int j = 0; j < NS; j++) {
          hqs[(sl * 2 + 0) * NS + j] = hv.hrow16[mF + size_t(j) * V + x];
          hqs[(sl * 2 + 1) * NS + j] = hv.hrow16[mB + size_t(j) * V + x];
        }
      }
    }
    invc = hv.invc[x];
    kt = hv.kt[x];
  }
  const HopRow16* hqF = hqs + (sl * 2 + 0) * NS;
  const HopRow16* hqB = hqs + (sl * 2 + 1) * NS;
  const float a = invc * kt;
  const auto in = src.bind();
  const float2 cjs = in.scale(cj);  // the injection's coefficient with what its values owe
  // the wrap terms, weighted: the forward row's on the first slice of lane 0, the backward
  // row's on the last slice of the last lane
  float2 wF = {0.f, 0.f}, wB = {0.f, 0.f};
  if constexpr (!FIRST) {
    if (act && s0 == 0) {
      const float2 w = ldu(d.uPoly + x);
      const float2 g = comp(gin[size_t(Lt - 1) * V + x], rF);
      const float2 t = cmul<float>({w.x, -w.y}, g);
      wF = {d.bc * kt * t.x, d.bc * kt * t.y};
    }
    if (act && len > 0 && s0 + len == Lt) {
      const float2 w = ldu(d.uPoly + x);
      const float2 g = comp(gin[x], rB);
      const float2 t = cmul<float>(w, g);
      wB = {d.bc * kt * t.x, d.bc * kt * t.y};
    }
  }
  constexpr int NW = (NS + 2) / 3;  // packed angle words per site and slice
  struct Loads {
    SpinH v, gc, gh0, gh1;
    SpinH o;  // kDots: the side field
    unsigned u[NW];  // the slots' link angles, packed
  };
  // slice s0 + j of this lane (own point, links, kDots: the side field) and slice c T + j of
  // the halo share.  Every thread issues the same loads at every step (lanes past the column's
  // end re-read its last slice, idle threads site 0; the values are not used), so that the
  // loads of the next kTcDepth steps stay in flight while this one computes; 32-bit offsets
  // (checked at launch).
  // buffer loads: the slice's byte offset in the scalar offset when it is wave-uniform (UL),
  // the thread's site or link in the vector offset
  const unsigned fbytes = unsigned(V) * unsigned(Lt) * unsigned(sizeof(SpinH));
  constexpr unsigned usz = sizeof(unsigned);
  const unsigned ubytes = unsigned(V) * unsigned(NW) * unsigned(Lt) * usz;
  const unsigned vx = xc * unsigned(sizeof(SpinH)), vh0 = hc0 * unsigned(sizeof(SpinH)),
                 vh1 = hc1 * unsigned(sizeof(SpinH)), vxu = xc * usz;
  auto load = [&](int j, Loads& Ld) {
    const unsigned s = unsigned(min(s0 + j, Lt - 1));
    const unsigned oG = s * unsigned(V) * unsigned(sizeof(SpinH));  // the slice's field offset
    const unsigned oU = s * unsigned(NW * V) * usz;  // ... and phases
    const unsigned sG = UL ? oG : 0u, sU = UL ? oU : 0u, vG = UL ? 0u : oG, vU = UL ? 0u : oU;
    Ld.v = in.bat(fbytes, vG + vx, sG);
    if constexpr (kDots) Ld.o = in.bside(fbytes, vG + vx, sG);
    if constexpr (!FIRST) {
#pragma unroll
      for (int w = 0; w < NW; w++) Ld.u[w] = bload_w(d.uSite10, ubytes, vU + vxu, sU + unsigned(w * V) * usz);
      Ld.gc = bload_h(gin, fbytes, vG + vx, sG);
      Ld.gh0 = bload_h(gin, fbytes, vG + vh0, sG);
      if constexpr (H2) Ld.gh1 = bload_h(gin, fbytes, vG + vh1, sG);
    }
  };
  auto transport = [&](const Loads& Ld, const SpinH* tb, const HopRow16* hq, int r) -> float2 {
    half2_t t16 = {(_Float16)0.f, (_Float16)0.f};
#pragma unroll
    for (int q = 0; q < NS; q++) {
      const half2_t u = angle_pair10(Ld.u[q / 3], q % 3);
      const SpinH g = tb[nb[q]];
      t16 += cmul16(u, cmul16(hq[q].s, comp16(g, r)) + cmul16(hq[q].f, comp16(g, 1 - r)));
    }
    return {float(t16.x), float(t16.y)};
  };
  float2 zF[T];   // forward row: the chunk's states (FP32)
  half2_t zB[T];  // backward row: its recurrence inputs, then its states (binary16)
  float2 prevF = {0.f, 0.f};
  float2 dots = {0.f, 0.f};  // kDots: this thread's |v|^2 and Re<o, v> over its chunk
  constexpr int NB = kTcDepth + 1;  // load buffers: this step's and the next kTcDepth steps'
  Loads lb[NB];
#pragma unroll
  for (int j = 0; j < kTcDepth && j < T; j++) load(j, lb[j]);
  if constexpr (WS && !FIRST) __syncthreads();  // the hop rows lane 0 staged, before any lane reads them
#pragma unroll
  for (int j = 0; j < T; j++) {
    if (j + kTcDepth < T) load(j + kTcDepth, lb[(j + kTcDepth) % NB]);
    const Loads& cur = lb[j % NB];
    SpinH* tb = tile + ((j & 1) * C + lane) * row;
    if constexpr (!FIRST) {
      if (j < lenC) tb[sl] = cur.gc;
      if (live && s0 + j < Lt) {
        if (hid0 >= 0) tb[S + sl] = cur.gh0;
        if constexpr (H2) {  // a halo longer than the block's range
          if (hid1 >= 0) tb[2 * S + sl] = cur.gh1;
          for (int q = sl + 2 * S; q < nh; q += S) tb[S + q] = gin[size_t(s0 + j) * V + halo.halo[h0 + q]];
        }
      }
      if constexpr (WS)
        wave_sync();
      else
        __syncthreads();
    }
    if (j < lenC) {
      if constexpr (kDots) {
        if (act) {
          const Spin<float> v = from_half(cur.v), o = from_half(cur.o);
          dots.x += v.a.x * v.a.x + v.a.y * v.a.y + v.b.x * v.b.x + v.b.y * v.b.y;
          dots.y += o.a.x * v.a.x + o.a.y * v.a.y + o.b.x * v.b.x + o.b.y * v.b.y;
        }
      }
      float2 yF = cmul<float>(cjs, comp(cur.v, rF)), yB = cmul<float>(cjs, comp(cur.v, rB));
      if constexpr (!FIRST) {
        const float2 tF = transport(cur, tb, hqF, rF), tB = transport(cur, tb, hqB, rB);
        yF.x -= tF.x;
        yF.y -= tF.y;
        yB.x -= tB.x;
        yB.y -= tB.y;
      }
      if (j == 0) {
        yF.x += wF.x;
        yF.y += wF.y;
      }
      if (j == lenC - 1) {
        yB.x += wB.x;
        yB.y += wB.y;
      }
      prevF = {a * prevF.x + invc * yF.x, a * prevF.y + invc * yF.y};
      zF[j] = prevF;
      zB[j] = half2_t{(_Float16)yB.x, (_Float16)yB.y};
    }
  }
  // local backward recurrence over the chunk held in registers
  float2 nextB = {0.f, 0.f};
#pragma unroll
  for (int j = T - 1; j >= 0; j--) {
    if (j < lenC) {
      nextB = {a * nextB.x + invc * float(zB[j].x), a * nextB.y + invc * float(zB[j].y)};
      zB[j] = half2_t{(_Float16)nextB.x, (_Float16)nextB.y};
    }
  }
  // carries across the lanes of each site
  float aT = 1.f;
#pragma unroll
  for (int j = 0; j < T; j++) aT *= a;
  carry[threadIdx.x] = prevF;                       // forward end value (0 for an empty lane)
  carry[blockDim.x + threadIdx.x] = nextB;          // backward end value
  __syncthreads();
  float2 cF = {0.f, 0.f}, cB = {0.f, 0.f};
  for (int c = 0; c < lane; c++) {
    const float2 e = carry[c * S + sl];
    cF = {aT * cF.x + e.x, aT * cF.y + e.y};
  }
  for (int c = C - 1; c > lane; c--) {
    const int lc = max(0, min(T, Lt - c * T));
    float ac = 1.f;
    for (int q = 0; q < lc; q++) ac *= a;
    const float2 e = carry[blockDim.x + c * S + sl];
    cB = {ac * cB.x + e.x, ac * cB.y + e.y};
  }
  // add the carries and store
  double2 acc = {double(dots.x), double(dots.y)};
  float pF = a;
#pragma unroll
  for (int j = 0; j < T; j++) {
    if (j < lenC) {
      zF[j].x += pF * cF.x;
      zF[j].y += pF * cF.y;
      pF *= a;
    }
  }
  // backward row: add a^(len - j) carry while storing, last slice first
  float pB = a;
#pragma unroll
  for (int j = T - 1; j >= 0; j--) {
    if (j < lenC) {
      const float2 b = {float(zB[j].x) + pB * cB.x, float(zB[j].y) + pB * cB.y};
      pB *= a;
      if (act) {
        const unsigned i = unsigned(s0 + j) * unsigned(V) + unsigned(x);
        if constexpr (rF == 0) {
          put(ptr32(gout, i), zF[j], b);
          epi(size_t(i), stored(gout, zF[j], b), acc);
        } else {
          put(ptr32(gout, i), b, zF[j]);
          epi(size_t(i), stored(gout, b, zF[j]), acc);
        }
      }
      if constexpr (FUSE) zB[j] = half2_t{(_Float16)b.x, (_Float16)b.y};  // row rB as stored
    }
  }
  if constexpr (Epi::kReduce) store_partial(acc, gr.partials, blockIdx.x);
  if constexpr (FUSE) {
    // K's first layer on the stored output (rows: ADJ's forward row rF = 1 is K's backward
    // row, ADJ's rB = 0 K's forward row): forward over cj2 out_0, backward over cj2 out_1
    float2 p0 = {0.f, 0.f};
#pragma unroll
    for (int j = 0; j < T; j++) {
      if (j < lenC) {
        const float2 h0 = {float(zB[j].x), float(zB[j].y)};
        const float2 h1 = {float((_Float16)zF[j].x), float((_Float16)zF[j].y)};
        const float2 y0 = cmul<float>(cj2, h0), y1 = cmul<float>(cj2, h1);
        p0 = {a * p0.x + invc * y0.x, a * p0.y + invc * y0.y};
        zF[j] = p0;
        zB[j] = half2_t{(_Float16)y1.x, (_Float16)y1.y};
      }
    }
    float2 p1 = {0.f, 0.f};
#pragma unroll
    for (int j = T - 1; j >= 0; j--) {
      if (j < lenC) {
        p1 = {a * p1.x + invc * float(zB[j].x), a * p1.y + invc * float(zB[j].y)};
        zB[j] = half2_t{(_Float16)p1.x, (_Float16)p1.y};
      }
    }
    __syncthreads();  // every lane has read the first exchange
    carry[threadIdx.x] = p0;
    carry[blockDim.x + threadIdx.x] = p1;
    __syncthreads();
    float2 c0 = {0.f, 0.f}, c1 = {0.f, 0.f};
    for (int c = 0; c < lane; c++) {
      const float2 e = carry[c * S + sl];
      c0 = {aT * c0.x + e.x, aT * c0.y + e.y};
    }
    for (int c = C - 1; c > lane; c--) {
      const int lc = max(0, min(T, Lt - c * T));
      float ac = 1.f;
      for (int q = 0; q < lc; q++) ac *= a;
      const float2 e = carry[blockDim.x + c * S + sl];
      c1 = {ac * c1.x + e.x, ac * c1.y + e.y};
    }
    float q0 = a;
#pragma unroll
    for (int j = 0; j < T; j++) {
      if (j < lenC) {
        zF[j].x += q0 * c0.x;
        zF[j].y += q0 * c0.y;
        q0 *= a;
      }
    }
    float q1 = a;
#pragma unroll
    for (int j = T - 1; j >= 0; j--) {
      if (j < lenC) {
        const float2 b = {float(zB[j].x) + q1 * c1.x, float(zB[j].y) + q1 * c1.y};
        q1 *= a;
        if (act) put(ptr32(gout2, unsigned(s0 + j) * unsigned(V) + unsigned(x)), zF[j], b);
      }
    }
  }
}

// Lanes per site of the parallel-in-time layer: C = ceil(Lt / T) rounded up to a power of
// two (C * S <= kTcBlock threads, S <= kTcSites sites); at most 64.
inline int tc_lanes(int Lt) {
  const int c = (Lt + kTcChunk - 1) / kTcChunk;
  int p = 1;
  while (p < c) p <<= 1;
  if (p > 64) throw std::invalid_argument("time extent too long for the parallel-in-time layer");
  return p;
}
inline int tc_sites(int Lt) { return std::min(kTcSites, kTcBlock / tc_lanes(Lt)); }

// gout2 != nullptr (ADJ only): also K's first layer with coefficient cj2 into gout2 (FUSE).
template <bool ADJ, bool FIRST, class Src, class Epi>
inline void launch_horner_layer_tc(const HornerView& hv, const DiracView<float, phasepair_t>& tables,
                                   const HaloView& halo, const Src& src, const SpinH* gin, SpinH* gout, float2 cj,
                                   const Epi& epi, GridReduce gr, hipStream_t st, const hipEvent_t* bracket = nullptr,
                                   SpinH* gout2 = nullptr, float2 cj2 = {0.f, 0.f}) {
  if (gout2 && !ADJ) throw std::logic_error("only a K^dagger layer carries K's first layer");
  if (gout2 && (gout2 == gout || gout2 == gin)) throw std::logic_error("fused output aliases an operand");
  const int C = tc_lanes(hv.Lt), S = tc_sites(hv.Lt);
  if (!tables.uPoly || !tables.uSite10)
    throw std::logic_error("parallel-in-time layer needs the temporal gauge");
  if (!FIRST && !halo.nbl) throw std::logic_error("parallel-in-time layer launched without halo tables");
  TileMap tm;
  tm.nxr = (hv.V + S - 1) / S;
  tm.nxr_pad = (tm.nxr + 7) / 8 * 8;
  tm.nst = 1;
  tm.remap = 1;
  const unsigned nblk = unsigned(tm.nxr_pad);
  if (Epi::kReduce && nblk > kReduceCapacity) throw std::invalid_argument("layer grid exceeds the reduction scratch");
  const size_t smem =
      (FIRST ? 0 : size_t(2) * C * (S + halo.hmax) * sizeof(SpinH)) + size_t(2) * kTcBlock * sizeof(float2) +
      (FIRST ? 0 : size_t(S) * 2 * hv.nslot * sizeof(HopRow16));
  if (smem > 64 * 1024) throw std::invalid_argument("parallel-in-time layer: halo too large for LDS");
  if (size_t(hv.V) * size_t(hv.Lt) * sizeof(SpinH) > 0xffffffffull ||
      size_t(hv.V) * size_t((hv.nslot + 2) / 3) * size_t(hv.Lt) * sizeof(unsigned) > 0xffffffffull)
    throw std::invalid_argument("parallel-in-time layer: fields beyond 32-bit offsets");
  if (bracket) QED3_HIP(hipEventRecord(bracket[0], st));
  auto go = [&](auto fuse, auto ul, auto h2, auto ws) {
    constexpr bool F = decltype(fuse)::value && ADJ;
    constexpr bool U = decltype(ul)::value;
    constexpr bool H = decltype(h2)::value && !FIRST;
    constexpr bool W = decltype(ws)::value && U && !FIRST;
    if (hv.nslot == 6)
      hipLaunchKernelGGL((k_horner_layer_tc<6, ADJ, FIRST, kTcChunk, Src, Epi, F, U, H, W>), dim3(nblk),
                         dim3(C * S), smem, st, hv, tables, halo, src, gin, gout, cj, tm, epi, gr, C, gout2, cj2);
    else
      hipLaunchKernelGGL((k_horner_layer_tc<3, ADJ, FIRST, kTcChunk, Src, Epi, F, U, H, W>), dim3(nblk),
                         dim3(C * S), smem, st, hv, tables, halo, src, gin, gout, cj, tm, epi, gr, C, gout2, cj2);
  };
  // the wave-uniform form where a wave's 64 threads belong to one lane (S a multiple of 64),
  // wave-synchronous where a lane is exactly one wave (S = 64); the second halo load only
  // where some block's halo is longer than its range
  const bool ul = S % 64 == 0, ws = S == 64, h2 = !FIRST && halo.hmax > S;
  auto go2 = [&](auto fuse) {
    if (ws && h2)
      go(fuse, std::true_type{}, std::true_type{}, std::true_type{});
    else if (ws)
      go(fuse, std::true_type{}, std::false_type{}, std::true_type{});
    else if (ul && h2)
      go(fuse, std::true_type{}, std::true_type{}, std::false_type{});
    else if (ul)
      go(fuse, std::true_type{}, std::false_type{}, std::false_type{});
    else if (h2)
      go(fuse, std::false_type{}, std::true_type{}, std::false_type{});
    else
      go(fuse, std::false_type{}, std::false_type{}, std::false_type{});
  };
  if (gout2)
    go2(std::true_type{});
  else
    go2(std::false_type{});
  if (bracket) QED3_HIP(hipEventRecord(bracket[1], st));
  if constexpr (Epi::kReduce)
    hipLaunchKernelGGL((k_finalize<Epi>), dim3(1), dim3(256), 0, st, gr.partials, nblk, epi);
}

}  // namespace hk
}  // namespace qed3
