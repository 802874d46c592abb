// Device side of the time-exact Horner preconditioner (horner.hpp): one layer
// g_out = T^{-1}(c v + S_th g_in) (or its adjoint form) per launch.
//
// Thread = site x.  A block owns a range of R Hilbert-consecutive sites; each thread walks
// its site's whole time column twice at once, once per spin row, each in the direction of
// that row's recurrence (forward for spin up under T, backward for spin down, reversed under
// T^dagger), all threads in step.  Per step the block stages the input field g_in of both
// slices (its range and spatial halo) in LDS, gathers the neighbours' spinors from there,
// forms y = c v - sum_j u_j (m_same g_a + m_flip g_abar) for each row and runs one step of
// each recurrence state = (y + kappa_t u^t state) / c_x, storing state as the layer's
// output.  The first layer of a network (FIRST) has no input field: no staging, no barrier.
#pragma once
#include "horner.hpp"
#include "stencil.hpp"

namespace qed3 {
namespace hk {

// planar complex storage in FP32 or binary16
__device__ __forceinline__ float2 ldc(const float2* p) { return *p; }
__device__ __forceinline__ float2 ldc(const half2_t* p) {
  const half2_t h = *p;
  return {float(h.x), float(h.y)};
}
__device__ __forceinline__ void stc(float2* p, float2 v) { *p = v; }
__device__ __forceinline__ void stc(half2_t* p, float2 v) {
  half2_t h;
  h.x = (_Float16)v.x;
  h.y = (_Float16)v.y;
  unsigned int bits;
  __builtin_memcpy(&bits, &h, 4);
  __builtin_nontemporal_store(bits, reinterpret_cast<unsigned int*>(p));
}
// value as stored (rounded to the storage type), for norms of what the next kernel reads
__device__ __forceinline__ float2 rnd(float2 v, const float2*) { return v; }
__device__ __forceinline__ float2 rnd(float2 v, const half2_t*) {
  return {float((_Float16)v.x), float((_Float16)v.y)};
}

// both spin components of point i of a planar field (n points per plane)
template <typename S>
__device__ __forceinline__ float4 ld2(const S* g, size_t i, size_t n) {
  const float2 a = ldc(g + i), b = ldc(g + n + i);
  return {a.x, a.y, b.x, b.y};
}

// Injection sources v_a(i).
template <typename S>
struct SrcPlanar {
  const S* p;
  size_t n;
  __device__ SrcPlanar bind() const { return *this; }
  __device__ float2 operator()(int a, size_t i) const { return ldc(p + size_t(a) * n + i); }
  __device__ S raw(int a, size_t i) const { return p[size_t(a) * n + i]; }
};
// planar binary16 field chosen on the device: buffer par[0] & 1 of two (the solver's
// double-buffered residual)
struct SrcPlanarPar {
  const half2_t* p0;
  const half2_t* p1;
  const int* par;
  size_t n;
  __device__ SrcPlanar<half2_t> bind() const { return {(*par & 1) ? p1 : p0, n}; }
};

struct EpiNone {
  static constexpr bool kReduce = false;
  __device__ bool skip() const { return false; }
  __device__ void operator()(int, size_t, float2, double2&) const {}
  __device__ void operator()(double2) const {}
};

// raw (unconverted) loads, so a slice's values can be fetched a step ahead and converted
// only when used
template <typename S>
struct Raw2 {
  S a, b;
};
template <typename S>
__device__ __forceinline__ Raw2<S> ldraw2(const S* g, size_t i, size_t n) {
  return {g[i], g[n + i]};
}
__device__ __forceinline__ float4 cvt(const Raw2<float2>& r) { return {r.a.x, r.a.y, r.b.x, r.b.y}; }
__device__ __forceinline__ float4 cvt(const Raw2<half2_t>& r) {
  return {float(r.a.x), float(r.a.y), float(r.b.x), float(r.b.y)};
}
__device__ __forceinline__ float2 cvt1(const float2& r) { return r; }
__device__ __forceinline__ float2 cvt1(const half2_t& r) { return {float(r.x), float(r.y)}; }

// The loads of one slice for one spin row, issued a step before they are used.
template <int NS, typename S, typename P>
struct SliceLoads {
  S v;
  P u[NS];
  P ut;
  Raw2<S> gc, gh0, gh1;
};

// Thread = site x, carrying both spin rows: the row whose recurrence runs forward in time
// (spin up under T, spin down under T^dagger) walks s = 0 .. Lt-1 while the other walks
// s = Lt-1 .. 0, both in the same loop, so a thread has two independent chains and the
// grid one block per site range.  Each step stages the input field of both slices (range
// and halo) in LDS, one barrier per step.
template <int NS, bool ADJ, bool FIRST, bool TG, typename S, class Src, class Epi, typename P>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NS == 3 ? 5 : 3)))
void k_horner_layer(HornerView hv, DiracView<float, P> d, HaloView halo,
                                                      Src src, const S* __restrict__ gin,
                                                      S* __restrict__ gout, float2 cj, TileMap tm, Epi epi,
                                                      GridReduce gr) {
  if (epi.skip()) return;  // uniform
  extern __shared__ float4 hk_tile[];
  int xr, st;
  tile_of(tm, blockIdx.x, xr, st);
  (void)st;
  const int V = hv.V, Lt = hv.Lt, R = blockDim.x;
  const size_t n = size_t(V) * Lt;
  const bool live = xr < tm.nxr;
  const int x0 = xr * R;
  const int x = live ? x0 + int(threadIdx.x) : V;
  const bool act = x < V;
  constexpr int rF = ADJ ? 1 : 0, rB = 1 - rF;  // forward- and backward-walking rows
  const int h0 = (live && !FIRST) ? halo.halo_off[xr] : 0;
  const int nh = (live && !FIRST) ? halo.halo_off[xr + 1] - h0 : 0;
  const int hid0 = (int(threadIdx.x) < nh) ? halo.halo[h0 + threadIdx.x] : -1;
  const int hid1 = (int(threadIdx.x) + R < nh) ? halo.halo[h0 + threadIdx.x + R] : -1;
  int nb[NS], lk[NS];
  float4 hrF[NS], hrB[NS];
  float invc = 0.f, kt = 0.f;
  if (act) {
    if constexpr (!FIRST) {
      const float4* rowsF = hv.hrow + (size_t((ADJ ? 2 : 0) + rF) * NS) * V;
      const float4* rowsB = hv.hrow + (size_t((ADJ ? 2 : 0) + rB) * NS) * V;
#pragma unroll
      for (int j = 0; j < NS; j++) {
        nb[j] = halo.nbl[j * V + x];
        lk[j] = d.lnk[j * V + x];
        hrF[j] = rowsF[size_t(j) * V + x];
        hrB[j] = rowsB[size_t(j) * V + x];
      }
    }
    invc = hv.invc[x];
    kt = hv.kt[x];
  }
  const auto in = src.bind();
  // Loads of slice s for row r: the injection, (original gauge) the temporal link the
  // recurrence crosses (s-1 -> s walking forward, s -> s+1 walking backward), the link
  // phases and the input field at the site and at its share of the halo.
  auto load = [&](int s, int r, bool fwd, bool first, SliceLoads<NS, S, P>& L) {
    const size_t sV = size_t(s) * V;
    if (act) {
      L.v = in.raw(r, sV + x);
      if constexpr (!TG) {
        if (!first) L.ut = fwd ? d.uT[size_t(s - 1) * V + x] : d.uT[sV + x];
      }
      if constexpr (!FIRST) {
        const P* uS = d.uS + size_t(s) * d.E;
#pragma unroll
        for (int j = 0; j < NS; j++) L.u[j] = uS[lk[j] >= 0 ? lk[j] : ~lk[j]];
        L.gc = ldraw2(gin, sV + x, n);
      }
    }
    if constexpr (!FIRST) {
      if (hid0 >= 0) L.gh0 = ldraw2(gin, sV + hid0, n);
      if (hid1 >= 0) L.gh1 = ldraw2(gin, sV + hid1, n);
    }
  };
  auto stage = [&](const SliceLoads<NS, S, P>& L, float4* tb, int s) {
    if (act) tb[threadIdx.x] = cvt(L.gc);
    if (hid0 >= 0) tb[R + threadIdx.x] = cvt(L.gh0);
    if (hid1 >= 0) tb[2 * R + threadIdx.x] = cvt(L.gh1);
    for (int q = int(threadIdx.x) + 2 * R; q < nh; q += R)
      tb[R + q] = ld2(gin, size_t(s) * V + halo.halo[h0 + q], n);
  };
  // one recurrence step of row r at slice s from its staged tile
  auto step = [&](const SliceLoads<NS, S, P>& L, const float4* tb, const float4* hr, int r, int s, bool fwd,
                  bool first, float2& state, double2& acc) {
    float2 y = cmul<float>(cj, cvt1(L.v));
    if constexpr (!FIRST) {
#pragma unroll
      for (int j = 0; j < NS; j++) {
        float2 u = ldu(&L.u[j]);
        if (lk[j] < 0) u.y = -u.y;
        const float4 g = tb[nb[j]];
        const float2 ga = r == 0 ? float2{g.x, g.y} : float2{g.z, g.w};
        const float2 gb = r == 0 ? float2{g.z, g.w} : float2{g.x, g.y};
        const float2 cm = cadd<float>(cmul<float>({hr[j].x, hr[j].y}, ga), cmul<float>({hr[j].z, hr[j].w}, gb));
        const float2 t = cmul<float>(u, cm);
        y.x -= t.x;
        y.y -= t.y;
      }
    }
    if (!first) {
      float2 t = state;
      if constexpr (!TG) {
        float2 ut = ldu(&L.ut);
        if (fwd) ut.y = -ut.y;
        t = cmul<float>(ut, state);
      }
      y.x += kt * t.x;
      y.y += kt * t.y;
    } else {
      y.x += state.x;  // the wrap term of S, already weighted (0 in a first layer)
      y.y += state.y;
    }
    state = {invc * y.x, invc * y.y};
    const size_t i = size_t(s) * V + x;
    stc(gout + size_t(r) * n + i, state);
    epi(r, i, rnd(state, gout), acc);
  };
  float2 stF = {0.f, 0.f}, stB = {0.f, 0.f};
  double2 acc = {0.0, 0.0};
  // S holds the time wrap (the link from Lt-1 to 0, which T leaves out): each chain starts
  // from bc kappa_t w g_r(the far end of its walk), w the wrap link's phase, conjugated for
  // the forward walk.
  if constexpr (!FIRST) {
    if (act) {
      cplx_t<float> w = TG ? ldu(d.uPoly + x) : ldu(d.uT + size_t(Lt - 1) * V + x);
      const float4 gl = cvt(ldraw2(gin, size_t(Lt - 1) * V + x, n));
      const float4 g0 = cvt(ldraw2(gin, x, n));
      const float2 gF = rF == 0 ? float2{gl.x, gl.y} : float2{gl.z, gl.w};
      const float2 gB = rB == 0 ? float2{g0.x, g0.y} : float2{g0.z, g0.w};
      const float2 tF = cmul<float>({w.x, -w.y}, gF), tB = cmul<float>(w, gB);
      stF = {d.bc * kt * tF.x, d.bc * kt * tF.y};
      stB = {d.bc * kt * tB.x, d.bc * kt * tB.y};
    }
  }
  SliceLoads<NS, S, P> cF, cB, nF, nB;
  load(0, rF, true, true, cF);
  load(Lt - 1, rB, false, true, cB);
  const int tstride = R + halo.hmax;
  for (int k = 0; k < Lt; k++) {
    const int sF = k, sB = Lt - 1 - k;
    if (k + 1 < Lt) {  // a step ahead: in flight during this step's work
      load(sF + 1, rF, true, false, nF);
      load(sB - 1, rB, false, false, nB);
    }
    float4* tF = hk_tile + (k & 1) * 2 * tstride;
    float4* tB = tF + tstride;
    if constexpr (!FIRST) {
      stage(cF, tF, sF);
      stage(cB, tB, sB);
      __syncthreads();
    }
    if (act) {
      step(cF, tF, hrF, rF, sF, true, k == 0, stF, acc);
      step(cB, tB, hrB, rB, sB, false, k == 0, stB, acc);
    }
    cF = nF;
    cB = nB;
  }
  if constexpr (Epi::kReduce) store_partial(acc, gr.partials, blockIdx.x);
}

// Launch one layer over the lattice.  `tables` supplies the spatial link phases (and, without
// TG, the temporal ones) of the gauge the fields live in.
template <bool ADJ, bool FIRST, bool TG, typename S, class Src, class Epi, typename P>
inline void launch_horner_layer(const HornerView& hv, const DiracView<float, P>& tables, const HaloView& halo,
                                const Src& src, const S* gin, S* gout, float2 cj, const Epi& epi,
                                GridReduce gr, hipStream_t st, const hipEvent_t* bracket = nullptr) {
  constexpr int R = 256;
  TileMap tm;
  tm.nxr = (hv.V + R - 1) / R;
  tm.nxr_pad = (tm.nxr + 7) / 8 * 8;
  tm.nst = 1;
  tm.remap = 1;
  const unsigned nblk = unsigned(tm.nxr_pad);
  if (Epi::kReduce && nblk > kReduceCapacity) throw std::invalid_argument("layer grid exceeds the reduction scratch");
  const size_t smem = FIRST ? 0 : size_t(4) * (R + halo.hmax) * sizeof(float4);
  if (bracket) QED3_HIP(hipEventRecord(bracket[0], st));
  if (hv.nslot == 6)
    hipLaunchKernelGGL((k_horner_layer<6, ADJ, FIRST, TG, S, Src, Epi, P>), dim3(nblk), dim3(R), smem, st, hv,
                       tables, halo, src, gin, gout, cj, tm, epi, gr);
  else
    hipLaunchKernelGGL((k_horner_layer<3, ADJ, FIRST, TG, S, Src, Epi, P>), dim3(nblk), dim3(R), smem, st, hv,
                       tables, halo, src, gin, gout, cj, tm, epi, gr);
  if (bracket) QED3_HIP(hipEventRecord(bracket[1], st));
  if constexpr (Epi::kReduce)
    hipLaunchKernelGGL((k_finalize<Epi>), dim3(1), dim3(256), 0, st, gr.partials, nblk, epi);
}

}  // namespace hk
}  // namespace qed3
