// Learned preconditioner kernels (gfx950): forward (input head, BF16-MFMA channel
// mixing layers, output head), backward (adjoint feature maps, MFMA input-gradient and
// weight-gradient GEMMs) and the Adam training loop.  See mpnn.hpp for the network.
#include <algorithm>
#include <cmath>
#include <complex>
#include <random>
#include <stdexcept>

#include "mpnn.hpp"

namespace qed3 {
namespace {

using bf16 = __bf16;
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

constexpr int C = kNetC;          // complex channels
constexpr int RW = 2 * C;         // real rows of a channel vector (32)
constexpr int KF = 5 * RW;        // real feature length (160)
constexpr int PTS = 64;           // points per block in the MFMA kernels
constexpr int LDSK = KF + 8;      // padded LDS row (bf16) for [col][k] feature tiles

struct NetGeo {
  const int* nbr;
  const int* lnk;
  const float4* slot;     // (C_, S_, cos a_xy, sin a_xy)
  const float2* slot_nb;  // (cos a_yx, sin a_yx)
  const float2* uS;
  const float2* uT;
  const float* kt;  // kappa_t(x)/2 per site
  int V, E, Lt, M;
  float bc;
  size_t np;  // B * Lt * V
};

__device__ __forceinline__ float2 cm(float2 a, float2 b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ float2 cmc(float2 a, float2 b) { return {a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x}; }  // conj(a) b
__device__ __forceinline__ float2 cj(float2 a) { return {a.x, -a.y}; }
__device__ __forceinline__ void cadd_(float2& a, float2 b) { a.x += b.x; a.y += b.y; }

struct PointIdx {
  int b, s, x;
  size_t base;  // (b*Lt + s) * V
  size_t Pp, Pm;
  float bp, bm;  // boundary sign times kappa_t(x)/2 of the two time hops
};

__device__ __forceinline__ PointIdx point_of(const NetGeo& g, size_t P) {
  PointIdx q;
  q.x = int(P % g.V);
  const size_t sg = P / g.V;
  q.b = int(sg / g.Lt);
  q.s = int(sg % g.Lt);
  q.base = sg * g.V;
  const int sp = (q.s == g.Lt - 1) ? 0 : q.s + 1, sm = (q.s == 0) ? g.Lt - 1 : q.s - 1;
  const size_t pb = size_t(q.b) * g.Lt;
  q.Pp = (pb + sp) * g.V + q.x;
  q.Pm = (pb + sm) * g.V + q.x;
  const float kh = g.kt[q.x];
  q.bp = (q.s == g.Lt - 1) ? g.bc * kh : kh;
  q.bm = (q.s == 0) ? g.bc * kh : kh;
  return q;
}

// transports of slot j at site x, slice s
struct SlotT {
  int y;
  float2 t0, t1;   // tau_xy,0 and tau_xy,1
  float2 ea, eb;   // e^{-i alpha}, e^{+i alpha} (alpha of the chosen end)
};

__device__ __forceinline__ SlotT slot_t(const NetGeo& g, int j, int x, int s, bool nb_alpha) {
  SlotT t;
  const size_t si = size_t(j) * g.V + x;
  t.y = g.nbr[si];
  const int l = g.lnk[si];
  float2 u = g.uS[size_t(s) * g.E + (l >= 0 ? l : ~l)];
  if (l < 0) u.y = -u.y;
  const float4 sl = g.slot[si];
  t.t0 = cm(u, float2{sl.x, -sl.y});
  t.t1 = cm(u, float2{sl.x, sl.y});
  float ca = sl.z, sa = sl.w;
  if (nb_alpha) {
    const float2 a2 = g.slot_nb[si];
    ca = a2.x;
    sa = a2.y;
  }
  t.ea = {ca, -sa};
  t.eb = {ca, sa};
  return t;
}

// The five feature maps of a single-channel spinor field at point P.  adj = true gives
// the adjoint maps (F0, F1, F2 with the neighbour's alpha, F4, F3 as derived from
// conj(tau_xy,a) = tau_yx,a).
template <bool ADJ>
__device__ __forceinline__ void feat1(const NetGeo& g, const Spin<float>* __restrict__ r, size_t P,
                                      float2 F[5][2]) {
  const PointIdx q = point_of(g, P);
  const Spin<float> r0 = r[P];
  F[0][0] = r0.a;
  F[0][1] = r0.b;
  F[1][0] = F[1][1] = F[2][0] = F[2][1] = float2{0.f, 0.f};
  for (int j = 0; j < g.M; j++) {
    const SlotT t = slot_t(g, j, q.x, q.s, ADJ);
    const Spin<float> h = r[q.base + t.y];
    cadd_(F[1][0], cm(t.t0, h.a));
    cadd_(F[1][1], cm(t.t1, h.b));
    if (!ADJ) {
      cadd_(F[2][0], cm(cm(t.ea, t.t1), h.b));
      cadd_(F[2][1], cm(cm(t.eb, t.t0), h.a));
    } else {
      cadd_(F[2][0], cm(cm(t.ea, t.t0), h.b));
      cadd_(F[2][1], cm(cm(t.eb, t.t1), h.a));
    }
  }
  const float2 up = g.uT[size_t(q.s) * g.V + q.x];
  const float2 um = cj(g.uT[size_t(q.s == 0 ? g.Lt - 1 : q.s - 1) * g.V + q.x]);
  const Spin<float> hp = r[q.Pp], hm = r[q.Pm];
  // forward map: F3 = b+ U_s h(s+1), F4 = b- conj(U_{s-1}) h(s-1); the adjoint swaps them
  const int i3 = ADJ ? 4 : 3, i4 = ADJ ? 3 : 4;
  F[i3][0] = cm(float2{q.bp * up.x, q.bp * up.y}, hp.a);
  F[i3][1] = cm(float2{q.bp * up.x, q.bp * up.y}, hp.b);
  F[i4][0] = cm(float2{q.bm * um.x, q.bm * um.y}, hm.a);
  F[i4][1] = cm(float2{q.bm * um.x, q.bm * um.y}, hm.b);
}

// ---------------------------------------------------------------- forward kernels
// ADJ = true runs the adjoint network's first stage: adjoint feature maps with the head
// weights prepared by upload_params (conjugated output-head weights, spins of the gamma
// term swapped).  The same holds for k_net_mid and k_net_out below.
template <bool ADJ>
__global__ __launch_bounds__(256) void k_net_in(NetGeo g, const Spin<float>* __restrict__ r,
                                                const float2* __restrict__ v, bf16* __restrict__ h1) {
  const size_t P = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (P >= g.np) return;
  float2 F[5][2];
  feat1<ADJ>(g, r, P, F);
  for (int a = 0; a < 2; a++) {
    bf16 o[RW];
#pragma unroll
    for (int c = 0; c < C; c++) {
      float2 z = {0.f, 0.f};
#pragma unroll
      for (int k = 0; k < 5; k++) cadd_(z, cm(v[(a * 5 + k) * C + c], F[k][a]));
      o[2 * c] = bf16(z.x);
      o[2 * c + 1] = bf16(z.y);
    }
    bf16x8* dst = reinterpret_cast<bf16x8*>(h1 + (P * 2 + a) * RW);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      bf16x8 w;
      for (int i = 0; i < 8; i++) w[i] = o[8 * q + i];
      dst[q] = w;
    }
  }
}

__device__ __forceinline__ void load4(const bf16* __restrict__ h, size_t P, int a, int grp, float2 o[4]) {
  const bf16x8 w = *reinterpret_cast<const bf16x8*>(h + (P * 2 + a) * RW + 8 * grp);
#pragma unroll
  for (int i = 0; i < 4; i++) o[i] = {float(w[2 * i]), float(w[2 * i + 1])};
}

// Features of a C-channel BF16 field for channels [4 grp, 4 grp + 4) at point P (ADJ: the
// adjoint maps, F_k^dagger in slot k, as in feat1).
template <bool ADJ>
__device__ __forceinline__ void feat4(const NetGeo& g, const bf16* __restrict__ h, size_t P, int grp,
                                      float2 F[5][2][4]) {
  const PointIdx q = point_of(g, P);
  load4(h, P, 0, grp, F[0][0]);
  load4(h, P, 1, grp, F[0][1]);
#pragma unroll
  for (int i = 0; i < 4; i++) F[1][0][i] = F[1][1][i] = F[2][0][i] = F[2][1][i] = float2{0.f, 0.f};
  for (int j = 0; j < g.M; j++) {
    const SlotT t = slot_t(g, j, q.x, q.s, ADJ);
    float2 h0[4], h1[4];
    load4(h, q.base + t.y, 0, grp, h0);
    load4(h, q.base + t.y, 1, grp, h1);
    const float2 k0 = ADJ ? cm(t.ea, t.t0) : cm(t.ea, t.t1), k1 = ADJ ? cm(t.eb, t.t1) : cm(t.eb, t.t0);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      cadd_(F[1][0][i], cm(t.t0, h0[i]));
      cadd_(F[1][1][i], cm(t.t1, h1[i]));
      cadd_(F[2][0][i], cm(k0, h1[i]));
      cadd_(F[2][1][i], cm(k1, h0[i]));
    }
  }
  const float2 up = g.uT[size_t(q.s) * g.V + q.x];
  const float2 um = cj(g.uT[size_t(q.s == 0 ? g.Lt - 1 : q.s - 1) * g.V + q.x]);
  const float2 wp = {q.bp * up.x, q.bp * up.y}, wm = {q.bm * um.x, q.bm * um.y};
  float2 t[4];
  const int i3 = ADJ ? 4 : 3, i4 = ADJ ? 3 : 4;
  for (int a = 0; a < 2; a++) {
    load4(h, q.Pp, a, grp, t);
#pragma unroll
    for (int i = 0; i < 4; i++) F[i3][a][i] = cm(wp, t[i]);
    load4(h, q.Pm, a, grp, t);
#pragma unroll
    for (int i = 0; i < 4; i++) F[i4][a][i] = cm(wm, t[i]);
  }
}

// MFMA operand layouts of v_mfma_f32_32x32x16_bf16 (wave64): A(32x16): lane l holds row
// l%32, k = 8(l/32)..+8; B(16x32): column l%32, k = 8(l/32)..+8; D(32x32): column l%32,
// rows 8(i/4) + 4(l/32) + i%4 for accumulator element i.
// Columns of a block are ordered (spin a, point p): n = a PTS + p, so waves 0-1 hold spin
// 0 and waves 2-3 spin 1, and each wave multiplies by its spin's weight matrix.
template <bool ADJ>
__global__ __launch_bounds__(256) void k_net_mid(NetGeo g, const bf16* __restrict__ hin,
                                                 const bf16* __restrict__ A /*[2][32][160]*/,
                                                 bf16* __restrict__ hout) {
  __shared__ __attribute__((aligned(16))) bf16 Fl[2 * PTS][LDSK];
  const int t = threadIdx.x;
  const size_t P0 = size_t(blockIdx.x) * PTS;
  {
    const int p = t >> 2, grp = t & 3;
    const size_t P = P0 + p;
    float2 F[5][2][4];
    if (P < g.np) {
      feat4<ADJ>(g, hin, P, grp, F);
    } else {
      for (int k = 0; k < 5; k++)
        for (int a = 0; a < 2; a++)
          for (int i = 0; i < 4; i++) F[k][a][i] = float2{0.f, 0.f};
    }
    for (int a = 0; a < 2; a++)
      for (int k = 0; k < 5; k++) {
        bf16x8 w;
        for (int i = 0; i < 4; i++) {
          w[2 * i] = bf16(F[k][a][i].x);
          w[2 * i + 1] = bf16(F[k][a][i].y);
        }
        *reinterpret_cast<bf16x8*>(&Fl[a * PTS + p][k * RW + 8 * grp]) = w;
      }
  }
  __syncthreads();
  const int lane = t & 63, w = t >> 6;
  const int m = lane & 31, hi = lane >> 5;
  const bf16* Aw = A + size_t(w >> 1) * RW * KF;
  f32x16 acc = {};
#pragma unroll
  for (int kk = 0; kk < KF / 16; kk++) {
    const bf16x8 a = *reinterpret_cast<const bf16x8*>(Aw + m * KF + 16 * kk + 8 * hi);
    const bf16x8 b = *reinterpret_cast<const bf16x8*>(&Fl[32 * w + m][16 * kk + 8 * hi]);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
  }
  const int n = 32 * w + m;  // column = a PTS + p
  const size_t P = P0 + (n % PTS);
  if (P < g.np) {
    bf16* dst = hout + (P * 2 + n / PTS) * RW;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 q4;
      for (int i = 0; i < 4; i++) q4[i] = bf16(acc[4 * j + i]);
      *reinterpret_cast<decltype(q4)*>(dst + 8 * j + 4 * hi) = q4;
    }
  }
}

// out = sum_k F_k(hhat_k) + skip r, hhat_k = sum_c u_k[c] h_c (the maps act channel-wise)
template <bool ADJ>
__global__ __launch_bounds__(256) void k_net_out(NetGeo g, const bf16* __restrict__ h,
                                                 const float2* __restrict__ u,
                                                 const Spin<float>* __restrict__ r, float skip,
                                                 Spin<float>* __restrict__ out) {
  const size_t P = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (P >= g.np) return;
  const PointIdx q = point_of(g, P);
  // sum_c u[ao][k][c] h_c at spin ai of point Q (ao: the output spin the weights belong to)
  auto proj = [&](size_t Q, int k, int ai, int ao) {
    const bf16x8* src = reinterpret_cast<const bf16x8*>(h + (Q * 2 + ai) * RW);
    const float2* uk = u + (ao * 5 + k) * C;
    float2 z = {0.f, 0.f};
#pragma unroll
    for (int part = 0; part < 4; part++) {
      const bf16x8 w8 = src[part];
#pragma unroll
      for (int i = 0; i < 4; i++) cadd_(z, cm(uk[4 * part + i], float2{float(w8[2 * i]), float(w8[2 * i + 1])}));
    }
    return z;
  };
  float2 o0 = proj(P, 0, 0, 0), o1 = proj(P, 0, 1, 1);
  for (int j = 0; j < g.M; j++) {
    const SlotT t = slot_t(g, j, q.x, q.s, ADJ);
    const size_t Q = q.base + t.y;
    const float2 a0 = proj(Q, 1, 0, 0), a1 = proj(Q, 1, 1, 1);
    const float2 b0 = proj(Q, 2, 0, 1), b1 = proj(Q, 2, 1, 0);
    cadd_(o0, cm(t.t0, a0));
    cadd_(o1, cm(t.t1, a1));
    cadd_(o0, cm(cm(t.ea, ADJ ? t.t0 : t.t1), b1));
    cadd_(o1, cm(cm(t.eb, ADJ ? t.t1 : t.t0), b0));
  }
  const float2 up = g.uT[size_t(q.s) * g.V + q.x];
  const float2 um = cj(g.uT[size_t(q.s == 0 ? g.Lt - 1 : q.s - 1) * g.V + q.x]);
  const float2 wp = {q.bp * up.x, q.bp * up.y}, wm = {q.bm * um.x, q.bm * um.y};
  const int k3 = ADJ ? 4 : 3, k4 = ADJ ? 3 : 4;  // the adjoint of F3 reads s-1, of F4 s+1
  cadd_(o0, cm(wp, proj(q.Pp, k3, 0, 0)));
  cadd_(o1, cm(wp, proj(q.Pp, k3, 1, 1)));
  cadd_(o0, cm(wm, proj(q.Pm, k4, 0, 0)));
  cadd_(o1, cm(wm, proj(q.Pm, k4, 1, 1)));
  const Spin<float> rp = r[P];
  Spin<float> s;
  s.a = {o0.x + skip * rp.a.x, o0.y + skip * rp.a.y};
  s.b = {o1.x + skip * rp.b.x, o1.y + skip * rp.b.y};
  out[P] = s;
}

// ---------------------------------------------------------------- backward kernels
// loss pieces: per pair |e|^2 and |out - e|^2
__global__ __launch_bounds__(256) void k_net_loss(const Spin<float>* __restrict__ out,
                                                  const Spin<float>* __restrict__ e, size_t np,
                                                  size_t per, double* acc) {
  const size_t P = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  double en = 0, dn = 0;
  int b = -1;
  if (P < np) {
    const Spin<float> o = out[P], t = e[P];
    en = double(t.a.x) * t.a.x + double(t.a.y) * t.a.y + double(t.b.x) * t.b.x + double(t.b.y) * t.b.y;
    const double d0 = o.a.x - t.a.x, d1 = o.a.y - t.a.y, d2 = o.b.x - t.b.x, d3 = o.b.y - t.b.y;
    dn = d0 * d0 + d1 * d1 + d2 * d2 + d3 * d3;
    b = int(P / per);
  }
  if (b >= 0) {
    atomicAdd(acc + 2 * b, en);
    atomicAdd(acc + 2 * b + 1, dn);
  }
}

__global__ __launch_bounds__(256) void k_net_gout(const Spin<float>* __restrict__ out,
                                                  const Spin<float>* __restrict__ e, size_t np,
                                                  size_t per, const double* acc, int B,
                                                  Spin<float>* __restrict__ gout) {
  const size_t P = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (P >= np) return;
  const int b = int(P / per);
  const float sc = float(2.0 / acc[2 * b] / B);
  const Spin<float> o = out[P], t = e[P];
  Spin<float> gg;
  gg.a = {sc * (o.a.x - t.a.x), sc * (o.a.y - t.a.y)};
  gg.b = {sc * (o.b.x - t.b.x), sc * (o.b.y - t.b.y)};
  gout[P] = gg;
}

// Output head backward: A_k = F_k^dagger gout.  Input spin a of F_k feeds output spin
// o_k(a) = a (k != 2) or 1 - a (k = 2, the gamma map), so dh[c,a] = sum_k conj(u[o_k(a)][k][c])
// A_k,a and du[o][k][c] = A_k,a conj(h_c,a) with o = o_k(a) (per-point contributions,
// summed by k_colsum).
__global__ __launch_bounds__(256) void k_net_out_bwd(NetGeo g, const Spin<float>* __restrict__ gout,
                                                     const bf16* __restrict__ h,
                                                     const float2* __restrict__ u,
                                                     float* __restrict__ dh, float* __restrict__ contrib) {
  const size_t P = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (P >= g.np) return;
  float2 A[5][2];
  feat1<true>(g, gout, P, A);
  for (int c = 0; c < C; c++) {
    for (int a = 0; a < 2; a++) {
      float2 z = {0.f, 0.f};
      for (int k = 0; k < 5; k++) cadd_(z, cmc(u[((k == 2 ? 1 - a : a) * 5 + k) * C + c], A[k][a]));
      dh[(P * 2 + a) * RW + 2 * c] = z.x;
      dh[(P * 2 + a) * RW + 2 * c + 1] = z.y;
    }
    if (!contrib) continue;  // adjoint apply: no weight gradient
    float* cdst = contrib + P * (20 * C);
    float2 hc[2];
    for (int a = 0; a < 2; a++) {
      const bf16* src = h + (P * 2 + a) * RW + 2 * c;
      hc[a] = {float(src[0]), float(src[1])};
    }
    for (int a = 0; a < 2; a++)
      for (int k = 0; k < 5; k++) {
        const int o = k == 2 ? 1 - a : a;
        const float2 z = cmc(hc[a], A[k][a]);
        cdst[2 * ((o * 5 + k) * C + c)] = z.x;
        cdst[2 * ((o * 5 + k) * C + c) + 1] = z.y;
      }
  }
}

// Input head backward: contributions to dv[a][k][c] = dh1[c,a] conj(F_k(r)_a)
__global__ __launch_bounds__(256) void k_net_in_bwd(NetGeo g, const Spin<float>* __restrict__ r,
                                                    const float* __restrict__ dh,
                                                    float* __restrict__ contrib) {
  const size_t P = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (P >= g.np) return;
  float2 F[5][2];
  feat1<false>(g, r, P, F);
  float* cdst = contrib + P * (20 * C);
  for (int c = 0; c < C; c++)
    for (int a = 0; a < 2; a++) {
      const float2 ga = {dh[(P * 2 + a) * RW + 2 * c], dh[(P * 2 + a) * RW + 2 * c + 1]};
      for (int k = 0; k < 5; k++) {
        const float2 z = cmc(F[k][a], ga);
        cdst[2 * ((a * 5 + k) * C + c)] = z.x;
        cdst[2 * ((a * 5 + k) * C + c) + 1] = z.y;
      }
    }
}

// Input gradient of a mixing layer, G = R_a^T dz per column of spin a (MFMA, M = 160 in 5
// tiles, K = 32): G[P][k][a][32] FP32.  Columns ordered (spin, point) as in k_net_mid.
__global__ __launch_bounds__(256) void k_net_mid_G(size_t np, const float* __restrict__ dz,
                                                   const bf16* __restrict__ AT /*[2][160][32]*/,
                                                   float* __restrict__ G) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int m = lane & 31, hi = lane >> 5;
  const int n = 32 * w + m;
  const size_t P = size_t(blockIdx.x) * PTS + (n % PTS);
  const int a = n / PTS;
  const bf16* ATa = AT + size_t(w >> 1) * KF * RW;
  bf16x8 b[2];
  for (int kk = 0; kk < 2; kk++)
    for (int i = 0; i < 8; i++)
      b[kk][i] = bf16(P < np ? dz[(P * 2 + a) * RW + 16 * kk + 8 * hi + i] : 0.f);
  for (int tile = 0; tile < 5; tile++) {
    f32x16 acc = {};
    for (int kk = 0; kk < 2; kk++) {
      const bf16x8 av = *reinterpret_cast<const bf16x8*>(ATa + (32 * tile + m) * RW + 16 * kk + 8 * hi);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, b[kk], acc, 0, 0, 0);
    }
    if (P < np) {
      float* dst = G + ((P * 5 + tile) * 2 + a) * RW;
      for (int i = 0; i < 16; i++) dst[8 * (i / 4) + 4 * hi + (i % 4)] = acc[i];
    }
  }
}

// Weight gradient of a mixing layer, dR_a (32 x 160) = sum over the block's 64 columns of
// spin a of dz (32) F^T (160) (MFMA with K over those columns), one partial per block:
// partial[block][a][32][160].
__global__ __launch_bounds__(256) void k_net_mid_dW(NetGeo g, const bf16* __restrict__ hin,
                                                    const float* __restrict__ dz,
                                                    float* __restrict__ partial) {
  __shared__ __attribute__((aligned(16))) bf16 FT[KF][2 * PTS + 8];   // [feature][col]
  __shared__ __attribute__((aligned(16))) bf16 ZT[RW][2 * PTS + 8];   // [row][col]
  const int t = threadIdx.x;
  const size_t P0 = size_t(blockIdx.x) * PTS;
  {
    const int p = t >> 2, grp = t & 3;
    const size_t P = P0 + p;
    float2 F[5][2][4];
    const bool ok = P < g.np;
    if (ok) feat4<false>(g, hin, P, grp, F);
    for (int a = 0; a < 2; a++) {
      const int col = a * PTS + p;
      for (int k = 0; k < 5; k++)
        for (int i = 0; i < 4; i++) {
          const int f = k * RW + 2 * (4 * grp + i);
          FT[f][col] = bf16(ok ? F[k][a][i].x : 0.f);
          FT[f + 1][col] = bf16(ok ? F[k][a][i].y : 0.f);
        }
      for (int i = 0; i < 8; i++) {
        const int row = 8 * grp + i;
        ZT[row][col] = bf16(ok ? dz[(P * 2 + a) * RW + row] : 0.f);
      }
    }
  }
  __syncthreads();
  const int lane = t & 63, w = t >> 6;
  const int m = lane & 31, hi = lane >> 5;
  // 10 jobs (5 feature tiles x 2 spins) over 4 waves: wave w takes jobs w, w+4, w+8
  for (int job = w; job < 10; job += 4) {
    const int tile = job % 5, a = job / 5;
    f32x16 acc = {};
    for (int kk = a * (PTS / 16); kk < (a + 1) * (PTS / 16); kk++) {
      const bf16x8 av = *reinterpret_cast<const bf16x8*>(&ZT[m][16 * kk + 8 * hi]);
      const bf16x8 bv = *reinterpret_cast<const bf16x8*>(&FT[32 * tile + m][16 * kk + 8 * hi]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
    }
    float* dst = partial + (size_t(blockIdx.x) * 2 + a) * (RW * KF);
    for (int i = 0; i < 16; i++) {
      const int row = 8 * (i / 4) + 4 * hi + (i % 4);
      dst[row * KF + 32 * tile + m] = acc[i];
    }
  }
}

// dh_in(y) = sum_k F_k^dagger G_k  (gather form; thread per point and 4-channel group)
__global__ __launch_bounds__(256) void k_net_adj(NetGeo g, const float* __restrict__ G,
                                                 float* __restrict__ dh) {
  const size_t T = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t P = T >> 2;
  const int grp = int(T & 3);
  if (P >= g.np) return;
  const PointIdx q = point_of(g, P);
  auto ld = [&](size_t Q, int k, int a, float2 o[4]) {
    const float4* src = reinterpret_cast<const float4*>(G + ((Q * 5 + k) * 2 + a) * RW + 8 * grp);
    const float4 v0 = src[0], v1 = src[1];
    o[0] = {v0.x, v0.y};
    o[1] = {v0.z, v0.w};
    o[2] = {v1.x, v1.y};
    o[3] = {v1.z, v1.w};
  };
  float2 r0[4], r1[4], t[4], t2[4];
  ld(P, 0, 0, r0);
  ld(P, 0, 1, r1);
  for (int j = 0; j < g.M; j++) {
    const SlotT s = slot_t(g, j, q.x, q.s, true);
    const size_t Q = q.base + s.y;
    // spin 0: tau_yx,0 G1_0 + e^{-i a_xy} tau_yx,0 G2_1 ; spin 1: tau_yx,1 G1_1 + e^{i a_xy} tau_yx,1 G2_0
    ld(Q, 1, 0, t);
    ld(Q, 2, 1, t2);
    const float2 k0 = cm(s.ea, s.t0), k1 = cm(s.eb, s.t1);
    for (int i = 0; i < 4; i++) {
      cadd_(r0[i], cm(s.t0, t[i]));
      cadd_(r0[i], cm(k0, t2[i]));
    }
    ld(Q, 1, 1, t);
    ld(Q, 2, 0, t2);
    for (int i = 0; i < 4; i++) {
      cadd_(r1[i], cm(s.t1, t[i]));
      cadd_(r1[i], cm(k1, t2[i]));
    }
  }
  // adjoint of F3 from s-1: bm conj(U_{s-1}) G3(s-1);  adjoint of F4 from s+1: bp U_s G4(s+1)
  const float2 up = g.uT[size_t(q.s) * g.V + q.x];
  const float2 um = cj(g.uT[size_t(q.s == 0 ? g.Lt - 1 : q.s - 1) * g.V + q.x]);
  const float2 wm = {q.bm * um.x, q.bm * um.y}, wp = {q.bp * up.x, q.bp * up.y};
  for (int a = 0; a < 2; a++) {
    float2* acc = a == 0 ? r0 : r1;
    ld(q.Pm, 3, a, t);
    for (int i = 0; i < 4; i++) cadd_(acc[i], cm(wm, t[i]));
    ld(q.Pp, 4, a, t);
    for (int i = 0; i < 4; i++) cadd_(acc[i], cm(wp, t[i]));
  }
  for (int a = 0; a < 2; a++) {
    const float2* acc = a == 0 ? r0 : r1;
    float4* dst = reinterpret_cast<float4*>(dh + (P * 2 + a) * RW + 8 * grp);
    dst[0] = float4{acc[0].x, acc[0].y, acc[1].x, acc[1].y};
    dst[1] = float4{acc[2].x, acc[2].y, acc[3].x, acc[3].y};
  }
}

// out[j] = sum_i buf[i][j] in FP64, one block per column group
__global__ __launch_bounds__(256) void k_colsum(const float* __restrict__ buf, size_t rows, int cols,
                                                double* __restrict__ out) {
  __shared__ double sh[256];
  const int j = blockIdx.x;
  double s = 0;
  for (size_t i = threadIdx.x; i < rows; i += blockDim.x) s += buf[i * cols + j];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = sh[0];
}

template <typename T>
T* dmalloc(size_t n) {
  T* p = nullptr;
  QED3_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  return p;
}

}  // namespace

Net::Net(Dirac& D, int layers, int seed) : D_(D), L_(layers) {
  if (layers < 2 || layers > 16) throw std::invalid_argument("layers must be in [2,16]");
  const HopTables& h = D.tables();
  const size_t MS = size_t(h.max_nn) * h.V;
  std::vector<float> sl(4 * MS), nb(2 * MS);
  for (size_t i = 0; i < MS; i++) {
    sl[4 * i] = float(h.hopD[4 * i]);
    sl[4 * i + 1] = float(h.hopD[4 * i + 1]);
    sl[4 * i + 2] = float(std::cos(h.slot_alpha[i]));
    sl[4 * i + 3] = float(std::sin(h.slot_alpha[i]));
    nb[2 * i] = float(std::cos(h.slot_alpha_nb[i]));
    nb[2 * i + 1] = float(std::sin(h.slot_alpha_nb[i]));
  }
  // Gershgorin bound on |D|_2: the 2x2 blocks have norms |c_x|, kappa_xy/2 (r+1) per
  // spatial hop and kappa_t(x)/2 (r+1) per time hop; row and column sums agree.
  double bound = 0;
  for (int x = 0; x < h.V; x++) {
    double row = std::abs(h.diag[x]) + (h.Lt > 1 ? 2.0 : 1.0) * 0.5 * h.kappa_t[x] * (h.r + 1);
    for (int j = 0; j < h.nn_count[x]; j++) {
      const size_t i = size_t(j) * h.V + x;
      row += std::hypot(h.hopD[4 * i], h.hopD[4 * i + 1]) * (h.r + 1);
    }
    bound = std::max(bound, row);
  }
  skip_ = bound > 0 ? float(1.0 / bound) : 0.f;
  slot_ = dmalloc<float>(sl.size());
  slot_nb_ = dmalloc<float>(nb.size());
  QED3_HIP(hipMemcpy(slot_, sl.data(), sl.size() * 4, hipMemcpyHostToDevice));
  QED3_HIP(hipMemcpy(slot_nb_, nb.data(), nb.size() * 4, hipMemcpyHostToDevice));
  std::vector<float> kth(h.V);
  for (int x = 0; x < h.V; x++) kth[x] = float(0.5 * h.kappa_t[x]);
  kt_ = dmalloc<float>(h.V);
  QED3_HIP(hipMemcpy(kt_, kth.data(), kth.size() * 4, hipMemcpyHostToDevice));
  heads_ = dmalloc<float>(80 * C);
  const int nmid = std::max(L_ - 2, 0);
  QED3_HIP(hipMalloc(&mixA_, std::max(nmid, 1) * 2 * RW * KF * sizeof(bf16)));
  QED3_HIP(hipMalloc(&mixAT_, std::max(nmid, 1) * 2 * RW * KF * sizeof(bf16)));
  QED3_HIP(hipMalloc(&mixAH_, std::max(nmid, 1) * 2 * RW * KF * sizeof(bf16)));
  // initialisation: heads and mixing weights ~ N(0, sigma^2) complex, small
  p_.assign(2 * nparams(), 0.0);
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd;
  for (size_t i = 0; i < 10 * C; i++) {  // v
    p_[2 * i] = 0.3 * nd(rng);
    p_[2 * i + 1] = 0.3 * nd(rng);
  }
  const size_t nw = size_t(nmid) * 10 * C * C;
  for (size_t i = 0; i < nw; i++) {
    p_[2 * (10 * C + i)] = nd(rng) / std::sqrt(5.0 * C);
    p_[2 * (10 * C + i) + 1] = nd(rng) / std::sqrt(5.0 * C);
  }
  for (size_t i = 0; i < 10 * C; i++) {  // u
    p_[2 * (10 * C + nw + i)] = 0.3 * nd(rng) / C;
    p_[2 * (10 * C + nw + i) + 1] = 0.3 * nd(rng) / C;
  }
  m_.assign(p_.size(), 0.0);
  v_.assign(p_.size(), 0.0);
  upload_params();
}

Net::~Net() {
  for (void* p : {(void*)slot_, (void*)slot_nb_, (void*)kt_, (void*)heads_, mixA_, mixAT_, mixAH_, act_, (void*)gfield_,
                  (void*)gfield2_, (void*)G_, (void*)out_, (void*)gout_, (void*)partial_, (void*)red_})
    if (p) (void)hipFree(p);
}

void Net::get_params(double* p) const { std::copy(p_.begin(), p_.end(), p); }

void Net::set_params(const double* p) {
  std::copy(p, p + p_.size(), p_.begin());
  upload_params();
}

// Real form of z = W f on interleaved (re, im) rows: R[2c][2d] = Re W, R[2c][2d+1] = -Im W,
// R[2c+1][2d] = Im W, R[2c+1][2d+1] = Re W, for the five feature blocks side by side; one
// matrix per layer and output spin.  The adjoint network N^dagger uses the same kernels
// with the adjoint feature maps and, at output spin a, the weights W_a,k^dagger (k != 2) and
// W_{1-a},2^dagger (the gamma map swaps spins); its heads are the conjugated output and
// input heads with the same spin swap.
void Net::upload_params() {
  const size_t nw = size_t(std::max(L_ - 2, 0)) * 10 * C * C;
  auto cpar = [&](size_t i) { return std::complex<double>(p_[2 * i], p_[2 * i + 1]); };
  auto v_at = [&](int a, int k, int c) { return cpar((size_t(a) * 5 + k) * C + c); };
  auto u_at = [&](int a, int k, int c) { return cpar(10 * C + nw + (size_t(a) * 5 + k) * C + c); };
  auto w_at = [&](int l, int a, int k, int c, int d) {
    return cpar(10 * C + (((size_t(l) * 2 + a) * 5 + k) * C + c) * C + d);
  };
  // heads: v, u, then the adjoint network's input head (from u) and output head (from v)
  std::vector<float> heads(80 * C);
  for (int a = 0; a < 2; a++)
    for (int k = 0; k < 5; k++)
      for (int c = 0; c < C; c++) {
        const size_t i = (size_t(a) * 5 + k) * C + c;
        const int sa = k == 2 ? 1 - a : a;
        const std::complex<double> vv = v_at(a, k, c), uu = u_at(a, k, c);
        const std::complex<double> vh = std::conj(u_at(sa, k, c)), uh = std::conj(v_at(sa, k, c));
        heads[2 * i] = float(vv.real());
        heads[2 * i + 1] = float(vv.imag());
        heads[20 * C + 2 * i] = float(uu.real());
        heads[20 * C + 2 * i + 1] = float(uu.imag());
        heads[40 * C + 2 * i] = float(vh.real());
        heads[40 * C + 2 * i + 1] = float(vh.imag());
        heads[60 * C + 2 * i] = float(uh.real());
        heads[60 * C + 2 * i + 1] = float(uh.imag());
      }
  QED3_HIP(hipMemcpy(heads_, heads.data(), heads.size() * 4, hipMemcpyHostToDevice));
  const int nmid = std::max(L_ - 2, 0);
  std::vector<bf16> A(size_t(std::max(nmid, 1)) * 2 * RW * KF), AT(A.size()), AH(A.size());
  auto put = [](std::vector<bf16>& M, size_t base, int c, int d, int k, std::complex<double> w, bool tr) {
    const float re = float(w.real()), im = float(w.imag());
    const int col = k * RW + 2 * d;
    if (!tr) {
      M[base + (2 * c) * KF + col] = bf16(re);
      M[base + (2 * c) * KF + col + 1] = bf16(-im);
      M[base + (2 * c + 1) * KF + col] = bf16(im);
      M[base + (2 * c + 1) * KF + col + 1] = bf16(re);
    } else {
      M[base + size_t(col) * RW + 2 * c] = bf16(re);
      M[base + size_t(col + 1) * RW + 2 * c] = bf16(-im);
      M[base + size_t(col) * RW + 2 * c + 1] = bf16(im);
      M[base + size_t(col + 1) * RW + 2 * c + 1] = bf16(re);
    }
  };
  for (int l = 0; l < nmid; l++)
    for (int sa = 0; sa < 2; sa++)
      for (int k = 0; k < 5; k++)
        for (int c = 0; c < C; c++)
          for (int d = 0; d < C; d++) {
            const size_t base = (size_t(l) * 2 + sa) * RW * KF;
            const std::complex<double> w = w_at(l, sa, k, c, d);
            put(A, base, c, d, k, w, false);
            put(AT, base, c, d, k, w, true);
            put(AH, base, c, d, k, std::conj(w_at(l, k == 2 ? 1 - sa : sa, k, d, c)), false);
          }
  QED3_HIP(hipMemcpy(mixA_, A.data(), A.size() * sizeof(bf16), hipMemcpyHostToDevice));
  QED3_HIP(hipMemcpy(mixAT_, AT.data(), AT.size() * sizeof(bf16), hipMemcpyHostToDevice));
  QED3_HIP(hipMemcpy(mixAH_, AH.data(), AH.size() * sizeof(bf16), hipMemcpyHostToDevice));
}

void Net::ensure(int B, bool training) {
  if (B <= capB_ && (!training || cap_train_)) return;
  const size_t np = size_t(std::max(B, capB_)) * D_.points();
  for (void* p : {act_, (void*)gfield_, (void*)gfield2_, (void*)G_, (void*)out_, (void*)gout_,
                  (void*)partial_, (void*)red_})
    if (p) (void)hipFree(p);
  act_ = nullptr;
  gfield_ = gfield2_ = G_ = nullptr;
  out_ = gout_ = nullptr;
  partial_ = red_ = nullptr;
  const int nact = training ? std::max(L_ - 1, 2) : 2;
  QED3_HIP(hipMalloc(&act_, size_t(nact) * np * 2 * RW * sizeof(bf16)));
  if (training) {
    gfield_ = dmalloc<float>(np * 2 * RW);
    gfield2_ = dmalloc<float>(np * 2 * RW);
    G_ = dmalloc<float>(np * 5 * 2 * RW);
    out_ = dmalloc<Spin<float>>(np);
    gout_ = dmalloc<Spin<float>>(np);
    const size_t nblk = (np + PTS - 1) / PTS;
    partial_ = reinterpret_cast<double*>(dmalloc<float>(std::max(nblk * 2 * RW * KF, np * 20 * C)));
    red_ = dmalloc<double>(2 * RW * KF + 2 * size_t(std::max(B, capB_)));
  }
  capB_ = std::max(B, capB_);
  cap_train_ = cap_train_ || training;
}

void Net::apply(Spin<float>* out, const Spin<float>* r, int B) {
  ensure(B, false);
  const DiracView<float> dv = D_.view<float>();
  const HopTables& h = D_.tables();
  NetGeo g{dv.nbr, dv.lnk, reinterpret_cast<const float4*>(slot_), reinterpret_cast<const float2*>(slot_nb_),
           dv.uS, dv.uT, kt_, h.V, h.E, h.Lt, h.max_nn, dv.bc, size_t(B) * D_.points()};
  hipStream_t st = D_.stream();
  bf16* a0 = static_cast<bf16*>(act_);
  bf16* a1 = a0 + g.np * 2 * RW;
  const unsigned nb = unsigned((g.np + 255) / 256), nbm = unsigned((g.np + PTS - 1) / PTS);
  const float2* v = reinterpret_cast<const float2*>(heads_);
  const float2* u = v + 10 * C;
  hipLaunchKernelGGL(k_net_in<false>, dim3(nb), dim3(256), 0, st, g, r, v, a0);
  for (int l = 0; l < L_ - 2; l++) {
    hipLaunchKernelGGL(k_net_mid<false>, dim3(nbm), dim3(256), 0, st, g, a0,
                       static_cast<const bf16*>(mixA_) + size_t(l) * 2 * RW * KF, a1);
    std::swap(a0, a1);
  }
  hipLaunchKernelGGL(k_net_out<false>, dim3(nb), dim3(256), 0, st, g, a0, u, r, skip_, out);
  QED3_HIP(hipGetLastError());
}

// out = N^dagger(r): the forward kernels in adjoint mode, layers in reverse order (the
// net is linear, so no forward activations are needed); costs the same as N.
void Net::apply_adj(Spin<float>* out, const Spin<float>* r, int B) {
  ensure(B, false);
  const DiracView<float> dv = D_.view<float>();
  const HopTables& h = D_.tables();
  NetGeo g{dv.nbr, dv.lnk, reinterpret_cast<const float4*>(slot_), reinterpret_cast<const float2*>(slot_nb_),
           dv.uS, dv.uT, kt_, h.V, h.E, h.Lt, h.max_nn, dv.bc, size_t(B) * D_.points()};
  hipStream_t st = D_.stream();
  bf16* a0 = static_cast<bf16*>(act_);
  bf16* a1 = a0 + g.np * 2 * RW;
  const unsigned nb = unsigned((g.np + 255) / 256), nbm = unsigned((g.np + PTS - 1) / PTS);
  const float2* vh = reinterpret_cast<const float2*>(heads_) + 20 * C;
  const float2* uh = vh + 10 * C;
  hipLaunchKernelGGL(k_net_in<true>, dim3(nb), dim3(256), 0, st, g, r, vh, a0);
  for (int l = L_ - 3; l >= 0; l--) {
    hipLaunchKernelGGL(k_net_mid<true>, dim3(nbm), dim3(256), 0, st, g, a0,
                       static_cast<const bf16*>(mixAH_) + size_t(l) * 2 * RW * KF, a1);
    std::swap(a0, a1);
  }
  hipLaunchKernelGGL(k_net_out<true>, dim3(nb), dim3(256), 0, st, g, a0, uh, r, skip_, out);
  QED3_HIP(hipGetLastError());
}

double Net::loss_grad(const Spin<float>* r, const Spin<float>* e, int B, std::vector<double>& grad) {
  ensure(B, true);
  const DiracView<float> dv = D_.view<float>();
  const HopTables& h = D_.tables();
  NetGeo g{dv.nbr, dv.lnk, reinterpret_cast<const float4*>(slot_), reinterpret_cast<const float2*>(slot_nb_),
           dv.uS, dv.uT, kt_, h.V, h.E, h.Lt, h.max_nn, dv.bc, size_t(B) * D_.points()};
  hipStream_t st = D_.stream();
  const size_t np = g.np, per = D_.points();
  const unsigned nb = unsigned((np + 255) / 256), nbm = unsigned((np + PTS - 1) / PTS);
  const float2* v = reinterpret_cast<const float2*>(heads_);
  const float2* u = v + 10 * C;
  const int nmid = L_ - 2;
  auto act = [&](int j) { return static_cast<bf16*>(act_) + size_t(j) * np * 2 * RW; };
  // forward, keeping h1..h^{L-1}
  hipLaunchKernelGGL(k_net_in<false>, dim3(nb), dim3(256), 0, st, g, r, v, act(0));
  for (int l = 0; l < nmid; l++)
    hipLaunchKernelGGL(k_net_mid<false>, dim3(nbm), dim3(256), 0, st, g, act(l),
                       static_cast<const bf16*>(mixA_) + size_t(l) * 2 * RW * KF, act(l + 1));
  hipLaunchKernelGGL(k_net_out<false>, dim3(nb), dim3(256), 0, st, g, act(nmid), u, r, skip_, out_);
  // loss
  QED3_HIP(hipMemsetAsync(red_, 0, 2 * B * sizeof(double), st));
  hipLaunchKernelGGL(k_net_loss, dim3(nb), dim3(256), 0, st, out_, e, np, per, red_);
  hipLaunchKernelGGL(k_net_gout, dim3(nb), dim3(256), 0, st, out_, e, np, per, red_, B, gout_);
  std::vector<double> pairs(2 * B);
  QED3_HIP(hipMemcpyAsync(pairs.data(), red_, pairs.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  // output head
  float* contrib = reinterpret_cast<float*>(partial_);
  double* red = red_ + 2 * B;
  grad.assign(p_.size(), 0.0);
  const size_t nw = size_t(nmid) * 10 * C * C;
  std::vector<double> tmp(2 * RW * KF);
  hipLaunchKernelGGL(k_net_out_bwd, dim3(nb), dim3(256), 0, st, g, gout_, act(nmid), u, gfield_, contrib);
  hipLaunchKernelGGL(k_colsum, dim3(20 * C), dim3(256), 0, st, contrib, np, 20 * C, red);
  QED3_HIP(hipMemcpyAsync(tmp.data(), red, 20 * C * sizeof(double), hipMemcpyDeviceToHost, st));
  QED3_HIP(hipStreamSynchronize(st));
  for (int i = 0; i < 20 * C; i++) grad[2 * (10 * C + nw) + i] = tmp[i];
  // middle layers, last to first
  float* gin = gfield_;
  float* gnext = gfield2_;
  for (int l = nmid - 1; l >= 0; l--) {
    hipLaunchKernelGGL(k_net_mid_dW, dim3(nbm), dim3(256), 0, st, g, act(l), gin, contrib);
    hipLaunchKernelGGL(k_colsum, dim3(2 * RW * KF), dim3(256), 0, st, contrib, size_t(nbm), 2 * RW * KF, red);
    QED3_HIP(hipMemcpyAsync(tmp.data(), red, 2 * RW * KF * sizeof(double), hipMemcpyDeviceToHost, st));
    hipLaunchKernelGGL(k_net_mid_G, dim3(nbm), dim3(256), 0, st, np, gin,
                       static_cast<const bf16*>(mixAT_) + size_t(l) * 2 * RW * KF, G_);
    hipLaunchKernelGGL(k_net_adj, dim3(unsigned((4 * np + 255) / 256)), dim3(256), 0, st, g, G_, gnext);
    QED3_HIP(hipStreamSynchronize(st));
    for (int sa = 0; sa < 2; sa++)
    for (int k = 0; k < 5; k++)
      for (int c = 0; c < C; c++)
        for (int d = 0; d < C; d++) {
          const int col = k * RW + 2 * d;
          const double* t2 = tmp.data() + size_t(sa) * RW * KF;
          const double rr = t2[(2 * c) * KF + col], ri = t2[(2 * c) * KF + col + 1];
          const double ir = t2[(2 * c + 1) * KF + col], ii = t2[(2 * c + 1) * KF + col + 1];
          const size_t wi = 10 * C + (((size_t(l) * 2 + sa) * 5 + k) * C + c) * C + d;
          grad[2 * wi] = rr + ii;
          grad[2 * wi + 1] = ir - ri;
        }
    std::swap(gin, gnext);
  }
  // input head
  hipLaunchKernelGGL(k_net_in_bwd, dim3(nb), dim3(256), 0, st, g, r, gin, contrib);
  hipLaunchKernelGGL(k_colsum, dim3(20 * C), dim3(256), 0, st, contrib, np, 20 * C, red);
  QED3_HIP(hipMemcpyAsync(tmp.data(), red, 20 * C * sizeof(double), hipMemcpyDeviceToHost, st));
  QED3_HIP(hipStreamSynchronize(st));
  QED3_HIP(hipGetLastError());
  for (int i = 0; i < 20 * C; i++) grad[i] = tmp[i];
  double loss = 0;
  for (int b = 0; b < B; b++) loss += pairs[2 * b + 1] / pairs[2 * b] / B;
  return loss;
}

double Net::train(const Spin<float>* r, const Spin<float>* e, int npairs, int B, int epochs,
                  double lr, std::vector<double>* history) {
  const size_t N = D_.points();
  const double b1 = 0.9, b2 = 0.999, eps = 1e-8;
  double last = 0;
  std::vector<double> grad;
  for (int ep = 0; ep < epochs; ep++) {
    double tot = 0;
    int nb = 0;
    for (int i = 0; i + B <= npairs; i += B) {
      const double loss = loss_grad(r + size_t(i) * N, e + size_t(i) * N, B, grad);
      tot += loss;
      nb++;
      step_++;
      const double c1 = 1.0 - std::pow(b1, double(step_)), c2 = 1.0 - std::pow(b2, double(step_));
      for (size_t k = 0; k < p_.size(); k++) {
        m_[k] = b1 * m_[k] + (1 - b1) * grad[k];
        v_[k] = b2 * v_[k] + (1 - b2) * grad[k] * grad[k];
        p_[k] -= lr * (m_[k] / c1) / (std::sqrt(v_[k] / c2) + eps);
      }
      upload_params();
    }
    last = nb ? tot / nb : 0;
    if (history) history->push_back(last);
  }
  return last;
}

}  // namespace qed3
