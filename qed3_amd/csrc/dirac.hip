// Dirac operator host class, gauge-phase preparation, FP64 dot and the all-links
// pseudofermion force kernel (gfx950).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "dirac.hpp"
#include "stencil.hpp"

namespace qed3 {

namespace {
constexpr size_t kReduceCap = kReduceCapacity;

template <typename T>
__global__ void k_phases(const double* __restrict__ ang, cplx_t<T>* __restrict__ u, size_t n) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) {
    double sn, cs;
    sincos(ang[i], &sn, &cs);
    u[i] = {T(cs), T(sn)};
  }
}

// 16-bit fixed-point angle in units of 2 pi / 65536, wrapped into [-pi, pi)
__device__ __forceinline__ phase16_t angle16(double a) {
  const long long k = llrint(a * (65536.0 / (2.0 * M_PI)));
  return phase16_t{static_cast<short>(static_cast<unsigned short>(k & 0xFFFF))};
}

__global__ void k_phases_half(const double* __restrict__ ang, phase16_t* __restrict__ u, size_t n) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) u[i] = angle16(ang[i]);
}

// Temporal gauge, per site: phi_0 = 0, phi_{s+1} = phi_s + theta^t_s (FP64 and 16-bit), and
// the wrap link's angle phi_{Lt-1} + theta^t_{Lt-1} = the sum of the site's temporal angles.
__global__ void k_tg_sites(const double* __restrict__ angT, double* __restrict__ phi64,
                           phase16_t* __restrict__ rot, phase16_t* __restrict__ poly, int V, int Lt) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= V) return;
  double phi = 0.0;
  for (int s = 0; s < Lt; s++) {
    const size_t i = size_t(s) * V + x;
    phi64[i] = phi;
    rot[i] = angle16(phi);
    phi += angT[i];
  }
  poly[x] = angle16(phi);
}

// Temporal gauge, per spatial link l = (x -> y) at slice s: theta_l + phi_s(x) - phi_s(y).
__global__ void k_tg_links(const double* __restrict__ angS, const double* __restrict__ phi64,
                           const int4* __restrict__ ends, phase16_t* __restrict__ out, int E, int V, int Lt) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= size_t(Lt) * E) return;
  const int s = int(i / E), l = int(i % E);
  const int4 e = ends[l];
  out[i] = angle16(angS[i] + phi64[size_t(s) * V + e.x] - phi64[size_t(s) * V + e.y]);
}

// (cos, sin) of 16-bit angles as binary16 pairs
__global__ void k_angle_pairs(const phase16_t* __restrict__ a, phasepair_t* __restrict__ out, size_t n) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float2 u = ldu(a + i);
  phasepair_t p;
  p.v.x = (_Float16)u.x;
  p.v.y = (_Float16)u.y;
  out[i] = p;
}

// Per site and slot, the 16-bit angle of the slot's link as the site sees it ([Lt][M][V], M
// the slots per site): negated where the link table gives the link reversed (~id).
__global__ void k_site_angles(const phase16_t* __restrict__ uS, const int* __restrict__ lnk,
                              phase16_t* __restrict__ out, int V, int E, int M, int Lt) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= size_t(Lt) * M * V) return;
  const int x = int(i % V);
  const size_t sq = i / V;
  const int q = int(sq % M), s = int(sq / M);
  const int l = lnk[size_t(q) * V + x];
  phase16_t a = uS[size_t(s) * E + (l >= 0 ? l : ~l)];
  if (l < 0) a.v = short(-a.v);
  out[i] = a;
}

// The per-site angles rounded to 10 bits (to nearest, 2 pi wrapping to 0) and packed three
// slots to a word: word w of site x at slice s holds slot 3w + k in bits 10k..10k+9 ([Lt][W][V],
// W = ceil(M / 3), slots past M zero).
__global__ void k_site_angles10(const phase16_t* __restrict__ a, unsigned* __restrict__ out, int V, int M, int Lt) {
  const int W = (M + 2) / 3;
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= size_t(Lt) * W * V) return;
  const int x = int(i % V);
  const size_t sw = i / V;
  const int w = int(sw % W), s = int(sw / W);
  unsigned word = 0;
  for (int k = 0; k < 3 && 3 * w + k < M; k++) {
    const unsigned u16 = unsigned(static_cast<unsigned short>(a[(size_t(s) * M + 3 * w + k) * V + x].v));
    word |= (((u16 + 32u) >> 6) & 1023u) << (10 * k);
  }
  out[i] = word;
}

// <x,y> = sum conj(x) y over complex double fields of n points.
__global__ __launch_bounds__(256) void k_dot(const Spin<double>* __restrict__ x,
                                             const Spin<double>* __restrict__ y, size_t n,
                                             GridReduce gr) {
  double2 acc = {0.0, 0.0};
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x) {
    const Spin<double> a = x[i], b = y[i];
    const double2 p = cconjmul<double>(a.a, b.a), q = cconjmul<double>(a.b, b.b);
    acc.x += p.x + q.x;
    acc.y += p.y + q.y;
  }
  store_partial(acc, gr.partials, blockIdx.x);
}

struct FinDot {
  double2* out;
  __device__ bool skip() const { return false; }
  __device__ void operator()(double2 t) const { *out = t; }
};

// Re <w, dD/dtheta_i eta> for link i: spatial links (s,l) first, then temporal links (s,x)
// (the two derivative blocks of dirac_ext.h:233-269 for a spatial link, :272-302 for a
// temporal one).  i < Lt (E + V).
__device__ __forceinline__ double link_bilinear(const DiracView<double>& d, const int4* __restrict__ ends,
                                                size_t i, const Spin<double>* __restrict__ w,
                                                const Spin<double>* __restrict__ eta) {
  using C = double2;
  const size_t nS = size_t(d.Lt) * d.E;
  const double r = d.r;
  if (i < nS) {
    const int s = int(i / d.E), l = int(i % d.E);
    const int4 e = ends[l];  // x, y, slot of y at x, slot of x at y
    const int x = e.x, y = e.y;
    const C u = d.uS[i];
    const size_t ox = size_t(s) * d.V + x, oy = size_t(s) * d.V + y;
    const Spin<double> ex = eta[ox], ey = eta[oy], wx = w[ox], wy = w[oy];
    auto hmul = [&](const double4 h, const C p0, const C p1, C& o0, C& o1) {
      const double rC = r * h.x, rS = r * h.y, A = h.z, B = h.w;
      o0 = {-rC * p0.x - rS * p0.y + A * p1.x + B * p1.y, -rC * p0.y + rS * p0.x + A * p1.y - B * p1.x};
      o1 = {A * p0.x - B * p0.y - rC * p1.x + rS * p1.y, A * p0.y + B * p0.x - rC * p1.y - rS * p1.x};
    };
    // row x: i e^{i theta} H_xy eta(y)
    C a0, a1;
    hmul(d.hop[size_t(e.z) * d.V + x], cmul<double>(u, ey.a), cmul<double>(u, ey.b), a0, a1);
    // row y: -i e^{-i theta} H_yx eta(x)
    const C uc = {u.x, -u.y};
    C b0, b1;
    hmul(d.hop[size_t(e.w) * d.V + y], cmul<double>(uc, ex.a), cmul<double>(uc, ex.b), b0, b1);
    // Re <w, i v> = -Im <w, v>
    const C t1 = cadd<double>(cconjmul<double>(wx.a, a0), cconjmul<double>(wx.b, a1));
    const C t2 = cadd<double>(cconjmul<double>(wy.a, b0), cconjmul<double>(wy.b, b1));
    return -t1.y + t2.y;
  }
  const size_t k = i - nS;
  const int s = int(k / d.V), x = int(k % d.V);
  const int sp = (s == d.Lt - 1) ? 0 : s + 1;
  const double sg = (s == d.Lt - 1) ? double(d.bc) : 1.0;
  const double kh = d.kth[x];
  const double m0 = kh * (-r - 1.0), m1 = kh * (-r + 1.0);  // T- = diag(m0,m1), T+ = diag(m1,m0)
  const C u = d.uT[k];
  const size_t o0 = size_t(s) * d.V + x, o1 = size_t(sp) * d.V + x;
  const Spin<double> e0 = eta[o0], e1 = eta[o1], w0 = w[o0], w1 = w[o1];
  // dD[s+1 <- s] = -i sg T- e^{-i theta};  dD[s <- s+1] = +i sg T+ e^{+i theta}
  const C uc = {u.x, -u.y};
  const C p0 = cmul<double>(uc, e0.a), p1 = cmul<double>(uc, e0.b);
  const C q0 = cmul<double>(u, e1.a), q1 = cmul<double>(u, e1.b);
  const C tP = cadd<double>(cconjmul<double>(w1.a, cscale<double>(m0, p0)),
                            cconjmul<double>(w1.b, cscale<double>(m1, p1)));
  const C tM = cadd<double>(cconjmul<double>(w0.a, cscale<double>(m1, q0)),
                            cconjmul<double>(w0.b, cscale<double>(m0, q1)));
  // Re<w, -i sg tP> = sg Im(tP);  Re<w, i sg tM> = -sg Im(tM)
  return sg * (tP.y - tM.y);
}

// Pseudofermion force, one thread per spatial link (s,l) followed by one per temporal
// link (s,x):  F = -2 Re <w, dD/dtheta eta> with w = D eta  (dirac_pf.h:107-121).
__global__ __launch_bounds__(256) void k_force(DiracView<double> d, const int4* __restrict__ ends,
                                               const Spin<double>* __restrict__ eta,
                                               const Spin<double>* __restrict__ w,
                                               double* __restrict__ fs, double* __restrict__ ft) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t nS = size_t(d.Lt) * d.E, nT = size_t(d.Lt) * d.V;
  if (i >= nS + nT) return;
  const double f = -2.0 * link_bilinear(d, ends, i, w, eta);
  if (i < nS)
    fs[i] = f;
  else
    ft[i - nS] = f;
}

// f_i (+)= c (Re <u1, dD/dtheta_i v1> + Re <u2, dD/dtheta_i v2>) for every link; the second
// pair is skipped when u2 is null, f is overwritten when init.
__global__ __launch_bounds__(256) void k_force_acc(DiracView<double> d, const int4* __restrict__ ends,
                                                   const Spin<double>* __restrict__ u1,
                                                   const Spin<double>* __restrict__ v1,
                                                   const Spin<double>* __restrict__ u2,
                                                   const Spin<double>* __restrict__ v2, double c, int init,
                                                   double* __restrict__ fs, double* __restrict__ ft) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t nS = size_t(d.Lt) * d.E, nT = size_t(d.Lt) * d.V;
  if (i >= nS + nT) return;
  double t = link_bilinear(d, ends, i, u1, v1);
  if (u2) t += link_bilinear(d, ends, i, u2, v2);
  double* f = i < nS ? fs + i : ft + (i - nS);
  *f = init ? c * t : *f + c * t;
}

template <typename T>
T* dalloc(size_t n) {
  T* p = nullptr;
  if (n) QED3_HIP(hipMalloc(&p, n * sizeof(T)));
  return p;
}

template <typename T, typename S>
T* dupload(const std::vector<S>& v, hipStream_t st) {
  std::vector<T> tmp(v.begin(), v.end());
  T* p = dalloc<T>(tmp.size());
  if (!tmp.empty()) {
    QED3_HIP(hipMemcpyAsync(p, tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice, st));
    QED3_HIP(hipStreamSynchronize(st));
  }
  return p;
}

}  // namespace

size_t reduce_capacity() { return kReduceCap; }

namespace {
// Plain streaming copy: the calibration kernel for HBM counters and achievable bandwidth.
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ a, float4* __restrict__ b,
                                              size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x)
    b[i] = a[i];
}
}  // namespace

double time_copy(size_t bytes, int reps) {
  const size_t n = bytes / sizeof(float4);
  float4 *a = nullptr, *b = nullptr;
  QED3_HIP(hipMalloc(&a, n * sizeof(float4)));
  QED3_HIP(hipMalloc(&b, n * sizeof(float4)));
  QED3_HIP(hipMemset(a, 0, n * sizeof(float4)));
  hipEvent_t e0, e1;
  QED3_HIP(hipEventCreate(&e0));
  QED3_HIP(hipEventCreate(&e1));
  const unsigned grid = unsigned(std::min<size_t>((n + 255) / 256, 256 * 1024));
  hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n);
  QED3_HIP(hipEventRecord(e0, 0));
  for (int i = 0; i < reps; i++) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n);
  QED3_HIP(hipEventRecord(e1, 0));
  QED3_HIP(hipEventSynchronize(e1));
  float ms = 0;
  QED3_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return double(ms) / std::max(reps, 1);
}

Dirac::Dirac(const Geometry& g, int lattice, int Lt, int bc, double r, double M5, double at)
    : h_(make_hop_tables(g, lattice, Lt, r, M5, at)), bc_(bc) {
  QED3_HIP(hipStreamCreate(&own_stream_));
  stream_ = own_stream_;
  upload_tables(d64_);
  upload_tables(d32_);
  // link ends for the force kernel
  std::vector<int> ends(size_t(h_.E) * 4, -1);
  for (int x = 0; x < h_.V; x++)
    for (int j = 0; j < h_.nn_count[x]; j++) {
      const int sl = h_.lnk[size_t(j) * h_.V + x];
      if (sl >= 0) {
        ends[size_t(sl) * 4 + 0] = x;
        ends[size_t(sl) * 4 + 1] = h_.nbr[size_t(j) * h_.V + x];
        ends[size_t(sl) * 4 + 2] = j;
      } else {
        ends[size_t(~sl) * 4 + 3] = j;
      }
    }
  link_ends_ = dupload<int>(ends, stream_);
  angS_ = dalloc<double>(size_t(h_.Lt) * h_.E);
  angT_ = dalloc<double>(size_t(h_.Lt) * h_.V);
  partials_ = dalloc<double2>(kReduceCap);
  counter_ = dalloc<unsigned int>(1);
  dot_out_ = dalloc<double2>(1);
  QED3_HIP(hipMemsetAsync(counter_, 0, sizeof(unsigned int), stream_));
  QED3_HIP(hipStreamSynchronize(stream_));
  for (auto& a : tile_)
    for (auto& b : a) b = StencilTile{256, std::min(4, h_.Lt)};
}

template <typename T>
void Dirac::upload_tables(DeviceTables<T>& d) {
  d.nbr = dupload<int>(h_.nbr, stream_);
  d.lnk = dupload<int>(h_.lnk, stream_);
  const size_t ns = size_t(h_.max_nn) * h_.V;
  std::vector<real4_t<T>> hd(ns), hdd(ns);
  for (size_t i = 0; i < ns; i++) {
    hd[i] = {T(h_.hopD[4 * i]), T(h_.hopD[4 * i + 1]), T(h_.hopD[4 * i + 2]), T(h_.hopD[4 * i + 3])};
    hdd[i] = {T(h_.hopDdag[4 * i]), T(h_.hopDdag[4 * i + 1]), T(h_.hopDdag[4 * i + 2]),
              T(h_.hopDdag[4 * i + 3])};
  }
  d.hop = dupload<real4_t<T>>(hd, stream_);
  d.hopd = dupload<real4_t<T>>(hdd, stream_);
  d.diag = dupload<T>(h_.diag, stream_);
  std::vector<double> kth(h_.V);
  for (int x = 0; x < h_.V; x++) kth[x] = 0.5 * h_.kappa_t[x];
  d.kth = dupload<T>(kth, stream_);
  d.uS = dalloc<cplx_t<T>>(size_t(h_.Lt) * h_.E);
  d.uT = dalloc<cplx_t<T>>(size_t(h_.Lt) * h_.V);
}

void Dirac::set_stream(hipStream_t s) {
  if (!s) s = own_stream_;
  if (s == stream_) return;
  QED3_HIP(hipStreamSynchronize(stream_));
  stream_ = s;
}

void* Dirac::scratch(int slot, size_t bytes) {
  if (slot < 0 || slot >= kScratchSlots) throw std::logic_error("bad scratch slot");
  if (scratch_bytes_[slot] < bytes) {
    QED3_HIP(hipStreamSynchronize(stream_));  // the old buffer may still be in use
    if (scratch_[slot]) QED3_HIP(hipFree(scratch_[slot]));
    scratch_[slot] = nullptr;
    scratch_bytes_[slot] = 0;
    QED3_HIP(hipMalloc(&scratch_[slot], bytes));
    scratch_bytes_[slot] = bytes;
  }
  return scratch_[slot];
}

void Dirac::release_scratch() {
  QED3_HIP(hipStreamSynchronize(stream_));
  for (int i = 0; i < kScratchSlots; i++) {
    if (scratch_[i]) QED3_HIP(hipFree(scratch_[i]));
    scratch_[i] = nullptr;
    scratch_bytes_[i] = 0;
  }
}

CGScalars* Dirac::host_scalars() {
  if (!hsc_) QED3_HIP(hipHostMalloc((void**)&hsc_, sizeof(CGScalars), hipHostMallocDefault));
  return hsc_;
}

Dirac::~Dirac() {
  if (stream_) (void)hipStreamSynchronize(stream_);  // work queued on a caller's stream
  auto f = [](void* p) {
    if (p) (void)hipFree(p);
  };
  for (void* p : scratch_) f(p);
  if (hsc_) (void)hipHostFree(hsc_);
  for (void* p : {(void*)d64_.nbr, (void*)d64_.lnk, (void*)d64_.hop, (void*)d64_.hopd,
                  (void*)d64_.diag, (void*)d64_.kth, (void*)d64_.uS, (void*)d64_.uT,
                  (void*)d32_.nbr, (void*)d32_.lnk, (void*)d32_.hop, (void*)d32_.hopd,
                  (void*)d32_.diag, (void*)d32_.kth, (void*)d32_.uS, (void*)d32_.uT,
                  (void*)uSh_, (void*)uTh_, (void*)uStg_, (void*)uPolyH_, (void*)rotH_, (void*)phi64_,
                  (void*)uPolyP_, (void*)uSiteH_, (void*)uSite10_,
                  (void*)angS_, (void*)angT_, (void*)link_ends_, (void*)partials_,
                  (void*)counter_, (void*)dot_out_, (void*)tmp64_})
    f(p);
  for (auto& kv : halos_) {
    f(kv.second.nbl);
    f(kv.second.off);
    f(kv.second.sites);
  }
  if (own_stream_) (void)hipStreamDestroy(own_stream_);
}

HaloView Dirac::halo(int R) {
  auto it = halos_.find(R);
  if (it == halos_.end()) {
    const int V = h_.V, M = h_.max_nn, nr = (V + R - 1) / R;
    std::vector<unsigned short> nbl(size_t(M) * V);
    std::vector<int> off(nr + 1, 0), sites;
    int hmax = 0;
    for (int r = 0; r < nr; r++) {
      const int x0 = r * R, x1 = std::min(V, x0 + R);
      std::vector<int> ext;
      for (int x = x0; x < x1; x++)
        for (int j = 0; j < M; j++) {
          const int y = h_.nbr[size_t(j) * V + x];
          if (y < x0 || y >= x1) ext.push_back(y);
        }
      std::sort(ext.begin(), ext.end());
      ext.erase(std::unique(ext.begin(), ext.end()), ext.end());
      off[r] = int(sites.size());
      for (int x = x0; x < x1; x++)
        for (int j = 0; j < M; j++) {
          const int y = h_.nbr[size_t(j) * V + x];
          int loc;
          if (y >= x0 && y < x1)
            loc = y - x0;
          else
            loc = R + int(std::lower_bound(ext.begin(), ext.end(), y) - ext.begin());
          if (loc > 65535) throw std::runtime_error("halo index exceeds 16 bits");
          nbl[size_t(j) * V + x] = (unsigned short)loc;
        }
      sites.insert(sites.end(), ext.begin(), ext.end());
      hmax = std::max(hmax, int(ext.size()));
    }
    off[nr] = int(sites.size());
    HaloTables t;
    t.nbl = dupload<unsigned short>(nbl, stream_);
    t.off = dupload<int>(off, stream_);
    t.sites = dupload<int>(sites.empty() ? std::vector<int>{0} : sites, stream_);
    t.hmax = hmax;
    it = halos_.emplace(R, t).first;
  }
  HaloView v;
  v.nbl = it->second.nbl;
  v.halo_off = it->second.off;
  v.halo = it->second.sites;
  v.hmax = it->second.hmax;
  return v;
}

void Dirac::set_gauge(const double* spatial, const double* temporal, bool on_device) {
  const size_t nS = size_t(h_.Lt) * h_.E, nT = size_t(h_.Lt) * h_.V;
  const hipMemcpyKind k = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  QED3_HIP(hipMemcpyAsync(angS_, spatial, nS * sizeof(double), k, stream_));
  QED3_HIP(hipMemcpyAsync(angT_, temporal, nT * sizeof(double), k, stream_));
  const int B = 256;
  hipLaunchKernelGGL(k_phases<double>, dim3((nS + B - 1) / B), dim3(B), 0, stream_, angS_, d64_.uS, nS);
  hipLaunchKernelGGL(k_phases<double>, dim3((nT + B - 1) / B), dim3(B), 0, stream_, angT_, d64_.uT, nT);
  hipLaunchKernelGGL(k_phases<float>, dim3((nS + B - 1) / B), dim3(B), 0, stream_, angS_, d32_.uS, nS);
  hipLaunchKernelGGL(k_phases<float>, dim3((nT + B - 1) / B), dim3(B), 0, stream_, angT_, d32_.uT, nT);
  QED3_HIP(hipGetLastError());
  if (uSh_) fill_half_phases();
  if (uStg_) fill_temporal_gauge();
  QED3_HIP(hipStreamSynchronize(stream_));
  has_gauge_ = true;
}

void Dirac::fill_half_phases() {
  const size_t nS = size_t(h_.Lt) * h_.E, nT = size_t(h_.Lt) * h_.V;
  const int B = 256;
  hipLaunchKernelGGL(k_phases_half, dim3((nS + B - 1) / B), dim3(B), 0, stream_, angS_, uSh_, nS);
  hipLaunchKernelGGL(k_phases_half, dim3((nT + B - 1) / B), dim3(B), 0, stream_, angT_, uTh_, nT);
  QED3_HIP(hipGetLastError());
}

void Dirac::fill_temporal_gauge() {
  const int B = 256;
  hipLaunchKernelGGL(k_tg_sites, dim3((h_.V + B - 1) / B), dim3(B), 0, stream_, angT_, phi64_, rotH_, uPolyH_,
                     h_.V, h_.Lt);
  const size_t nS = size_t(h_.Lt) * h_.E;
  hipLaunchKernelGGL(k_tg_links, dim3((nS + B - 1) / B), dim3(B), 0, stream_, angS_, phi64_,
                     reinterpret_cast<const int4*>(link_ends_), uStg_, h_.E, h_.V, h_.Lt);
  const size_t nq = size_t(h_.Lt) * h_.max_nn * h_.V;
  hipLaunchKernelGGL(k_site_angles, dim3((nq + B - 1) / B), dim3(B), 0, stream_, uStg_, d32_.lnk, uSiteH_, h_.V,
                     h_.E, h_.max_nn, h_.Lt);
  if (uPolyP_)
    hipLaunchKernelGGL(k_angle_pairs, dim3((h_.V + B - 1) / B), dim3(B), 0, stream_, uPolyH_, uPolyP_,
                       size_t(h_.V));
  if (uSite10_) {
    const size_t nw = size_t(h_.Lt) * ((h_.max_nn + 2) / 3) * h_.V;
    hipLaunchKernelGGL(k_site_angles10, dim3((nw + B - 1) / B), dim3(B), 0, stream_, uSiteH_, uSite10_, h_.V,
                       h_.max_nn, h_.Lt);
  }
  QED3_HIP(hipGetLastError());
}

DiracView<float, phasepair_t> Dirac::view_half_tg_pairs() {
  const DiracView<float, phase16_t> tg = view_half_temporal_gauge();
  if (!uPolyP_) {
    uPolyP_ = dalloc<phasepair_t>(size_t(h_.V));
    uSite10_ = dalloc<unsigned>(size_t(h_.Lt) * ((h_.max_nn + 2) / 3) * h_.V);
    if (has_gauge_) fill_temporal_gauge();
  }
  DiracView<float, phasepair_t> v;
  v.nbr = tg.nbr;
  v.lnk = tg.lnk;
  v.hop = tg.hop;
  v.hopd = tg.hopd;
  v.diag = tg.diag;
  v.kth = tg.kth;
  v.uS = nullptr;  // the spatial phases only per site and slot, as packed angles (uSite10)
  v.uT = nullptr;
  v.uPoly = uPolyP_;
  v.uSite10 = uSite10_;
  v.V = tg.V;
  v.E = tg.E;
  v.Lt = tg.Lt;
  v.r = tg.r;
  v.bc = tg.bc;
  return v;
}

DiracView<float, phase16_t> Dirac::view_half_temporal_gauge() {
  DiracView<float, phase16_t> v = view_half_phases();
  if (!uStg_) {
    uStg_ = dalloc<phase16_t>(size_t(h_.Lt) * h_.E);
    uSiteH_ = dalloc<phase16_t>(size_t(h_.Lt) * h_.max_nn * h_.V);
    uPolyH_ = dalloc<phase16_t>(size_t(h_.V));
    rotH_ = dalloc<phase16_t>(size_t(h_.Lt) * h_.V);
    phi64_ = dalloc<double>(size_t(h_.Lt) * h_.V);
    if (has_gauge_) fill_temporal_gauge();
  }
  v.uS = uStg_;
  v.uT = nullptr;
  v.uPoly = uPolyH_;
  v.uSite = uSiteH_;
  return v;
}

DiracView<float, phase16_t> Dirac::view_half_phases() {
  if (!uSh_) {
    uSh_ = dalloc<phase16_t>(size_t(h_.Lt) * h_.E);
    uTh_ = dalloc<phase16_t>(size_t(h_.Lt) * h_.V);
    if (has_gauge_) fill_half_phases();
  }
  const DiracView<float> f = view<float>();
  DiracView<float, phase16_t> v;
  v.nbr = f.nbr;
  v.lnk = f.lnk;
  v.hop = f.hop;
  v.hopd = f.hopd;
  v.diag = f.diag;
  v.kth = f.kth;
  v.uS = uSh_;
  v.uT = uTh_;
  v.V = f.V;
  v.E = f.E;
  v.Lt = f.Lt;
  v.r = f.r;
  v.bc = f.bc;
  return v;
}

template <typename T>
DiracView<T> Dirac::view() const {
  const DeviceTables<T>& d = [this]() -> const DeviceTables<T>& {
    if constexpr (std::is_same<T, double>::value) return d64_;
    else return d32_;
  }();
  DiracView<T> v;
  v.nbr = d.nbr;
  v.lnk = d.lnk;
  v.hop = d.hop;
  v.hopd = d.hopd;
  v.diag = d.diag;
  v.kth = d.kth;
  v.uS = d.uS;
  v.uT = d.uT;
  v.V = h_.V;
  v.E = h_.E;
  v.Lt = h_.Lt;
  v.r = T(h_.r);
  v.bc = bc_ == 1 ? T(1) : T(-1);
  return v;
}

template <typename T>
StencilTile Dirac::tile(bool dag) const {
  return tile_[std::is_same<T, float>::value ? 1 : 0][dag ? 1 : 0];
}

template <typename T>
void Dirac::apply(bool dag, Spin<T>* out, const Spin<T>* in, bool sync) {
  const DiracView<T> v = view<T>();
  EpiStore<T> e{out};
  if (dag)
    launch_stencil<T, true>(v, h_.max_nn, in, e, tile<T>(true), reducer(), stream_, halo_for<T>(true));
  else
    launch_stencil<T, false>(v, h_.max_nn, in, e, tile<T>(false), reducer(), stream_, halo_for<T>(false));
  QED3_HIP(hipGetLastError());
  if (sync) QED3_HIP(hipStreamSynchronize(stream_));
}

void Dirac::apply_half(bool dag, SpinH* out, const SpinH* in, bool sync) {
  const DiracView<float, phase16_t> v = view_half_phases();
  const EpiStoreH e{out};
  if (dag)
    launch_stencil_src<float, true>(v, h_.max_nn, SrcHalf{in}, e, tile<float>(true), reducer(), stream_,
                                    halo_for<float>(true));
  else
    launch_stencil_src<float, false>(v, h_.max_nn, SrcHalf{in}, e, tile<float>(false), reducer(), stream_,
                                     halo_for<float>(false));
  QED3_HIP(hipGetLastError());
  if (sync) QED3_HIP(hipStreamSynchronize(stream_));
}

template <typename T>
void Dirac::apply_dagd(Spin<T>* out, const Spin<T>* in, Spin<T>* tmp, bool sync) {
  apply<T>(false, tmp, in, false);
  apply<T>(true, out, tmp, sync);
}

template <typename T>
void Dirac::apply_ddag(Spin<T>* out, const Spin<T>* in, Spin<T>* tmp, bool sync) {
  apply<T>(true, tmp, in, false);
  apply<T>(false, out, tmp, sync);
}

void Dirac::dot(const Spin<double>* x, const Spin<double>* y, double* re, double* im) {
  const size_t n = points();
  const int B = 256;
  const size_t nb = std::min<size_t>((n + B - 1) / B, 4096);
  const unsigned nbl = unsigned(std::max<size_t>(nb, 1));
  hipLaunchKernelGGL(k_dot, dim3(nbl), dim3(B), 0, stream_, x, y, n, reducer());
  hipLaunchKernelGGL(k_finalize<FinDot>, dim3(1), dim3(256), 0, stream_, reducer().partials, nbl, FinDot{dot_out_});
  QED3_HIP(hipGetLastError());
  double2 h;
  QED3_HIP(hipMemcpyAsync(&h, dot_out_, sizeof(h), hipMemcpyDeviceToHost, stream_));
  QED3_HIP(hipStreamSynchronize(stream_));
  if (re) *re = h.x;
  if (im) *im = h.y;
}

void Dirac::force(const Spin<double>* eta, double* fs, double* ft) {
  if (!tmp64_) tmp64_ = dalloc<Spin<double>>(points());
  apply<double>(false, tmp64_, eta, false);  // precalc: w = D eta
  const size_t n = size_t(h_.Lt) * (h_.E + h_.V);
  const int B = 256;
  hipLaunchKernelGGL(k_force, dim3((n + B - 1) / B), dim3(B), 0, stream_, view<double>(),
                     reinterpret_cast<const int4*>(link_ends_), eta, tmp64_, fs, ft);
  QED3_HIP(hipGetLastError());
  QED3_HIP(hipStreamSynchronize(stream_));
}

void Dirac::force_bilinear(const Spin<double>* u1, const Spin<double>* v1, const Spin<double>* u2,
                           const Spin<double>* v2, double c, bool init, double* fs, double* ft) {
  if ((u2 == nullptr) != (v2 == nullptr)) throw std::invalid_argument("force_bilinear: second pair incomplete");
  const size_t n = size_t(h_.Lt) * (h_.E + h_.V);
  const int B = 256;
  hipLaunchKernelGGL(k_force_acc, dim3((n + B - 1) / B), dim3(B), 0, stream_, view<double>(),
                     reinterpret_cast<const int4*>(link_ends_), u1, v1, u2, v2, c, init ? 1 : 0, fs, ft);
  QED3_HIP(hipGetLastError());
}

double Dirac::time_apply(bool dag, int prec, int warmup, int reps) {
  const size_t n = points();
  hipEvent_t e0, e1;
  QED3_HIP(hipEventCreate(&e0));
  QED3_HIP(hipEventCreate(&e1));
  float ms = 0;
  auto run = [&](auto tag) {
    using T = decltype(tag);
    Spin<T>* a = dalloc<Spin<T>>(n);
    Spin<T>* b = dalloc<Spin<T>>(n);
    QED3_HIP(hipMemsetAsync(a, 0, n * sizeof(Spin<T>), stream_));
    for (int i = 0; i < warmup; i++) apply<T>(dag, b, a, false);
    QED3_HIP(hipEventRecord(e0, stream_));
    for (int i = 0; i < reps; i++) apply<T>(dag, b, a, false);
    QED3_HIP(hipEventRecord(e1, stream_));
    QED3_HIP(hipEventSynchronize(e1));
    QED3_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipFree(a);
    (void)hipFree(b);
  };
  if (prec == 2) {
    SpinH* a = dalloc<SpinH>(n);
    SpinH* b = dalloc<SpinH>(n);
    QED3_HIP(hipMemsetAsync(a, 0, n * sizeof(SpinH), stream_));
    for (int i = 0; i < warmup; i++) apply_half(dag, b, a, false);
    QED3_HIP(hipEventRecord(e0, stream_));
    for (int i = 0; i < reps; i++) apply_half(dag, b, a, false);
    QED3_HIP(hipEventRecord(e1, stream_));
    QED3_HIP(hipEventSynchronize(e1));
    QED3_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipFree(a);
    (void)hipFree(b);
  } else if (prec == 1) {
    run(float{});
  } else {
    run(double{});
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return double(ms) / std::max(reps, 1);
}

// Project-written autotuner: times each candidate (block size, time-tile) of the
// stencil on this lattice shape with HIP events and keeps the fastest.
template <typename T>
void Dirac::tune_one(bool dag) {
  const size_t n = points();
  Spin<T>* a = dalloc<Spin<T>>(n);
  Spin<T>* b = dalloc<Spin<T>>(n);
  QED3_HIP(hipMemsetAsync(a, 0, n * sizeof(Spin<T>), stream_));
  hipEvent_t e0, e1;
  QED3_HIP(hipEventCreate(&e0));
  QED3_HIP(hipEventCreate(&e1));
  StencilTile best = tile<T>(dag);
  float best_ms = 1e30f;
  const int f = std::is_same<T, float>::value ? 1 : 0;
  for (int lds : {0, 1})
  for (int remap : {1, 0})
  for (int block : {64, 128, 256}) {
    for (int ts : {1, 2, 4, 8, 16, 32, 64}) {
      if (ts > h_.Lt) continue;
      tile_[f][dag] = StencilTile{block, ts, remap, lds};
      for (int i = 0; i < 2; i++) apply<T>(dag, b, a, false);
      QED3_HIP(hipEventRecord(e0, stream_));
      const int reps = 5;
      for (int i = 0; i < reps; i++) apply<T>(dag, b, a, false);
      QED3_HIP(hipEventRecord(e1, stream_));
      QED3_HIP(hipEventSynchronize(e1));
      float ms;
      QED3_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best_ms) {
        best_ms = ms;
        best = tile_[f][dag];
      }
    }
  }
  tile_[f][dag] = best;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
}

void Dirac::autotune() {
  if (tuned_) return;
  tune_one<double>(false);
  tune_one<double>(true);
  tune_one<float>(false);
  tune_one<float>(true);
  tuned_ = true;
}

template DiracView<double> Dirac::view<double>() const;
template DiracView<float> Dirac::view<float>() const;
template StencilTile Dirac::tile<double>(bool) const;
template StencilTile Dirac::tile<float>(bool) const;
template void Dirac::apply<double>(bool, Spin<double>*, const Spin<double>*, bool);
template void Dirac::apply<float>(bool, Spin<float>*, const Spin<float>*, bool);
template void Dirac::apply_dagd<double>(Spin<double>*, const Spin<double>*, Spin<double>*, bool);
template void Dirac::apply_dagd<float>(Spin<float>*, const Spin<float>*, Spin<float>*, bool);
template void Dirac::apply_ddag<double>(Spin<double>*, const Spin<double>*, Spin<double>*, bool);
template void Dirac::apply_ddag<float>(Spin<float>*, const Spin<float>*, Spin<float>*, bool);

}  // namespace qed3
