// Time-exact Horner preconditioner: tables, the FP64 apply used by the tests, and the
// training loss and gradient on the GPU (see horner.hpp for the network).
#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "horner.hpp"
#include "horner_kernels.hpp"

namespace qed3 {

using cd = std::complex<double>;

void HornerParams::pack(cd* p) const {
  for (int j = 0; j < layers; j++) p[j] = c[j];
  p[layers] = A[0];
  p[layers + 1] = A[1];
  p[layers + 2] = B[0];
  p[layers + 3] = B[1];
}

void HornerParams::unpack(const cd* p) {
  c.assign(p, p + layers);
  A[0] = p[layers];
  A[1] = p[layers + 1];
  B[0] = p[layers + 2];
  B[1] = p[layers + 3];
}

namespace {



// Row coefficients of S_th for spin rows 0 and 1 from the compressed hop matrix
// H = [[-r(C - iS), a - ib], [a + ib, -r(C + iS)]] (D: hop, D^dagger: hopd):
//   mode 0 (K):       row a = (A_a H_aa, B_a H_a,abar)
//   mode 1 (K^dag):   row a = (conj(A_a) Hd_aa, conj(B_abar) Hd_a,abar)
//   mode 2 (unit):    row a = (H_aa, H_a,abar), for the gradient
__global__ void k_horner_tables(const float4* __restrict__ hop, const float4* __restrict__ hopd, float r,
                                float4 AB, float4 ABc, float4* out, HopRow16* out16, size_t ns) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  const float2 A0 = {AB.x, AB.y}, A1 = {AB.z, AB.w};
  const float2 B0 = {ABc.x, ABc.y}, B1 = {ABc.z, ABc.w};
  for (int mode = 0; mode < 3; mode++) {
    const float4 h = mode == 1 ? hopd[i] : hop[i];
    const float2 H00 = {-r * h.x, r * h.y}, H01 = {h.z, -h.w}, H10 = {h.z, h.w}, H11 = {-r * h.x, -r * h.y};
    float2 s0 = H00, f0 = H01, s1 = H11, f1 = H10;
    if (mode == 0) {
      s0 = cmul<float>(A0, H00);
      f0 = cmul<float>(B0, H01);
      s1 = cmul<float>(A1, H11);
      f1 = cmul<float>(B1, H10);
    } else if (mode == 1) {
      s0 = cmul<float>({A0.x, -A0.y}, H00);
      f0 = cmul<float>({B1.x, -B1.y}, H01);
      s1 = cmul<float>({A1.x, -A1.y}, H11);
      f1 = cmul<float>({B0.x, -B0.y}, H10);
    }
    out[(size_t(mode) * 2 + 0) * ns + i] = {s0.x, s0.y, f0.x, f0.y};
    out[(size_t(mode) * 2 + 1) * ns + i] = {s1.x, s1.y, f1.x, f1.y};
    if (mode < 2) {
      HopRow16 h0, h1;
      h0.s = half2_t{(_Float16)s0.x, (_Float16)s0.y};
      h0.f = half2_t{(_Float16)f0.x, (_Float16)f0.y};
      h1.s = half2_t{(_Float16)s1.x, (_Float16)s1.y};
      h1.f = half2_t{(_Float16)f1.x, (_Float16)f1.y};
      out16[(size_t(mode) * 2 + 0) * ns + i] = h0;
      out16[(size_t(mode) * 2 + 1) * ns + i] = h1;
    }
  }
}

__global__ void k_to_f32(const Spin<double>* __restrict__ in, Spin<float>* out, size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const Spin<double> v = in[i];
    Spin<float> f;
    f.a = {float(v.a.x), float(v.a.y)};
    f.b = {float(v.b.x), float(v.b.y)};
    out[i] = f;
  }
}

__global__ void k_to_f64(const Spin<float>* __restrict__ in, Spin<double>* out, size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const Spin<float> f = in[i];
    Spin<double> v;
    v.a = {f.a.x, f.a.y};
    v.b = {f.b.x, f.b.y};
    out[i] = v;
  }
}

// |e|^2 of an FP32 field (n points)
__global__ __launch_bounds__(256) void k_norm_f32(const Spin<float>* __restrict__ e, size_t n, GridReduce gr) {
  double2 acc = {0.0, 0.0};
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    acc.x += norm2(e[i]);
  store_partial(acc, gr.partials, blockIdx.x);
}

struct FinStore {
  double* out;
  __device__ bool skip() const { return false; }
  __device__ void operator()(double2 t) const { *out = t.x; }
};

// lambda = w (g - e) with w = wscale / |e|^2; accumulates w |g - e|^2
__global__ __launch_bounds__(256) void k_loss_f32(const Spin<float>* __restrict__ g, const Spin<float>* __restrict__ e,
                                                  Spin<float>* lam, const double* enorm2, double wscale, size_t n,
                                                  GridReduce gr) {
  const double w = wscale / *enorm2;
  double2 acc = {0.0, 0.0};
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const Spin<float> a = g[i], b = e[i];
    const double d0 = double(a.a.x) - b.a.x, d1 = double(a.a.y) - b.a.y, d2 = double(a.b.x) - b.b.x,
                 d3 = double(a.b.y) - b.b.y;
    acc.x += w * (d0 * d0 + d1 * d1 + d2 * d2 + d3 * d3);
    Spin<float> l;
    l.a = {float(w * d0), float(w * d1)};
    l.b = {float(w * d2), float(w * d3)};
    lam[i] = l;
  }
  store_partial(acc, gr.partials, blockIdx.x);
}

// Gradient partials of one Horner step t: per block of (site range, spin row a),
//   q[0..1] = sum conj(v_a) mu_a                      (-> the step's c)
//   q[2..3] = sum conj(P_a) mu_a, P_a = sum_j u_j H_aa g_a(nb_j)        (-> A_a, times -1)
//   q[4..5] = sum conj(Q_a) mu_a, Q_a = sum_j u_j H_a,abar g_abar(nb_j)  (-> B_a, times -1)
// with g the previous step's output (has_g) and mu = dloss / d conj(y_t).
template <int NS>
__global__ __launch_bounds__(256) void k_horner_grad(HornerView hv, DiracView<float> d, const Spin<float>* __restrict__ v,
                                                     const Spin<float>* __restrict__ mu,
                                                     const Spin<float>* __restrict__ g, int has_g, TileMap tm,
                                                     double* part) {
  int xr, a;
  tile_of(tm, blockIdx.x, xr, a);
  const int V = hv.V, Lt = hv.Lt;
  const int x = xr < tm.nxr ? xr * int(blockDim.x) + int(threadIdx.x) : V;
  double q[6] = {0, 0, 0, 0, 0, 0};
  if (x < V) {
    int nb[NS], lk[NS];
    float4 hr[NS];
    const float4* rows = hv.hrow + (size_t(2 * 2 + a) * NS) * V;  // mode 2: unit coefficients
#pragma unroll
    for (int j = 0; j < NS; j++) {
      nb[j] = d.nbr[j * V + x];
      lk[j] = d.lnk[j * V + x];
      hr[j] = rows[size_t(j) * V + x];
    }
    for (int s = 0; s < Lt; s++) {
      const size_t i = size_t(s) * V + x;
      const float2 m = hk::comp(mu[i], a);
      const float2 vv = hk::comp(v[i], a);
      q[0] += double(vv.x) * m.x + double(vv.y) * m.y;
      q[1] += double(vv.x) * m.y - double(vv.y) * m.x;
      if (has_g) {
        float2 P = {0.f, 0.f}, Q = {0.f, 0.f};
        const float2* uS = d.uS + size_t(s) * d.E;
#pragma unroll
        for (int j = 0; j < NS; j++) {
          const int l = lk[j];
          float2 u = ldu(uS + (l >= 0 ? l : ~l));
          if (l < 0) u.y = -u.y;
          const size_t iy = size_t(s) * V + nb[j];
          const Spin<float> gy = g[iy];
          const float2 ga = hk::comp(gy, a), gb = hk::comp(gy, 1 - a);
          P = cadd<float>(P, cmul<float>(u, cmul<float>({hr[j].x, hr[j].y}, ga)));
          Q = cadd<float>(Q, cmul<float>(u, cmul<float>({hr[j].z, hr[j].w}, gb)));
        }
        q[2] += double(P.x) * m.x + double(P.y) * m.y;
        q[3] += double(P.x) * m.y - double(P.y) * m.x;
        q[4] += double(Q.x) * m.x + double(Q.y) * m.y;
        q[5] += double(Q.x) * m.y - double(Q.y) * m.x;
      }
    }
  }
  for (int k = 0; k < 6; k++) {
    const double2 b = block_sum2({q[k], 0.0});
    if (threadIdx.x == 0) part[size_t(blockIdx.x) * 6 + k] = b.x;
  }
}

TileMap spin_tiles(int V) {
  TileMap tm;
  tm.nxr = (V + 255) / 256;
  tm.nxr_pad = (tm.nxr + 7) / 8 * 8;
  tm.nst = 2;
  tm.remap = 1;
  return tm;
}

}  // namespace

Horner::Horner(Dirac& D, int layers) : D_(D) {
  if (layers < 1 || layers > 16) throw std::invalid_argument("layers must be in [1,16]");
  if (std::abs(D.tables().r - 1.0) > 0) throw std::invalid_argument("the time-exact preconditioner needs r = 1");
  p_.layers = layers;
  p_.c.assign(layers, cd(1.0, 0.0));
  const int V = D.V(), NS = D.tables().max_nn;
  QED3_HIP(hipMalloc((void**)&hrow_, size_t(6) * NS * V * sizeof(float4)));
  QED3_HIP(hipMalloc((void**)&hrow16_, size_t(4) * NS * V * sizeof(HopRow16)));
  std::vector<float> invc(V), kt(V);
  for (int x = 0; x < V; x++) {
    invc[x] = float(1.0 / D.tables().diag[x]);
    kt[x] = float(D.tables().kappa_t[x]);  // r = 1: the temporal hop weight is kappa_t
  }
  QED3_HIP(hipMalloc((void**)&invc_, V * sizeof(float)));
  QED3_HIP(hipMalloc((void**)&kt_, V * sizeof(float)));
  QED3_HIP(hipMemcpyAsync(invc_, invc.data(), V * sizeof(float), hipMemcpyHostToDevice, D.stream()));
  QED3_HIP(hipMemcpyAsync(kt_, kt.data(), V * sizeof(float), hipMemcpyHostToDevice, D.stream()));
  upload_tables();
}

Horner::~Horner() {
  for (void* p : {(void*)hrow_, (void*)hrow16_, (void*)invc_, (void*)kt_, (void*)red_})
    if (p) (void)hipFree(p);
  for (void* p : work_)
    if (p) (void)hipFree(p);
}

void Horner::set_params(const HornerParams& p) {
  if (p.layers != p_.layers || int(p.c.size()) != p.layers) throw std::invalid_argument("parameter count");
  p_ = p;
  upload_tables();
}

void Horner::upload_tables() {
  const DiracView<float> f = D_.view<float>();
  const size_t ns = size_t(D_.tables().max_nn) * D_.V();
  const float4 AB = {float(p_.A[0].real()), float(p_.A[0].imag()), float(p_.A[1].real()), float(p_.A[1].imag())};
  const float4 Bc = {float(p_.B[0].real()), float(p_.B[0].imag()), float(p_.B[1].real()), float(p_.B[1].imag())};
  hipLaunchKernelGGL(k_horner_tables, dim3(unsigned((ns + 255) / 256)), dim3(256), 0, D_.stream(), f.hop, f.hopd,
                     float(D_.tables().r), AB, Bc, hrow_, hrow16_, ns);
  QED3_HIP(hipGetLastError());
  QED3_HIP(hipStreamSynchronize(D_.stream()));
}

HornerView Horner::view() const {
  HornerView v;
  v.hrow = hrow_;
  v.hrow16 = hrow16_;
  v.invc = invc_;
  v.kt = kt_;
  v.V = D_.V();
  v.Lt = D_.Lt();
  v.nslot = D_.tables().max_nn;
  return v;
}

void* Horner::work(int slot, size_t bytes) {
  if (slot < 0 || slot >= kWorkSlots) throw std::logic_error("bad work slot");
  if (work_bytes_[slot] < bytes) {
    QED3_HIP(hipStreamSynchronize(D_.stream()));
    if (work_[slot]) QED3_HIP(hipFree(work_[slot]));
    work_[slot] = nullptr;
    work_bytes_[slot] = 0;
    QED3_HIP(hipMalloc(&work_[slot], bytes));
    work_bytes_[slot] = bytes;
  }
  return work_[slot];
}

namespace {
// One network application on FP32 fields in the original gauge: out = K in (or
// K^dagger in); b0, b1 ping-pong buffers.  Every layer output is kept when keep != nullptr
// (keep[t] = output of step t).
void run_layers(Horner& K, bool adj, const Spin<float>* in, Spin<float>* const* bufs, Spin<float>* out,
                Spin<float>* const* keep) {
  Dirac& D = K.dirac();
  const HornerView hv = K.view();
  const DiracView<float> tab = D.view<float>();
  const HaloView halo = D.halo(256);
  const int L = K.layers();
  const hk::SrcF<Spin<float>> src{in};
  const GridReduce gr = D.reducer();
  Spin<float>* prev = nullptr;
  for (int t = 0; t < L; t++) {
    const cd c = K.params().c[L - 1 - t];
    const float2 cj = {float(c.real()), float(adj ? -c.imag() : c.imag())};
    Spin<float>* o = keep ? keep[t] : (t == L - 1 ? out : bufs[t & 1]);
    if (t == 0) {
      if (adj)
        hk::launch_horner_layer<true, true, false>(hv, tab, halo, src, (const Spin<float>*)nullptr, o, cj,
                                                   hk::EpiNone{}, gr, D.stream());
      else
        hk::launch_horner_layer<false, true, false>(hv, tab, halo, src, (const Spin<float>*)nullptr, o, cj,
                                                    hk::EpiNone{}, gr, D.stream());
    } else {
      if (adj)
        hk::launch_horner_layer<true, false, false>(hv, tab, halo, src, (const Spin<float>*)prev, o, cj,
                                                    hk::EpiNone{}, gr, D.stream());
      else
        hk::launch_horner_layer<false, false, false>(hv, tab, halo, src, (const Spin<float>*)prev, o, cj,
                                                     hk::EpiNone{}, gr, D.stream());
    }
    prev = o;
  }
  QED3_HIP(hipGetLastError());
}

unsigned stream_grid(size_t n) { return unsigned(std::min<size_t>((n + 255) / 256, 8192)); }
}  // namespace

void Horner::apply(bool adj, Spin<double>* out, const Spin<double>* in) {
  const size_t n = D_.points();
  const size_t fb = n * sizeof(Spin<float>);
  Spin<float>* pin = static_cast<Spin<float>*>(work(0, fb));
  Spin<float>* b0 = static_cast<Spin<float>*>(work(1, fb));
  Spin<float>* b1 = static_cast<Spin<float>*>(work(2, fb));
  Spin<float>* po = static_cast<Spin<float>*>(work(3, fb));
  hipStream_t st = D_.stream();
  hipLaunchKernelGGL(k_to_f32, dim3(stream_grid(n)), dim3(256), 0, st, in, pin, n);
  Spin<float>* bufs[2] = {b0, b1};
  run_layers(*this, adj, pin, bufs, po, nullptr);
  hipLaunchKernelGGL(k_to_f64, dim3(stream_grid(n)), dim3(256), 0, st, po, out, n);
  QED3_HIP(hipGetLastError());
  QED3_HIP(hipStreamSynchronize(st));
}

double Horner::loss_grad(const Spin<double>* v, const Spin<double>* e, int batch, std::vector<cd>* grad) {
  if (batch < 1) throw std::invalid_argument("batch must be positive");
  const int L = p_.layers, V = D_.V(), NS = D_.tables().max_nn;
  const size_t n = D_.points();
  const size_t fb = n * sizeof(Spin<float>);
  hipStream_t st = D_.stream();
  if (L > kWorkSlots - 6) throw std::invalid_argument("too many layers for the training buffers");
  Spin<float>* pv = static_cast<Spin<float>*>(work(0, fb));
  Spin<float>* pe = static_cast<Spin<float>*>(work(1, fb));
  Spin<float>* lam = static_cast<Spin<float>*>(work(2, fb));
  Spin<float>* mu[2] = {static_cast<Spin<float>*>(work(3, fb)), static_cast<Spin<float>*>(work(4, fb))};
  std::vector<Spin<float>*> G(L);
  for (int t = 0; t < L; t++) G[t] = static_cast<Spin<float>*>(work(6 + t, fb));
  double* dscal = static_cast<double*>(work(5, 2 * sizeof(double)));
  const TileMap tm = spin_tiles(V);
  const unsigned nblk = unsigned(tm.nxr_pad) * 2u;
  if (red_n_ < size_t(nblk) * 6) {
    if (red_) QED3_HIP(hipFree(red_));
    QED3_HIP(hipMalloc((void**)&red_, size_t(nblk) * 6 * sizeof(double)));
    red_n_ = size_t(nblk) * 6;
  }
  const GridReduce gr = D_.reducer();
  const HornerView hv = view();
  const DiracView<float> tab = D_.view<float>();
  const HaloView halo = D_.halo(256);
  std::vector<cd> g(p_.count(), cd(0, 0));
  std::vector<double> part(size_t(nblk) * 6);
  double loss = 0;
  const unsigned nbk = stream_grid(n);
  for (int b = 0; b < batch; b++) {
    hipLaunchKernelGGL(k_to_f32, dim3(nbk), dim3(256), 0, st, v + size_t(b) * n, pv, n);
    hipLaunchKernelGGL(k_to_f32, dim3(nbk), dim3(256), 0, st, e + size_t(b) * n, pe, n);
    run_layers(*this, false, pv, nullptr, nullptr, G.data());
    hipLaunchKernelGGL(k_norm_f32, dim3(nbk), dim3(256), 0, st, pe, n, gr);
    hipLaunchKernelGGL(k_finalize<FinStore>, dim3(1), dim3(256), 0, st, gr.partials, nbk, FinStore{dscal});
    hipLaunchKernelGGL(k_loss_f32, dim3(nbk), dim3(256), 0, st, G[L - 1], pe, lam, dscal, 1.0 / batch, n, gr);
    hipLaunchKernelGGL(k_finalize<FinStore>, dim3(1), dim3(256), 0, st, gr.partials, nbk, FinStore{dscal + 1});
    QED3_HIP(hipGetLastError());
    double hl[2];
    QED3_HIP(hipMemcpyAsync(hl, dscal, sizeof(hl), hipMemcpyDeviceToHost, st));
    // backward: mu_{L-1} = T^-dag lambda; mu_{t-1} = T^-dag (S_th^dag mu_t)
    const hk::SrcF<Spin<float>> srcl{lam};
    for (int t = L - 1; t >= 0; t--) {
      Spin<float>* mo = mu[t & 1];
      if (t == L - 1)
        hk::launch_horner_layer<true, true, false>(hv, tab, halo, srcl, (const Spin<float>*)nullptr, mo,
                                                   float2{1.f, 0.f}, hk::EpiNone{}, gr, st);
      else
        hk::launch_horner_layer<true, false, false>(hv, tab, halo, srcl, (const Spin<float>*)mu[(t + 1) & 1], mo,
                                                    float2{0.f, 0.f}, hk::EpiNone{}, gr, st);
      const Spin<float>* gp = t > 0 ? G[t - 1] : nullptr;
      if (NS == 6)
        hipLaunchKernelGGL(k_horner_grad<6>, dim3(nblk), dim3(256), 0, st, hv, tab, pv, mo, gp, int(t > 0), tm, red_);
      else
        hipLaunchKernelGGL(k_horner_grad<3>, dim3(nblk), dim3(256), 0, st, hv, tab, pv, mo, gp, int(t > 0), tm, red_);
      QED3_HIP(hipGetLastError());
      QED3_HIP(hipMemcpyAsync(part.data(), red_, part.size() * sizeof(double), hipMemcpyDeviceToHost, st));
      QED3_HIP(hipStreamSynchronize(st));
      double q0[6] = {0, 0, 0, 0, 0, 0}, q1[6] = {0, 0, 0, 0, 0, 0};
      for (unsigned k = 0; k < nblk; k++) {
        // block k's spin row, decoded as tile_of does
        int a;
        if (tm.remap) {
          const int per = tm.nxr_pad >> 3, j = int(k) >> 3;
          a = j / per;
        } else {
          a = int(k) / tm.nxr_pad;
        }
        double* q = a == 0 ? q0 : q1;
        for (int m = 0; m < 6; m++) q[m] += part[size_t(k) * 6 + m];
      }
      // step t used c[L-1-t]; gradients are 2 d loss / d conj(param)
      g[L - 1 - t] += 2.0 * cd(q0[0] + q1[0], q0[1] + q1[1]);
      if (t > 0) {
        g[L] += -2.0 * cd(q0[2], q0[3]);
        g[L + 1] += -2.0 * cd(q1[2], q1[3]);
        g[L + 2] += -2.0 * cd(q0[4], q0[5]);
        g[L + 3] += -2.0 * cd(q1[4], q1[5]);
      }
    }
    loss += hl[1];
  }
  if (grad) *grad = g;
  return loss;
}

std::vector<double> Horner::train(const Spin<double>* v, const Spin<double>* e, int npairs, int batch, int epochs,
                                  double lr) {
  if (npairs < 1 || batch < 1 || epochs < 1) throw std::invalid_argument("bad training sizes");
  const size_t n = D_.points();
  const int np = p_.count();
  std::vector<cd> p(np), m(np, 0.0), s2(np, 0.0), g;
  p_.pack(p.data());
  std::vector<double> hist;
  const double b1 = 0.9, b2 = 0.999, eps = 1e-12;
  int step = 0;
  for (int ep = 0; ep < epochs; ep++) {
    double tot = 0;
    int cnt = 0;
    for (int k0 = 0; k0 < npairs; k0 += batch) {
      const int bs = std::min(batch, npairs - k0);
      const double l = loss_grad(v + size_t(k0) * n, e + size_t(k0) * n, bs, &g);
      tot += l * bs;
      cnt += bs;
      step++;
      for (int i = 0; i < np; i++) {
        m[i] = b1 * m[i] + (1 - b1) * g[i];
        const double gr = g[i].real(), gi = g[i].imag();
        s2[i] = cd(b2 * s2[i].real() + (1 - b2) * gr * gr, b2 * s2[i].imag() + (1 - b2) * gi * gi);
        const cd mh = m[i] / (1 - std::pow(b1, step));
        const double vr = s2[i].real() / (1 - std::pow(b2, step)), vi = s2[i].imag() / (1 - std::pow(b2, step));
        p[i] -= cd(lr * mh.real() / (std::sqrt(vr) + eps), lr * mh.imag() / (std::sqrt(vi) + eps));
      }
      HornerParams q = p_;
      q.unpack(p.data());
      set_params(q);
    }
    hist.push_back(tot / std::max(cnt, 1));
  }
  return hist;
}

}  // namespace qed3
