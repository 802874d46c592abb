// Device gauge action and force (see gauge_dev.hpp).
#include <algorithm>
#include <stdexcept>

#include "device_common.hpp"
#include "gauge_dev.hpp"

namespace qed3 {
namespace {

__device__ __forceinline__ double dgrad(bool compact, double b, double P) {
  return compact ? b * sin(P) : b * P;
}
__device__ __forceinline__ double dact(bool compact, double b, double P) {
  return compact ? b * (1.0 - cos(P)) : 0.5 * b * P * P;
}

// temporal plaquette of link l = (x, y) at slice s: u_s(l) + t_s(y) - u_{s+1}(l) - t_s(x)
__device__ __forceinline__ double tplaq(const double* u, const double* t, int s, int l, int Lt, int E,
                                        int V, int2 xy) {
  const int s1 = (s + 1 == Lt) ? 0 : s + 1;
  return u[size_t(s) * E + l] + t[size_t(s) * V + xy.y] - u[size_t(s1) * E + l] - t[size_t(s) * V + xy.x];
}

__global__ __launch_bounds__(256) void k_plaq_grads(const double* __restrict__ u, const double* __restrict__ t,
                                                    const int* __restrict__ fe_off, const int* __restrict__ fe_link,
                                                    const int* __restrict__ fe_sign, const int2* __restrict__ ends,
                                                    const double* __restrict__ bs, const double* __restrict__ bt,
                                                    bool compact, int Lt, int V, int E, int F,
                                                    double* gf, double* gtp) {
  const size_t nf = size_t(Lt) * F, ne = size_t(Lt) * E;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nf + ne; i += size_t(gridDim.x) * blockDim.x) {
    if (i < nf) {
      const int s = int(i / F), f = int(i % F);
      const double* us = u + size_t(s) * E;
      double P = 0;
      for (int k = fe_off[f]; k < fe_off[f + 1]; k++) P += fe_sign[k] * us[fe_link[k]];
      gf[i] = dgrad(compact, bs[f], P);
    } else {
      const size_t j = i - nf;
      const int s = int(j / E), l = int(j % E);
      gtp[j] = dgrad(compact, bt[l], tplaq(u, t, s, l, Lt, E, V, ends[l]));
    }
  }
}

__global__ __launch_bounds__(256) void k_force_links(const double* __restrict__ gf, const double* __restrict__ gtp,
                                                     const int* __restrict__ lf_face, const int* __restrict__ lf_sign,
                                                     const int* __restrict__ st_off, const int* __restrict__ st_link,
                                                     const int* __restrict__ st_sign, int Lt, int V, int E, int F,
                                                     double* fs, double* ft) {
  const size_t ns = size_t(Lt) * E, nt = size_t(Lt) * V;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < ns + nt; i += size_t(gridDim.x) * blockDim.x) {
    if (i < ns) {
      const int s = int(i / E), l = int(i % E);
      const int sm = (s == 0) ? Lt - 1 : s - 1;
      double v = 0;
      for (int k = 0; k < 2; k++) {
        const int f = lf_face[2 * l + k];
        if (f >= 0) v += lf_sign[2 * l + k] * gf[size_t(s) * F + f];
      }
      v += gtp[size_t(s) * E + l] - gtp[size_t(sm) * E + l];
      fs[i] = v;
    } else {
      const size_t j = i - ns;
      const int s = int(j / V), x = int(j % V);
      double v = 0;
      for (int k = st_off[x]; k < st_off[x + 1]; k++) v += st_sign[k] * gtp[size_t(s) * E + st_link[k]];
      ft[j] = v;
    }
  }
}

__global__ __launch_bounds__(256) void k_action_parts(const double* __restrict__ u, const double* __restrict__ t,
                                                      const int* __restrict__ fe_off, const int* __restrict__ fe_link,
                                                      const int* __restrict__ fe_sign, const int2* __restrict__ ends,
                                                      const double* __restrict__ bs, const double* __restrict__ bt,
                                                      bool compact, int Lt, int V, int E, int F, double2* partials) {
  const size_t nf = size_t(Lt) * F, ne = size_t(Lt) * E;
  double2 acc = {0.0, 0.0};
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nf + ne; i += size_t(gridDim.x) * blockDim.x) {
    if (i < nf) {
      const int s = int(i / F), f = int(i % F);
      const double* us = u + size_t(s) * E;
      double P = 0;
      for (int k = fe_off[f]; k < fe_off[f + 1]; k++) P += fe_sign[k] * us[fe_link[k]];
      acc.x += dact(compact, bs[f], P);
    } else {
      const size_t j = i - nf;
      const int s = int(j / E), l = int(j % E);
      acc.x += dact(compact, bt[l], tplaq(u, t, s, l, Lt, E, V, ends[l]));
    }
  }
  store_partial(acc, partials, blockIdx.x);
}

__global__ __launch_bounds__(256) void k_norm2_parts(const double* __restrict__ a, size_t n, double2* partials) {
  double2 acc = {0.0, 0.0};
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    acc.x += a[i] * a[i];
  store_partial(acc, partials, blockIdx.x);
}

__global__ __launch_bounds__(256) void k_daxpy(double* y, const double* __restrict__ x, double c, size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    y[i] += c * x[i];
}

template <typename T>
T* upload(const std::vector<T>& v) {
  T* p = nullptr;
  QED3_HIP(hipMalloc(&p, std::max<size_t>(v.size(), 1) * sizeof(T)));
  if (!v.empty()) QED3_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return p;
}

unsigned blocks_for(size_t n) { return unsigned(std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 4096))); }

}  // namespace

DeviceGauge::DeviceGauge(const HopTables& h, const std::vector<double>& beta_s,
                         const std::vector<double>& beta_t, const std::vector<int>& fe_off,
                         const std::vector<int>& fe_link, const std::vector<int>& fe_sign, bool compact)
    : Lt_(h.Lt), V_(h.V), E_(h.E), F_(int(fe_off.size()) - 1), compact_(compact) {
  std::vector<int> lf_face(2 * size_t(E_), -1), lf_sign(2 * size_t(E_), 0);
  for (int f = 0; f < F_; f++)
    for (int k = fe_off[f]; k < fe_off[f + 1]; k++) {
      const int l = fe_link[k];
      const int slot = lf_face[2 * l] < 0 ? 0 : 1;
      if (lf_face[2 * l + slot] >= 0) throw std::runtime_error("a link lies on more than two faces");
      lf_face[2 * l + slot] = f;
      lf_sign[2 * l + slot] = fe_sign[k];
    }
  std::vector<int> st_off(V_ + 1, 0), st_link, st_sign;
  std::vector<std::vector<std::pair<int, int>>> at(V_);
  std::vector<int2> ends(E_);
  for (int l = 0; l < E_; l++) {
    const int x = h.links[l][0], y = h.links[l][1];
    ends[l] = int2{x, y};
    at[y].push_back({l, +1});  // ft[y] += g, ft[x] -= g (as GaugeAction::force)
    at[x].push_back({l, -1});
  }
  for (int x = 0; x < V_; x++) {
    for (auto& pr : at[x]) {
      st_link.push_back(pr.first);
      st_sign.push_back(pr.second);
    }
    st_off[x + 1] = int(st_link.size());
  }
  fe_off_ = upload(fe_off);
  fe_link_ = upload(fe_link);
  fe_sign_ = upload(fe_sign);
  lf_face_ = upload(lf_face);
  lf_sign_ = upload(lf_sign);
  st_off_ = upload(st_off);
  st_link_ = upload(st_link);
  st_sign_ = upload(st_sign);
  ends_ = upload(ends);
  beta_s_ = upload(beta_s);
  beta_t_ = upload(beta_t);
  QED3_HIP(hipMalloc(&gf_, size_t(Lt_) * F_ * sizeof(double)));
  QED3_HIP(hipMalloc(&gt_, size_t(Lt_) * E_ * sizeof(double)));
  nblk_ = 4096;
  QED3_HIP(hipMalloc(&partials_, (nblk_ + 1) * sizeof(double2)));
  QED3_HIP(hipHostMalloc((void**)&hpart_, (nblk_ + 1) * sizeof(double2), hipHostMallocDefault));
}

DeviceGauge::~DeviceGauge() {
  for (void* p : {(void*)fe_off_, (void*)fe_link_, (void*)fe_sign_, (void*)lf_face_, (void*)lf_sign_,
                  (void*)st_off_, (void*)st_link_, (void*)st_sign_, (void*)ends_, (void*)beta_s_,
                  (void*)beta_t_, (void*)gf_, (void*)gt_, (void*)partials_})
    if (p) (void)hipFree(p);
  if (hpart_) (void)hipHostFree(hpart_);
}

void DeviceGauge::force(const double* u, const double* t, double* fs, double* ft, hipStream_t st) {
  const size_t n1 = size_t(Lt_) * (F_ + E_), n2 = size_t(Lt_) * (E_ + V_);
  hipLaunchKernelGGL(k_plaq_grads, dim3(blocks_for(n1)), dim3(256), 0, st, u, t, fe_off_, fe_link_, fe_sign_,
                     ends_, beta_s_, beta_t_, compact_, Lt_, V_, E_, F_, gf_, gt_);
  hipLaunchKernelGGL(k_force_links, dim3(blocks_for(n2)), dim3(256), 0, st, gf_, gt_, lf_face_, lf_sign_,
                     st_off_, st_link_, st_sign_, Lt_, V_, E_, F_, fs, ft);
  QED3_HIP(hipGetLastError());
}

double DeviceGauge::action(const double* u, const double* t, hipStream_t st) {
  const size_t n = size_t(Lt_) * (F_ + E_);
  const unsigned nb = std::min(blocks_for(n), nblk_);
  hipLaunchKernelGGL(k_action_parts, dim3(nb), dim3(256), 0, st, u, t, fe_off_, fe_link_, fe_sign_, ends_,
                     beta_s_, beta_t_, compact_, Lt_, V_, E_, F_, partials_);
  QED3_HIP(hipMemcpyAsync(hpart_, partials_, nb * sizeof(double2), hipMemcpyDeviceToHost, st));
  QED3_HIP(hipStreamSynchronize(st));
  long double s = 0;
  for (unsigned i = 0; i < nb; i++) s += hpart_[i].x;
  return double(s);
}

double DeviceGauge::norm2(const double* a, size_t n, hipStream_t st) {
  const unsigned nb = std::min(blocks_for(n), nblk_);
  hipLaunchKernelGGL(k_norm2_parts, dim3(nb), dim3(256), 0, st, a, n, partials_);
  QED3_HIP(hipMemcpyAsync(hpart_, partials_, nb * sizeof(double2), hipMemcpyDeviceToHost, st));
  QED3_HIP(hipStreamSynchronize(st));
  long double s = 0;
  for (unsigned i = 0; i < nb; i++) s += hpart_[i].x;
  return double(s);
}

void daxpy(double* y, const double* x, double c, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_daxpy, dim3(blocks_for(n)), dim3(256), 0, st, y, x, c, n);
  QED3_HIP(hipGetLastError());
}

}  // namespace qed3
