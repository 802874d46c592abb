// The U(1) gauge action and its force on the device, for the HMC integrator (the host
// GaugeAction in hmc.hpp defines the same quantities; both follow action_ext.h:33-159 and
// the plaquettes of gauge_ext.h:124-143).  Force in gather form: one thread per spatial link
// (its two faces and two temporal plaquettes) and one per temporal link (the temporal
// plaquettes of the site's spatial links), from per-plaquette derivatives computed first.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "geometry.hpp"

namespace qed3 {

class DeviceGauge {
 public:
  // beta_s per face, beta_t per link, and the oriented edges of every face (fe_off/link/sign)
  DeviceGauge(const HopTables& h, const std::vector<double>& beta_s, const std::vector<double>& beta_t,
              const std::vector<int>& fe_off, const std::vector<int>& fe_link,
              const std::vector<int>& fe_sign, bool compact);
  ~DeviceGauge();
  DeviceGauge(const DeviceGauge&) = delete;
  DeviceGauge& operator=(const DeviceGauge&) = delete;
  // fs[Lt][E], ft[Lt][V] = dS/d(angle) (device arrays)
  void force(const double* u, const double* t, double* fs, double* ft, hipStream_t st);
  double action(const double* u, const double* t, hipStream_t st);  // FP64
  // sum of squares of n doubles (kinetic energy pieces), FP64
  double norm2(const double* a, size_t n, hipStream_t st);

 private:
  int Lt_, V_, E_, F_;
  bool compact_;
  int *fe_off_ = nullptr, *fe_link_ = nullptr, *fe_sign_ = nullptr;
  int *lf_face_ = nullptr, *lf_sign_ = nullptr;  // [E][2] faces of each link
  int *st_off_ = nullptr, *st_link_ = nullptr, *st_sign_ = nullptr;  // links at each site
  int2* ends_ = nullptr;
  double *beta_s_ = nullptr, *beta_t_ = nullptr;
  double *gf_ = nullptr, *gt_ = nullptr;  // per-plaquette derivatives [Lt][F], [Lt][E]
  double2* partials_ = nullptr;
  double2* hpart_ = nullptr;
  unsigned nblk_ = 0;
};

// Device leapfrog pieces: y += c x over n doubles.
void daxpy(double* y, const double* x, double c, size_t n, hipStream_t st);

}  // namespace qed3
