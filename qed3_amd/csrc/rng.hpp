// Seeded inputs in the reference's RNG scheme (src/both_3d/rng.h:6-128): a master
// std::mt19937_64 reseeds one stream per (t, link) and then one per (t, site), in that
// order; gauge angles draw from the link streams (temporal angles from the site
// streams, gauge_ext.h:146-149) and pseudofermion noise from the site streams.
#pragma once
#include <cstdint>
#include <random>
#include <vector>

namespace qed3 {

struct RngStream {
  std::mt19937_64 mt;
  std::normal_distribution<double> gauss;
  std::uniform_real_distribution<double> unif;
  double gaussian() { return gauss(mt); }
  double uniform() { return unif(mt); }
  // SingleRng::reseed(const int) narrows the master draw to int (rng.h:25)
  void reseed(int seed) { mt.seed(seed); }
};

// Persistent streams (HMC): Lt*(E+V) generators, as ParallelRngExt.
class ParallelRng {
 public:
  ParallelRng(int Lt, int E, int V, int seed);
  double gaussian_link(int s, int l) { return links_[size_t(s) * E_ + l].gaussian(); }
  double gaussian_site(int s, int x) { return sites_[size_t(s) * V_ + x].gaussian(); }
  double uniform() { return master_.uniform(); }
  RngStream& link(int s, int l) { return links_[size_t(s) * E_ + l]; }
  RngStream& site(int s, int x) { return sites_[size_t(s) * V_ + x]; }

 private:
  int Lt_, E_, V_;
  RngStream master_;
  std::vector<RngStream> links_, sites_;
};

// Best & Fisher, Appl. Statist. 28 (1979) 152: von Mises(mu=0, kappa) by rejection from
// a wrapped Cauchy envelope.  kappa <= 0 gives uniform on [-pi, pi).
double von_mises(RngStream& r, double kappa);

// One-pass draws with O(1) generator memory; identical to the first draws of the
// persistent streams for the same seed.
void draw_gauge(int Lt, int E, int V, int seed, double kappa, double* spatial, double* temporal);
// xi = (g1 + i g2)/sqrt2 per component from the site streams (rng.h:122-128); the real
// part is drawn first.
void draw_source(int Lt, int E, int V, int seed, double* xi);

}  // namespace qed3
