// Seeded inputs in the reference's RNG scheme (src/both_3d/rng.h:6-128): a master
// std::mt19937_64 reseeds one stream per (t, link) and then one per (t, site), in that
// order; gauge angles draw from the link streams (temporal angles from the site
// streams, gauge_ext.h:146-149) and pseudofermion noise from the site streams.
#pragma once
#include <cstdint>
#include <iosfwd>
#include <random>
#include <string>
#include <vector>

namespace qed3 {

struct RngStream {
  std::mt19937_64 mt;
  std::normal_distribution<double> gauss;
  std::uniform_real_distribution<double> unif;
  std::uniform_int_distribution<int> z2{1, 2};  // kept for the checkpoint layout (rng.h:10-13)
  double gaussian() { return gauss(mt); }
  double uniform() { return unif(mt); }
  // SingleRng::reseed(const int) narrows the master draw to int (rng.h:25)
  void reseed(int seed) { mt.seed(seed); }
  // Checkpoint text of one stream (rng.h:28-34): four lines, the engine and the three
  // distributions, each in the standard library's stream form.
  void write_text(std::ostream& os) const;
  // Reads the four lines back.  Only the engine state is restored and the distributions
  // are reset: the reference reader (rng.h:36-52) extracts the engine line to its end, the
  // shared stringstream then holds eofbit, and its three distribution extractions fail,
  // so a resumed reference run continues from fresh distributions on the saved engines.
  void read_text(std::istream& is);
};

// Persistent streams (HMC): Lt*(E+V) generators, as ParallelRngExt.
class ParallelRng {
 public:
  ParallelRng(int Lt, int E, int V, int seed);
  double gaussian_link(int s, int l) { return links_[size_t(s) * E_ + l].gaussian(); }
  double gaussian_site(int s, int x) { return sites_[size_t(s) * V_ + x].gaussian(); }
  double uniform() { return master_.uniform(); }
  double gaussian_master() { return master_.gaussian(); }
  int Lt() const { return Lt_; }
  int E() const { return E_; }
  int V() const { return V_; }
  RngStream& link(int s, int l) { return links_[size_t(s) * E_ + l]; }
  RngStream& site(int s, int x) { return sites_[size_t(s) * V_ + x]; }
  // Checkpoint file in the layout of ParallelRngExt::ckpoint/read (rng.h:131-168): master,
  // then links [t][l], then sites [t][x], four lines each.  Throws on I/O or parse errors
  // (the reference asserts).
  void save(const std::string& path) const;
  void load(const std::string& path);

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
