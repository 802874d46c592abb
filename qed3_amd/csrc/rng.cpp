#include "rng.hpp"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <thread>

namespace qed3 {

namespace {
// Run f(i) for i in [0, n) over the host cores; each index is independent.
template <class F>
void parallel_for(size_t n, F&& f) {
  const unsigned nt = std::max(1u, std::min(std::thread::hardware_concurrency(), 64u));
  if (n < 4096 || nt == 1) {
    for (size_t i = 0; i < n; i++) f(i);
    return;
  }
  std::vector<std::thread> th;
  const size_t chunk = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; t++) {
    const size_t a = t * chunk, b = std::min(n, a + chunk);
    if (a >= b) break;
    th.emplace_back([a, b, &f] {
      for (size_t i = a; i < b; i++) f(i);
    });
  }
  for (auto& t : th) t.join();
}

// The per-stream seeds a master generator hands out, in the reference's order.
std::vector<int> stream_seeds(int seed, size_t skip, size_t count) {
  std::mt19937_64 master;
  master.seed(seed);
  master.discard((unsigned long long)skip);
  std::vector<int> s(count);
  for (auto& v : s) v = int(master());
  return s;
}
}  // namespace

ParallelRng::ParallelRng(int Lt, int E, int V, int seed) : Lt_(Lt), E_(E), V_(V) {
  links_.resize(size_t(Lt) * E);
  sites_.resize(size_t(Lt) * V);
  master_.reseed(seed);
  for (auto& r : links_) r.reseed(int(master_.mt()));
  for (auto& r : sites_) r.reseed(int(master_.mt()));
}

double von_mises(RngStream& r, double kappa) {
  if (!(kappa > 0)) return -M_PI + 2.0 * M_PI * r.uniform();
  const double tau = 1.0 + std::sqrt(1.0 + 4.0 * kappa * kappa);
  const double rho = (tau - std::sqrt(2.0 * tau)) / (2.0 * kappa);
  const double rr = (1.0 + rho * rho) / (2.0 * rho);
  for (;;) {
    const double u1 = r.uniform(), u2 = r.uniform(), u3 = r.uniform();
    const double z = std::cos(M_PI * u1);
    const double f = (1.0 + rr * z) / (rr + z);
    const double c = kappa * (rr - f);
    if (c * (2.0 - c) - u2 > 0.0 || std::log(c / u2) + 1.0 - c >= 0.0) {
      const double th = std::acos(std::max(-1.0, std::min(1.0, f)));
      return (u3 > 0.5) ? th : -th;
    }
  }
}

void draw_gauge(int Lt, int E, int V, int seed, double kappa, double* spatial, double* temporal) {
  const size_t nS = size_t(Lt) * E, nT = size_t(Lt) * V;
  const std::vector<int> seeds = stream_seeds(seed, 0, nS + nT);
  parallel_for(nS + nT, [&](size_t i) {
    RngStream s;
    s.reseed(seeds[i]);
    const double v = von_mises(s, kappa);
    if (i < nS)
      spatial[i] = v;
    else
      temporal[i - nS] = v;
  });
}

void draw_source(int Lt, int E, int V, int seed, double* xi) {
  const size_t nT = size_t(Lt) * V;
  const std::vector<int> seeds = stream_seeds(seed, size_t(Lt) * E, nT);
  const double inv = 1.0 / std::sqrt(2.0);
  parallel_for(nT, [&](size_t k) {
    RngStream s;
    s.reseed(seeds[k]);
    for (int a = 0; a < 2; a++) {
      const double re = s.gaussian();
      const double im = s.gaussian();
      xi[4 * k + 2 * a] = re * inv;
      xi[4 * k + 2 * a + 1] = im * inv;
    }
  });
}

void RngStream::write_text(std::ostream& os) const {
  os << mt << '\n' << gauss << '\n' << unif << '\n' << z2 << '\n';
}

void RngStream::read_text(std::istream& is) {
  std::string line;
  if (!std::getline(is, line)) throw std::runtime_error("rng checkpoint: missing engine line");
  std::istringstream ss(line);
  std::mt19937_64 m;
  if (!(ss >> m)) throw std::runtime_error("rng checkpoint: bad engine line");
  for (int k = 0; k < 3; k++)
    if (!std::getline(is, line)) throw std::runtime_error("rng checkpoint: missing distribution line");
  mt = m;
  gauss.reset();
  gauss = std::normal_distribution<double>();
  unif = std::uniform_real_distribution<double>();
  z2 = std::uniform_int_distribution<int>(1, 2);
}

void ParallelRng::save(const std::string& path) const {
  std::ofstream os(path);
  if (!os) throw std::runtime_error("cannot write " + path);
  master_.write_text(os);
  for (const auto& r : links_) r.write_text(os);
  for (const auto& r : sites_) r.write_text(os);
  if (!os) throw std::runtime_error("write failed: " + path);
}

void ParallelRng::load(const std::string& path) {
  std::ifstream is(path);
  if (!is) throw std::runtime_error("cannot read " + path);
  master_.read_text(is);
  for (auto& r : links_) r.read_text(is);
  for (auto& r : sites_) r.read_text(is);
}

}  // namespace qed3
