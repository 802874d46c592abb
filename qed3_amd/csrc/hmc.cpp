#include "hmc.hpp"

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <stdexcept>

#include "host_parallel.hpp"

namespace qed3 {

GaugeAction::GaugeAction(const HopTables& h, double gsq, double at, bool compact)
    : h_(h), compact_(compact), Lt_(h.Lt) {
  if (!(gsq > 0)) throw std::invalid_argument("gsq must be positive");
  beta_s_.resize(h.F);
  beta_t_.resize(h.E);
  for (int f = 0; f < h.F; f++) beta_s_[f] = at / h.face_area[f] / gsq;   // action_ext.h:152-154
  for (int l = 0; l < h.E; l++)
    beta_t_[l] = 2.0 / at * h.link_volume[l] / (h.ell[l] * h.ell[l]) / gsq;  // :155-157
  std::map<std::pair<int, int>, int> id;
  for (int l = 0; l < h.E; l++) id[{h.links[l][0], h.links[l][1]}] = l;
  fe_off_.push_back(0);
  for (const auto& ring : h.face_rings) {
    for (size_t k = 0; k < ring.size(); k++) {
      const int a = ring[k], b = ring[(k + 1) % ring.size()];
      const bool fwd = a < b;
      fe_link_.push_back(id.at(fwd ? std::make_pair(a, b) : std::make_pair(b, a)));
      fe_sign_.push_back(fwd ? 1 : -1);
    }
    fe_off_.push_back(int(fe_link_.size()));
  }
}

// gauge.cu:319-335: C(t) = < P_f(s) P_f(s+t) + P_f(s) P_f(s-t) > / vol_f^2 averaged over
// slices s and faces f (2 Lt F terms), times taken mod Lt.
void GaugeAction::plaquette_correlator(const double* u, double* corr) const {
  const int F = int(h_.face_rings.size());
  std::vector<double> P(size_t(Lt_) * F);
  for (int s = 0; s < Lt_; s++)
    for (int f = 0; f < F; f++) P[size_t(s) * F + f] = face_angle(u, s, f);
  for (int t = 0; t < Lt_; t++) {
    double acc = 0;
    for (int s = 0; s < Lt_; s++) {
      const double* a = &P[size_t(s) * F];
      const double* b = &P[size_t((s + t) % Lt_) * F];
      const double* c = &P[size_t(((s - t) % Lt_ + Lt_) % Lt_) * F];
      for (int f = 0; f < F; f++) {
        const double w = 1.0 / (h_.face_area[f] * h_.face_area[f]);
        acc += a[f] * b[f] * w + a[f] * c[f] * w;
      }
    }
    corr[t] = acc / (2.0 * Lt_ * F);
  }
}

double GaugeAction::face_angle(const double* u, int s, int f) const {
  const double* us = u + size_t((s + Lt_) % Lt_) * h_.E;
  double sum = 0;
  for (int k = fe_off_[f]; k < fe_off_[f + 1]; k++) sum += fe_sign_[k] * us[fe_link_[k]];
  return sum;
}

double GaugeAction::link_angle(const double* u, const double* t, int s, int l) const {
  const int s0 = (s + Lt_) % Lt_, s1 = (s + 1) % Lt_;
  const int x = h_.links[l][0], y = h_.links[l][1];
  return u[size_t(s0) * h_.E + l] + t[size_t(s0) * h_.V + y] - u[size_t(s1) * h_.E + l] -
         t[size_t(s0) * h_.V + x];
}

double GaugeAction::S(const double* u, const double* t) const {
  const int F = h_.F, E = h_.E;
  std::vector<long double> part(size_t(Lt_) * 2, 0.0L);
  parallel_tasks(size_t(Lt_) * 2, [&](size_t k) {
    const int s = int(k / 2);
    long double acc = 0;
    if (k % 2 == 0) {
      for (int f = 0; f < F; f++) {
        const double P = face_angle(u, s, f);
        acc += compact_ ? beta_s_[f] * (1.0 - std::cos(P)) : 0.5 * beta_s_[f] * P * P;
      }
    } else {
      for (int l = 0; l < E; l++) {
        const double P = link_angle(u, t, s, l);
        acc += compact_ ? beta_t_[l] * (1.0 - std::cos(P)) : 0.5 * beta_t_[l] * P * P;
      }
    }
    part[k] = acc;
  });
  long double tot = 0;
  for (auto v : part) tot += v;
  return double(tot);
}

void GaugeAction::force(const double* u, const double* t, double* fs, double* ft) const {
  const int F = h_.F, E = h_.E, V = h_.V;
  auto g = [&](double b, double P) { return compact_ ? b * std::sin(P) : b * P; };
  // one slice per task: every write lands in slice s (spatial faces), or in slice s of
  // fs/ft from the temporal plaquettes at s and s-1, so slices are independent.
  parallel_tasks(size_t(Lt_), [&](size_t ss) {
    const int s = int(ss);
    double* fss = fs + size_t(s) * E;
    double* fts = ft + size_t(s) * V;
    for (int l = 0; l < E; l++) fss[l] = 0;
    for (int x = 0; x < V; x++) fts[x] = 0;
    for (int f = 0; f < F; f++) {
      const double grad = g(beta_s_[f], face_angle(u, s, f));
      for (int k = fe_off_[f]; k < fe_off_[f + 1]; k++) fss[fe_link_[k]] += grad * fe_sign_[k];
    }
    for (int l = 0; l < E; l++) {
      const double grad = g(beta_t_[l], link_angle(u, t, s, l));
      const double grad2 = g(beta_t_[l], link_angle(u, t, s - 1, l));
      fss[l] += grad - grad2;
      fts[h_.links[l][1]] += grad;
      fts[h_.links[l][0]] -= grad;
    }
  });
}

HMC::HMC(Dirac& D, const HMCParams& p)
    : D_(D), p_(p), Sg_(D.tables(), p.gsq, D.tables().at, p.compact) {
  // npf = 0 is the pure-gauge HMC (HMCPureGauge, hmc.h:212-296): with no fermion kicks the
  // two-level leapfrog is one leapfrog of nsteps * nsteps_inner gauge steps
  if (p.npf < 0 || p.nsteps < 1 || p.nsteps_inner < 1) throw std::invalid_argument("bad HMC parameters");
  const HopTables& h = D.tables();
  rng_ = std::make_unique<ParallelRng>(h.Lt, h.E, h.V, p.seed);
  if (p.pcg) K_ = std::make_unique<Horner>(D, 4);
  const size_t n = D.points();
  phi_.resize(p.npf);
  eta_.resize(p.npf);
  for (int k = 0; k < p.npf; k++) {
    QED3_HIP(hipMalloc(&phi_[k], n * sizeof(Spin<double>)));
    QED3_HIP(hipMalloc(&eta_[k], n * sizeof(Spin<double>)));
  }
  QED3_HIP(hipMalloc(&xi_, n * sizeof(Spin<double>)));
  QED3_HIP(hipMalloc(&dfs_, size_t(h.Lt) * h.E * sizeof(double)));
  QED3_HIP(hipMalloc(&dft_, size_t(h.Lt) * h.V * sizeof(double)));
  dg_ = std::make_unique<DeviceGauge>(h, Sg_.beta_s(), Sg_.beta_t(), Sg_.face_edge_offsets(),
                                      Sg_.face_edge_links(), Sg_.face_edge_signs(), Sg_.compact());
  const size_t nS = size_t(h.Lt) * h.E, nT = size_t(h.Lt) * h.V;
  for (double** p : {&du_, &dpu_, &dgs_}) QED3_HIP(hipMalloc(p, nS * sizeof(double)));
  for (double** p : {&dt_, &dpt_, &dgt_}) QED3_HIP(hipMalloc(p, nT * sizeof(double)));
}

HMC::~HMC() {
  for (auto* p : phi_) (void)hipFree(p);
  for (auto* p : eta_) (void)hipFree(p);
  (void)hipFree(xi_);
  (void)hipFree(dfs_);
  (void)hipFree(dft_);
  for (double* p : {du_, dt_, dpu_, dpt_, dgs_, dgt_}) (void)hipFree(p);
}

void HMC::update_eta(int k) {
  CGOptions o;
  o.tol = p_.tol;
  o.mixed = p_.mixed;
  o.half = p_.half;
  // Wilson: CG on D^dag D; overlap: FP64 CG on D_ov + D_ov^dag (PseudoFermion::update_eta)
  if (K_) o.delta = 0.001;  // the learned method's reliable-update threshold
  const CGResult r = ov_   ? ov_->solve_dhd(eta_[k], phi_[k], p_.tol, o.maxiter)
                     : K_ ? pcg_solve(*K_, eta_[k], phi_[k], o)
                          : cg_solve(D_, eta_[k], phi_[k], o);
  rep_.cg_solves++;
  rep_.cg_iterations += r.iterations;
  rep_.max_resid = std::max(rep_.max_resid, r.true_resid);
  if (r.status != 0) throw std::runtime_error("HMC: pseudofermion solve failed");
}

// PseudoFermion::gen (pseudofermion.h:36-49): xi from the site streams, phi = D^dag xi
void HMC::pf_gen(int k) {
  const HopTables& h = D_.tables();
  const size_t n = D_.points();
  std::vector<double> xi(4 * n);
  const double inv = 1.0 / std::sqrt(2.0);
  parallel_ranges(n, [&](size_t p0, size_t p1) {  // one site stream per point, drawn in order
    for (size_t P = p0; P < p1; P++) {
      const int s = int(P / h.V), x = int(P % h.V);
      for (int a = 0; a < 2; a++) {
        const double re = rng_->gaussian_site(s, x);
        const double im = rng_->gaussian_site(s, x);
        xi[4 * P + 2 * a] = re * inv;
        xi[4 * P + 2 * a + 1] = im * inv;
      }
    }
  });
  QED3_HIP(hipMemcpy(xi_, xi.data(), xi.size() * sizeof(double), hipMemcpyHostToDevice));
  if (ov_)
    ov_->apply(true, phi_[k], xi_);
  else
    D_.apply<double>(true, phi_[k], xi_, true);
  update_eta(k);
}

void HMC::fermion_force(int k) {
  if (ov_)
    ov_->force(eta_[k], dfs_, dft_);
  else
    D_.force(eta_[k], dfs_, dft_);
}

void HMC::gauge_changed() {
  if (ov_ && ov_refresh_) ov_->refresh_bounds();
}

double HMC::Sf(int k) {  // Re <phi, eta> (pseudofermion.h:53-58)
  double re, im;
  D_.dot(phi_[k], eta_[k], &re, &im);
  return re;
}

double HMC::H() {  // hmc.h:159-166: 1/2 |pi|^2 + S_g + sum_k S_f (device state, FP64)
  const HopTables& h = D_.tables();
  hipStream_t st = D_.stream();
  const size_t nS = size_t(h.Lt) * h.E, nT = size_t(h.Lt) * h.V;
  long double res = 0.5L * ((long double)dg_->norm2(dpu_, nS, st) + dg_->norm2(dpt_, nT, st));
  res += dg_->action(du_, dt_, st);
  for (int k = 0; k < p_.npf; k++) res += Sf(k);
  return double(res);
}

HMCReport HMC::trajectory(double* u, double* t, bool no_reject) {
  const auto t0 = std::chrono::steady_clock::now();
  const HopTables& h = D_.tables();
  const size_t nS = size_t(h.Lt) * h.E, nT = size_t(h.Lt) * h.V;
  hipStream_t st = D_.stream();
  rep_ = HMCReport();
  // per-phase host wall time, printed with QED3_HMC_TRACE set in the environment
  const bool trace = std::getenv("QED3_HMC_TRACE") != nullptr;
  double t_rng = 0, t_gauge = 0, t_fermion = 0, t_h = 0;
  auto now = [] { return std::chrono::steady_clock::now(); };
  auto since = [&](std::chrono::steady_clock::time_point a) {
    QED3_HIP(hipStreamSynchronize(st));
    return std::chrono::duration<double>(now() - a).count();
  };
  auto ta = now();
  pu_.resize(nS);
  pt_.resize(nT);
  // pi.gaussian(rng) (gauge_ext.h:146-149); every entry has a stream of its own, so the
  // draws run in parallel without changing any stream's sequence
  parallel_ranges(nS, [&](size_t a, size_t b) {
    for (size_t i = a; i < b; i++) pu_[i] = rng_->gaussian_link(int(i / h.E), int(i % h.E));
  });
  parallel_ranges(nT, [&](size_t a, size_t b) {
    for (size_t i = a; i < b; i++) pt_[i] = rng_->gaussian_site(int(i / h.V), int(i % h.V));
  });
  // the integrator's state lives on the device from here to the accept step
  QED3_HIP(hipMemcpyAsync(dpu_, pu_.data(), nS * sizeof(double), hipMemcpyHostToDevice, st));
  QED3_HIP(hipMemcpyAsync(dpt_, pt_.data(), nT * sizeof(double), hipMemcpyHostToDevice, st));
  QED3_HIP(hipMemcpyAsync(du_, u, nS * sizeof(double), hipMemcpyHostToDevice, st));
  QED3_HIP(hipMemcpyAsync(dt_, t, nT * sizeof(double), hipMemcpyHostToDevice, st));
  D_.set_gauge(du_, dt_, true);
  gauge_changed();
  t_rng += since(ta);
  ta = now();
  for (int k = 0; k < p_.npf; k++) pf_gen(k);
  t_fermion += since(ta);
  ta = now();
  const double h0 = H();
  t_h += since(ta);

  // ExplicitLeapfrogML::integrate, multi-pseudofermion form (integrator.h:132-168)
  const double tau = p_.tau / p_.nsteps, tau_in = tau / p_.nsteps_inner;
  auto kick = [&](const double* fs, const double* ft, double c) {
    daxpy(dpu_, fs, c, nS, st);
    daxpy(dpt_, ft, c, nT, st);
  };
  auto gauge_force = [&] { dg_->force(du_, dt_, dgs_, dgt_, st); };
  ta = now();
  for (int k = 0; k < p_.npf; k++) {
    fermion_force(k);
    kick(dfs_, dft_, -0.5 * tau);
  }
  t_fermion += since(ta);
  for (int n = 0; n < p_.nsteps; n++) {
    ta = now();
    gauge_force();
    kick(dgs_, dgt_, -0.5 * tau_in);
    for (int ni = 0; ni < p_.nsteps_inner; ni++) {
      daxpy(du_, dpu_, tau_in, nS, st);
      daxpy(dt_, dpt_, tau_in, nT, st);
      gauge_force();
      kick(dgs_, dgt_, (ni != p_.nsteps_inner - 1) ? -1.0 * tau_in : -0.5 * tau_in);
    }
    D_.set_gauge(du_, dt_, true);
    gauge_changed();
    t_gauge += since(ta);
    ta = now();
    for (int k = 0; k < p_.npf; k++) {
      update_eta(k);
      fermion_force(k);
      kick(dfs_, dft_, (n != p_.nsteps - 1) ? -1.0 * tau : -0.5 * tau);
    }
    t_fermion += since(ta);
  }
  ta = now();
  const double h1 = H();
  t_h += since(ta);
  if (trace)
    std::fprintf(stderr, "hmc trajectory: rng+upload %.3fs gauge %.3fs fermion %.3fs H %.3fs\n",
                 t_rng, t_gauge, t_fermion, t_h);
  rep_.h0 = h0;
  rep_.h1 = h1;
  rep_.dH = h1 - h0;
  const double r = std::min(1.0, std::exp(-rep_.dH));
  const double a = rng_->uniform();
  rep_.accept = (a < r || no_reject) ? 1 : 0;
  if (rep_.accept) {
    QED3_HIP(hipMemcpyAsync(u, du_, nS * sizeof(double), hipMemcpyDeviceToHost, st));
    QED3_HIP(hipMemcpyAsync(t, dt_, nT * sizeof(double), hipMemcpyDeviceToHost, st));
    QED3_HIP(hipStreamSynchronize(st));
  } else {
    D_.set_gauge(u, t, false);
    gauge_changed();
  }
  rep_.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return rep_;
}

}  // namespace qed3
