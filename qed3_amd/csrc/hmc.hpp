// Host-side HMC over the GPU fermion path: pseudofermions, U(1) gauge action (the
// reference's non-compact form and the compact Wilson form), the multi-level leapfrog
// and the accept/reject step.  Roles of the reference's PseudoFermion
// (src/both_3d/pseudofermion.h:6-61), U1WilsonExt (action_ext.h:4-161),
// ExplicitLeapfrogML (integrator.h:58-170) and HMC2 (hmc.h:103-201).
// The Hamiltonian and accept step are FP64 on the host; the fermion solves and forces
// run on the GPU.
#pragma once
#include <memory>
#include <vector>

#include "cg.hpp"
#include "dirac.hpp"
#include "horner.hpp"
#include "gauge_dev.hpp"
#include "msolve.hpp"
#include "rng.hpp"

namespace qed3 {

// U(1) plaquette action on the lattice of a Dirac operator.
class GaugeAction {
 public:
  GaugeAction(const HopTables& h, double gsq, double at, bool compact);
  double S(const double* u, const double* t) const;
  void force(const double* u, const double* t, double* fs, double* ft) const;
  // plaquette-angle correlator over time separations t = 0..Lt-1 (corr[Lt])
  void plaquette_correlator(const double* u, double* corr) const;
  const std::vector<double>& beta_s() const { return beta_s_; }
  const std::vector<double>& beta_t() const { return beta_t_; }
  const std::vector<int>& face_edge_offsets() const { return fe_off_; }
  const std::vector<int>& face_edge_links() const { return fe_link_; }
  const std::vector<int>& face_edge_signs() const { return fe_sign_; }
  bool compact() const { return compact_; }

 private:
  double face_angle(const double* u, int s, int f) const;
  double link_angle(const double* u, const double* t, int s, int l) const;
  const HopTables& h_;
  bool compact_;
  int Lt_;
  std::vector<double> beta_s_, beta_t_;
  std::vector<int> fe_off_, fe_link_, fe_sign_;  // oriented edges of each face
};

struct HMCParams {
  double gsq = 1.0 / 1.4;
  bool compact = true;
  int npf = 1;
  double tau = 1.0;
  int nsteps = 20;
  int nsteps_inner = 10;
  double tol = 1e-10;
  bool mixed = true;
  bool half = false;  // binary16-stored inner CG fields
  bool pcg = false;   // flexible CG with the time-exact preconditioner (r = 1 only)
  int seed = 1;
};

struct HMCReport {
  double dH = 0, h0 = 0, h1 = 0;
  int accept = 0;
  int cg_solves = 0;
  long cg_iterations = 0;
  double seconds = 0;
  double max_resid = 0;
};

class HMC {
 public:
  HMC(Dirac& D, const HMCParams& p);
  ~HMC();
  // one trajectory; u/t (host, gauge layout) are replaced by the new field on accept
  HMCReport trajectory(double* u, double* t, bool no_reject = false);
  // Overlap fermions (Fermion = Overlap<WilsonDirac>, hmc.cu:155-186) on the same context:
  // phi = D_ov^dag xi, eta = (D_ov + D_ov^dag)^{-1} phi, the overlap force.  With refresh the
  // spectral bounds and Zolotarev coefficients are recomputed after every gauge update, as
  // Overlap::update does in the reference; without, they stay those of the overlap
  // handle, so H is a smooth function of the links.  ov = nullptr returns to Wilson fermions.
  void set_overlap(Overlap* ov, bool refresh) {
    ov_ = ov;
    ov_refresh_ = refresh;
  }
  void set_steps(double tau, int nsteps, int nsteps_inner) {
    p_.tau = tau;
    p_.nsteps = nsteps;
    p_.nsteps_inner = nsteps_inner;
  }
  ParallelRng& rng() { return *rng_; }
  const ParallelRng& rng() const { return *rng_; }

 private:
  void set_gauge(const std::vector<double>& u, const std::vector<double>& t);
  void pf_gen(int k);
  void update_eta(int k);
  double Sf(int k);
  double H();
  void fermion_force(int k);   // dfs_, dft_ = the force of pseudofermion k
  void gauge_changed();        // after D_.set_gauge: refresh the overlap bounds if asked

  Dirac& D_;
  HMCParams p_;
  GaugeAction Sg_;
  std::unique_ptr<ParallelRng> rng_;
  std::unique_ptr<Horner> K_;  // the preconditioner of the pcg solver
  std::vector<Spin<double>*> phi_, eta_;
  Spin<double>* xi_ = nullptr;
  double* dfs_ = nullptr;
  double* dft_ = nullptr;
  std::vector<double> pu_, pt_;  // host momenta (drawn from the host streams)
  // device state of the integrator: angles, momenta, gauge force
  std::unique_ptr<DeviceGauge> dg_;
  double *du_ = nullptr, *dt_ = nullptr, *dpu_ = nullptr, *dpt_ = nullptr;
  double *dgs_ = nullptr, *dgt_ = nullptr;
  HMCReport rep_;
  Overlap* ov_ = nullptr;
  bool ov_refresh_ = false;
};

}  // namespace qed3
