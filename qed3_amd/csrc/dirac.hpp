// Host-side Wilson-Dirac operator on S^2 x time for one GPU: owns the hop tables, the
// gauge phases and the kernels' scratch.  Plays the role of the reference's
// DiracExt + DWDevice + DiracPf trio (src/both_3d/dirac_ext.h, sparse_dirac.h,
// dirac_pf.h) with a matrix-free stencil instead of CSR.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <vector>

#include "device_common.hpp"
#include "geometry.hpp"

namespace qed3 {

// Device copy of the time-independent tables in one precision; P is the stored type of
// the link phases (T's complex, or binary16 pairs for the half-storage solver).
template <typename T, typename P = cplx_t<T>>
struct DiracView {
  const int* nbr = nullptr;          // [nslot][V]
  const int* lnk = nullptr;          // [nslot][V] link id or ~id
  const real4_t<T>* hop = nullptr;   // [nslot][V] (C,S,a,b) for D
  const real4_t<T>* hopd = nullptr;  // [nslot][V] for D^dagger
  const T* diag = nullptr;           // [V]
  const T* kth = nullptr;            // [V] kappa_t/2
  const P* uS = nullptr;             // [Lt][E] exp(i theta_s(l))
  const P* uT = nullptr;             // [Lt][V] exp(i theta^t_s(x))
  // Temporal gauge (set instead of uT): every temporal link is 1 except those from the last
  // slice to the first, whose phase is the product of the site's temporal phases ([V]).
  const P* uPoly = nullptr;
  // Per site and slot (set by the temporal-gauge pair view): the phase of the slot's link as
  // the site sees it, conjugated where the link is stored in the opposite orientation, so
  // that a wave reads consecutive entries ([Lt][nslot][V]).
  const P* uSite = nullptr;
  // The pair view's only spatial phases: the per-site angles rounded to 10 bits and packed
  // three slots to a 32-bit word (slot 3w + k in bits 10k..10k+9 of word w; [Lt][W][V],
  // W = ceil(nslot / 3)), for the learned preconditioner's layers.
  const unsigned* uSite10 = nullptr;
  int V = 0, E = 0, Lt = 1;
  T r = 1;
  T bc = -1;  // sign on the time wrap: -1 antiperiodic, +1 periodic
};

// Launch shape of one stencil kernel: threads per block and time slices per thread.
struct StencilTile {
  int block = 256;
  int ts = 4;
  int remap = 1;  // XCD-aware block order
  int lds = 0;    // stage the time-slice spinor tile (range + halo) in LDS
};

// Halo tables of the LDS stencil for one range size R (see k_stencil_lds): per slot the
// 16-bit index of the neighbour in the block's LDS tile (range sites first, then the
// range's halo), and the halo site lists of all ranges.
struct HaloView {
  const unsigned short* nbl = nullptr;  // [nslot][V]
  const int* halo_off = nullptr;        // [nranges+1]
  const int* halo = nullptr;            // halo site ids, grouped by range
  int hmax = 0;
};

template <typename T>
struct DeviceTables {
  int* nbr = nullptr;
  int* lnk = nullptr;
  real4_t<T>* hop = nullptr;
  real4_t<T>* hopd = nullptr;
  T* diag = nullptr;
  T* kth = nullptr;
  cplx_t<T>* uS = nullptr;
  cplx_t<T>* uT = nullptr;
};

// Scalars of a device-driven CG; read and written only by kernels.
struct CGScalars {
  double mu;       // |r|^2 of the iterated residual
  double gam;      // |D p|^2 = <p, D^dag D p>
  double alpha, beta;
  double maxr2;    // max |r|^2 since the last reliable update
  double bnorm2, tol2, delta2;
  double true_r2;  // last true FP64 |b - A x|^2
  double scale;    // half-storage CG: true field = scale * stored field (1 otherwise)
  double rfac;     // half-storage CG: factor from the kept residual to the new stored one
  double alpha_pend[8];  // alphas of the iterations whose x update is pending (0 = none):
                         // the half-storage CG uses 3, the preconditioned CG 8 (its ring)
  double rho;      // preconditioned CG: <r, z> of the last iteration (stored units)
  int done;        // 0 running, 1 converged, 2 breakdown / NaN, 3 maxiter
  int need_ru;     // a reliable update is due in this iteration
  int iter, n_ru, maxiter;
  int rescale;     // half-storage CG: the residual is to be re-stored at the new scale
  int last_pos;    // half-storage CG: group position of the last iteration run
  int fresh;       // preconditioned CG: no previous direction (beta = 0)
  int x_set;       // preconditioned CG: x holds a flushed value (else it is taken as 0)
  int xin_set;     // preconditioned CG: x_inner holds folded terms (else it is taken as 0)
  double rpend;    // preconditioned CG: factor the stored residual still owes (1 = none)
};

class Dirac {
 public:
  Dirac(const Geometry& g, int lattice, int Lt, int bc, double r, double M5, double at);
  ~Dirac();
  Dirac(const Dirac&) = delete;
  Dirac& operator=(const Dirac&) = delete;

  const HopTables& tables() const { return h_; }
  int V() const { return h_.V; }
  int E() const { return h_.E; }
  int Lt() const { return h_.Lt; }
  size_t points() const { return size_t(h_.V) * h_.Lt; }
  hipStream_t stream() const { return stream_; }
  // Queue this operator's work on `s` from now on (nullptr: its own stream again), after
  // the work already queued on the previous stream has finished.  `s` must not be the null
  // stream (the solvers capture graphs on it) and must outlive its use here.
  void set_stream(hipStream_t s);
  bool has_gauge() const { return has_gauge_; }

  // angles in the reference's GaugeExt layout; on host or device
  void set_gauge(const double* spatial, const double* temporal, bool on_device);

  template <typename T>
  DiracView<T> view() const;
  // FP32 tables with 16-bit fixed-point link angles (kept up to date with the gauge field
  // from the first call on)
  DiracView<float, phase16_t> view_half_phases();
  // The same operator in the temporal gauge for the half-storage solver: with
  // psi'(s,x) = exp(i phi_s(x)) psi(s,x), phi_0 = 0, phi_{s+1} = phi_s + theta^t_s(x), the
  // spatial angles become theta_l + phi_s(x_l) - phi_s(y_l) and the temporal links 1 but at
  // the wrap (kept up to date with the gauge field from the first call on).
  DiracView<float, phase16_t> view_half_temporal_gauge();
  // phi_s(x) as 16-bit angles, [Lt][V] (valid once view_half_temporal_gauge was called)
  const phase16_t* half_gauge_rotation() const { return rotH_; }
  // The temporal-gauge spatial links and wrap links of view_half_temporal_gauge as binary16
  // (cos, sin) pairs (kept up to date with the gauge field from the first call on)
  DiracView<float, phasepair_t> view_half_tg_pairs();
  template <typename T>
  StencilTile tile(bool dag) const;

  // out = D in / D^dag in (no reduction); blocking on the stream when sync
  template <typename T>
  void apply(bool dag, Spin<T>* out, const Spin<T>* in, bool sync = true);
  template <typename T>
  void apply_dagd(Spin<T>* out, const Spin<T>* in, Spin<T>* tmp, bool sync = true);
  template <typename T>
  void apply_ddag(Spin<T>* out, const Spin<T>* in, Spin<T>* tmp, bool sync = true);
  // binary16 fields, FP32 arithmetic, 16-bit link angles (the half-storage CG's stencil)
  void apply_half(bool dag, SpinH* out, const SpinH* in, bool sync = true);

  // complex <x,y> in FP64
  void dot(const Spin<double>* x, const Spin<double>* y, double* re, double* im);

  // pseudofermion force for all links from eta (FP64)
  void force(const Spin<double>* eta, double* fs, double* ft);
  // f_l (+)= c (Re <u1, dD/dtheta_l v1> + Re <u2, dD/dtheta_l v2>) for every link, in the
  // force layout (spatial [Lt][E], temporal [Lt][V]); (u2, v2) may both be null; f is
  // overwritten when init.  Asynchronous on the context's stream.
  void force_bilinear(const Spin<double>* u1, const Spin<double>* v1, const Spin<double>* u2,
                      const Spin<double>* v2, double c, bool init, double* fs, double* ft);

  // ms per launch of one D or D^dag apply; prec 0 FP64, 1 FP32, 2 binary16 storage
  double time_apply(bool dag, int prec, int warmup, int reps);

  // scratch for reductions
  GridReduce reducer() const { return {partials_, counter_}; }
  // Device work buffer `slot` of at least `bytes`, kept across calls (grown on demand,
  // freed with the context), so repeated solves allocate nothing; contents are undefined.
  // Slots are owned by the CG solvers (cg.hip); calls on one context are sequential.
  static constexpr int kScratchSlots = 14;
  void* scratch(int slot, size_t bytes);
  void release_scratch();
  // pinned host copy of the solver scalars (same lifetime)
  CGScalars* host_scalars();
  // LDS halo tables for ranges of R sites (built on first use)
  HaloView halo(int R);
  bool has_halo(int R) const { return halos_.count(R) != 0; }
  template <typename T>
  HaloView halo_for(bool dag) {
    const StencilTile t = tile<T>(dag);
    return t.lds ? halo(t.block) : HaloView{};
  }
  // autotune the stencil launch shapes for this lattice (persisted in memory per shape)
  void autotune();
  bool cg_tuned = false;  // the FP32 / binary16 shapes were refined on the CG iteration
  void set_tile(bool f32, bool dag, StencilTile t) {
    tile_[f32 ? 1 : 0][dag ? 1 : 0] = t;
    tuned_ = true;
  }

 private:
  template <typename T>
  void upload_tables(DeviceTables<T>& d);
  template <typename T>
  void tune_one(bool dag);

  HopTables h_;
  int bc_ = 0;
  hipStream_t stream_ = nullptr;      // where the work goes
  hipStream_t own_stream_ = nullptr;  // created and destroyed here
  DeviceTables<double> d64_;
  DeviceTables<float> d32_;
  phase16_t* uSh_ = nullptr;  // [Lt][E] 16-bit fixed-point angles (allocated by view_half_phases)
  phase16_t* uTh_ = nullptr;  // [Lt][V]
  void fill_half_phases();
  phase16_t* uStg_ = nullptr;   // [Lt][E] spatial angles in the temporal gauge
  phase16_t* uPolyH_ = nullptr; // [V] the wrap link's angle in the temporal gauge
  phase16_t* rotH_ = nullptr;   // [Lt][V] phi_s(x)
  phasepair_t* uPolyP_ = nullptr;  // [V] uPolyH_ as (cos, sin) pairs
  phase16_t* uSiteH_ = nullptr;    // [Lt][nslot][V] uStg_ per site and slot (DiracView::uSite)
  unsigned* uSite10_ = nullptr;    // [Lt][W][V] uSiteH_ packed at 10 bits (DiracView::uSite10)
  double* phi64_ = nullptr;     // [Lt][V] phi_s(x) in FP64 while the tables are built
  void fill_temporal_gauge();
  double* angS_ = nullptr;  // [Lt][E] FP64 angles (device copy)
  double* angT_ = nullptr;  // [Lt][V]
  int* link_ends_ = nullptr;  // [E][4]: x, y, slot of y in x, slot of x in y
  double2* partials_ = nullptr;
  unsigned int* counter_ = nullptr;
  double2* dot_out_ = nullptr;
  Spin<double>* tmp64_ = nullptr;  // D eta for the force
  bool has_gauge_ = false;
  StencilTile tile_[2][2];  // [f32][dag]
  bool tuned_ = false;
  struct HaloTables {
    unsigned short* nbl = nullptr;
    int* off = nullptr;
    int* sites = nullptr;
    int hmax = 0;
  };
  std::map<int, HaloTables> halos_;
  void* scratch_[kScratchSlots] = {};
  size_t scratch_bytes_[kScratchSlots] = {};
  CGScalars* hsc_ = nullptr;
};

size_t reduce_capacity();  // max blocks a reducing kernel may use
double time_copy(size_t bytes, int reps);  // ms per launch of a plain device copy
constexpr size_t kReduceCapacity = size_t(1) << 20;

}  // namespace qed3
