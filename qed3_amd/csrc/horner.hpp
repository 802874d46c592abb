// Time-exact Horner preconditioner for CG on D^dagger D: a gauge-covariant linear
// message-passing network on the lattice graph whose layers alternate the exact inverse of
// the operator's time part with one spatial message-passing step.
//
// With r = 1 the temporal hops of D are spin projectors (dirac_ext.h:158-187): the
// spin-up component only hops from s-1 to s and the spin-down component only from s+1
// to s.  Write D = T - S with
//   T = c_x + the temporal hops, the wrap link from Lt-1 to 0 left out ("open in time"),
//   S = -(the spatial hops + the wrap hop),
// so that for every site and spin row T is bidiagonal in time and T^{-1} is a first-order
// recurrence along the site's time column: forward in time for spin up, backward for spin
// down (reversed for T^dagger).  Keeping the wrap in S rather than in T costs nothing (a
// layer's recurrence starts from the previous layer's output at the far end of the column)
// and keeps D K = 1 - (S T^{-1})^L exact for the Neumann coefficients; leaving the wrap out
// altogether would leave an O(kappa_t) error on the first and last slices (measured: 36
// instead of 13 iterations with four layers at dual n=8, Lt=16).  The network is
//   K v = sum_{j<L} c_j (T^{-1} S_th)^j T^{-1} v,
// evaluated in Horner form, g <- T^{-1}(c_j v + S_th g) for j = L-1 .. 0 (g = 0 before the
// first layer), where S_th is S with the same-spin spatial entries of spin row a scaled by
// A[a] and the spin-flip ones by B[a] (the wrap hop is not scaled).  With c_j = A = B = 1
// it is the Neumann series
// D^{-1} = sum_j (T^{-1} S)^j T^{-1} truncated after L terms; the learned parameters are
// trained on the solver's own (residual, error) pairs.  K^dagger has the same Horner form
// with T^dagger, S_th^dagger and conj(c_j), so M = K K^dagger is Hermitian positive
// definite and serves as the preconditioner of flexible CG on D^dagger D.
//
// Every parameter and transport commutes with U(1) gauge transformations (K[U^g] =
// G K[U] G^dagger) and with local spin-frame rotations (A and B act per spin row, which
// frame rotations exp(i phi sigma3 / 2) leave alone), as the network of the learned
// preconditioner must.
#pragma once
#include <complex>
#include <vector>

#include "cg.hpp"
#include "dirac.hpp"

namespace qed3 {

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

// Spinor fields of the preconditioner are stored spin-planar: plane a holds the spin
// component a of every point (s, x) at [a * n + s * V + x], so the thread that walks the
// time column of one spin row reads and writes contiguous values.
struct HornerParams {
  int layers = 4;
  std::vector<std::complex<double>> c;  // [layers], c[j] multiplies the input at layer j
  std::complex<double> A[2] = {1.0, 1.0}, B[2] = {1.0, 1.0};
  int count() const { return layers + 4; }  // complex parameters: c[0..L), A0, A1, B0, B1
  void pack(std::complex<double>* p) const;
  void unpack(const std::complex<double>* p);
};

// One spin row's coefficients of one hop (same-spin, spin-flip) in binary16
struct HopRow16 {
  half2_t s, f;
};

// Device tables of one network on one operator.
struct HornerView {
  const float4* hrow = nullptr;  // [3 (K, K^dag, unit)][2 (spin row)][nslot][V] (same, flip) coefficients
  const HopRow16* hrow16 = nullptr;  // [2 (K, K^dag)][2][nslot][V] the same in binary16
  const float* invc = nullptr;   // [V] 1 / c_x
  const float* kt = nullptr;     // [V] kappa_t(x) (the temporal hop weight for r = 1)
  int V = 0, Lt = 1, nslot = 3;
};

class Horner {
 public:
  // The operator must have r = 1 (the projector structure above); throws otherwise.
  Horner(Dirac& D, int layers);
  ~Horner();
  Horner(const Horner&) = delete;
  Horner& operator=(const Horner&) = delete;

  Dirac& dirac() { return D_; }
  int layers() const { return p_.layers; }
  const HornerParams& params() const { return p_; }
  void set_params(const HornerParams& p);
  HornerView view() const;

  // out = K in (adj = false) or K^dagger in, FP64 fields in the reference layout, computed
  // in FP32 on the operator's own link phases (original gauge).
  void apply(bool adj, Spin<double>* out, const Spin<double>* in);

  // loss = mean_b |K v_b - e_b|^2 / |e_b|^2 over `batch` pairs of FP64 fields, and its
  // gradient d loss / d Re p + i d loss / d Im p (p packed as HornerParams::pack)
  double loss_grad(const Spin<double>* v, const Spin<double>* e, int batch,
                   std::vector<std::complex<double>>* grad);

  // Adam over npairs (v, e) device pairs in mini-batches; returns the mean loss per epoch
  std::vector<double> train(const Spin<double>* v, const Spin<double>* e, int npairs, int batch,
                            int epochs, double lr);

  // ---- building blocks of the preconditioned CG (pcg.hip) ----
  // Planar FP32 / binary16 work fields for n points (slot-owned, grown on demand)
  void* work(int slot, size_t bytes);

 private:
  void upload_tables();
  Dirac& D_;
  HornerParams p_;
  float4* hrow_ = nullptr;
  HopRow16* hrow16_ = nullptr;
  float* invc_ = nullptr;
  float* kt_ = nullptr;
  static constexpr int kWorkSlots = 12;
  void* work_[kWorkSlots] = {};
  size_t work_bytes_[kWorkSlots] = {};
  double* red_ = nullptr;  // reduction partials of the gradient kernels
  size_t red_n_ = 0;
};

// Flexible CG (Notay's FCG(1)) on D^dagger D preconditioned by M = K K^dagger with the
// binary16-stored inner fields of the half-storage CG (FP32 arithmetic, FP64 reliable
// updates and true-residual stop), the network applied in the solver's temporal gauge.
CGResult pcg_solve(Horner& K, Spin<double>* x, const Spin<double>* b, const CGOptions& o,
                   KernelTimes* layer_times = nullptr);

// out = K in (which = 0), K^dagger in (1) or K K^dagger in (2, the solver's sequence with K's
// first layer fused into K^dagger's last) with the solver's kernels: the input rounded to
// binary16 in the temporal gauge, the parallel-in-time layers, the output rotated back to
// the original gauge in FP64 (for checking those kernels).
void horner_apply_solver(Horner& K, int which, Spin<double>* out, const Spin<double>* in);

}  // namespace qed3
