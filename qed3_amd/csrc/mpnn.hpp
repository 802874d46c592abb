// Learned preconditioner: a gauge-covariant linear message-passing network on the
// lattice graph of a Dirac operator (new component; the reference has no ML part).
//
// Feature maps of a field h (any number of complex channels, two spin components) at
// (s,x), spin a -- each commutes with U(1) gauge and local spin-frame rotations:
//   F0 = h(x)_a
//   F1 = sum_y tau_xy,a h(y)_a,      tau_xy,a = e^{i theta_s(x,y)} kappa_xy/2 e^{-+i omega_xy/2}
//   F2 = sum_y (gamma_xy tau_xy h(y))_a  (gamma_xy = cos(alpha_xy) s1 + sin(alpha_xy) s2)
//   F3 = b+ kappa_t(x)/2 e^{+i theta^t_s(x)} h(s+1,x),  F4 = b- kappa_t(x)/2 e^{-i theta^t_{s-1}(x)} h(s-1,x)
// Every weight below may differ between the two output spin components a: a diagonal
// spin matrix commutes with the frame rotations exp(i phi sigma3/2), so covariance holds,
// and with it the net can represent the spin projectors (-r -+ sigma3)/2 of the time
// hops, hence D and D^dagger themselves.
// Network (C = 16 complex channels, L >= 2 layers), at output spin a:
//   h1_a   = sum_k v_a,k (x) F_k(r)_a                   input head, 1 -> C
//   h^l_a  = sum_k W^l_a,k F_k(h^{l-1})_a, l = 2..L-1   C -> C channel mixing on BF16 MFMA
//   out_a  = sum_k sum_c u_a,k[c] F_k(h^{L-1})_c,a + c0 r_a   output head, C -> 1, plus skip
// The net is trained as an approximate inverse of D (pairs (D v, v)).  The skip
// c0 = 1/B (B a Gershgorin bound on |D|_2, so c0 <= 1/sigma_max(D)) gives every component
// at least a scaled-identity term; the layers learn the rest.  Linear in r (so
// scale-equivariant); flexible CG on D^dag D uses M = N N^dagger.  Parameters are stored
// as FP32 complex masters; the mixing matrices are rounded to BF16 for the MFMA tiles.
#pragma once
#include <vector>

#include "dirac.hpp"

namespace qed3 {

constexpr int kNetC = 16;

struct NetParamsView;

class Net {
 public:
  Net(Dirac& D, int layers, int seed);
  ~Net();
  Net(const Net&) = delete;
  Net& operator=(const Net&) = delete;

  int layers() const { return L_; }
  size_t nparams() const { return 20 * kNetC + size_t(L_ - 2) * 10 * kNetC * kNetC; }  // complex
  void get_params(double* p) const;  // interleaved complex
  void set_params(const double* p);
  double skip() const { return skip_; }
  void set_skip(double c) { skip_ = float(c); }

  // out = N(r) for B fields of D.points() FP32 spinors each (device)
  void apply(Spin<float>* out, const Spin<float>* r, int B);
  // out = N^dagger(r) (same layout and cost as apply)
  void apply_adj(Spin<float>* out, const Spin<float>* r, int B);
  // loss = (1/B) sum_b |N(r_b) - e_b|^2/|e_b|^2 (N including the skip term) and its gradient (d/dRe + i d/dIm)
  double loss_grad(const Spin<float>* r, const Spin<float>* e, int B, std::vector<double>& grad);
  // Adam over `npairs` device pairs in mini-batches of B; returns the last epoch's mean loss
  double train(const Spin<float>* r, const Spin<float>* e, int npairs, int B, int epochs,
               double lr, std::vector<double>* history);

 private:
  void upload_params();
  void ensure(int B, bool training);
  Dirac& D_;
  int L_;
  std::vector<double> p_;  // interleaved complex master parameters
  float skip_ = 0.f;       // c0 of the skip term
  // device
  float* slot_ = nullptr;      // [M][V] float4 (C, S, cos a_xy, sin a_xy)
  float* slot_nb_ = nullptr;   // [M][V] float2 (cos a_yx, sin a_yx)
  float* kt_ = nullptr;        // [V] kappa_t(x)/2
  float* heads_ = nullptr;     // v, u and the adjoint network's two heads, [2][5][C] complex FP32 each
  void* mixA_ = nullptr;       // per middle layer and output spin: R (32 x 160) bf16
  void* mixAT_ = nullptr;      // per middle layer and output spin: R^T (160 x 32) bf16
  void* mixAH_ = nullptr;      // adjoint network: per layer and output spin (32 x 160) bf16
  void* act_ = nullptr;        // activations h1..h^{L-1}: [L-1][B*points][2][32] bf16
  float* gfield_ = nullptr;    // gradient field [B*points][2][32] FP32 (two buffers)
  float* gfield2_ = nullptr;
  float* G_ = nullptr;         // [B*points][5][2][32] FP32
  Spin<float>* out_ = nullptr;
  Spin<float>* gout_ = nullptr;
  double* partial_ = nullptr;  // per-block weight-gradient partials
  double* red_ = nullptr;
  int capB_ = 0;
  bool cap_train_ = false;
  // Adam state
  std::vector<double> m_, v_;
  long step_ = 0;
};

}  // namespace qed3
