// Batched half-storage CG over systems of different shapes (see batch.hpp).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>

#include "batch.hpp"
#include "cg_epi.hpp"
#include "stencil.hpp"

namespace qed3 {
namespace {

using namespace cgepi;

constexpr int kBlock = 256;
constexpr long kLargeBlocks = 256;  // one block per CU of a 256-CU part

// One system of the batch: its operator views, LDS halo tables for ranges of kBlock
// sites, its blocks in the stencil grids and the streaming grids, and its fields.
struct Sys {
  DiracView<float, phase16_t> vh;
  DiracView<double> v64;
  HaloView hv;
  int ns;                  // neighbour slots (3 dual, 6 simplicial)
  int nxr, nxr_pad;        // site ranges of kBlock, padded to a multiple of 8
  int ts32, nst32;         // time tiling of the FP32-arithmetic (binary16-field) stencils
  int ts64, nst64;         // ... of the FP64 reliable-update stencils
  int blk32, nblk32;       // first block / blocks in the binary16 stencil grids
  int blk64, nblk64;       // ... in the FP64 stencil grids
  int sblk, nsblk;         // ... in the streaming grids
  size_t n;                // points
  double inv_ncomp;        // 1 / (4 n)
  SpinH* r;
  SpinH* w;
  SpinH* ring[kXGroup];    // search directions, slot k % kXGroup for iteration k
  Spin<float>* xin;        // x accumulated since the last reliable update (scaled)
  Spin<float>* rt;         // true residual at the outgoing scale (reliable update)
  Spin<double>* x;
  Spin<double>* t;
  const Spin<double>* b;
  CGScalars* sc;
  unsigned int* counter;
};

// index of the system owning global block b (first[k] ascending, first[0] = 0)
__device__ __forceinline__ int sys_of(const int* first, int nsys, int b) {
  int lo = 0, hi = nsys - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first[mid] <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

enum Phase { kK1 = 0, kK2 = 1, kU2 = 2, kU3 = 3 };

template <int PH, int POS>
struct PhaseEpi;  // the epilogue of each phase, built from a system

template <int POS>
struct PhaseEpi<kK1, POS> {
  __device__ static EpiK1PH make(const Sys& S) { return EpiK1PH{S.w, S.ring[POS], S.sc}; }
};
template <int POS>
struct PhaseEpi<kK2, POS> {
  __device__ static EpiK2H<POS> make(const Sys& S) {
    EpiK2H<POS> e{};
    e.x = S.xin;
    e.r = S.r;
    for (int j = 0; j < kXGroup; j++) e.ring[j] = S.ring[j];
    e.sc = S.sc;
    return e;
  }
};
template <int POS>
struct PhaseEpi<kU2, POS> {
  __device__ static EpiStoreIfRU make(const Sys& S) { return EpiStoreIfRU{S.t, S.sc}; }
};
template <int POS>
struct PhaseEpi<kU3, POS> {
  __device__ static EpiResidH make(const Sys& S) { return EpiResidH{S.b, S.rt, S.sc, S.inv_ncomp, 1.0}; }
};

// One launch of a stencil phase for the whole batch.  Virtual block vb (the batch's
// blocks laid out system after system) finds its system, its (site range, time tile) in
// that system's XCD-aware tile map, and runs the LDS stencil block with the phase's
// source and epilogue; a finished system's blocks exit at once.  The K1/K2 launches have
// one block per virtual block (vb = blockIdx.x); the reliable-update phases, idle in most
// iterations, run a capped grid that strides over the virtual blocks, so an idle phase
// does not dispatch a block per tile of the whole batch.
template <int PH, int POS, int NS>
__global__ __launch_bounds__(kBlock) void k_batch_stencil(const Sys* __restrict__ sys, const int* first,
                                                          int nsys, int vtotal, double2* partials) {
  extern __shared__ unsigned char smem_raw[];
  constexpr bool F64 = PH == kU2 || PH == kU3;
  constexpr bool DAG = PH == kK2 || PH == kU3;
  // K1/K2: one block per virtual block, no loop (fewer live registers)
  const int vstride = F64 ? int(gridDim.x) : vtotal;
  for (int vb = blockIdx.x; vb < vtotal; vb += vstride) {
    const int k = sys_of(first, nsys, vb);
    const Sys& S = sys[k];
    const int blk0 = F64 ? S.blk64 : S.blk32;
    const int lb = vb - blk0;
    TileMap tm;
    tm.nxr = S.nxr;
    tm.nxr_pad = S.nxr_pad;
    tm.nst = F64 ? S.nst64 : S.nst32;
    tm.remap = 1;  // every system's first block is a multiple of 8, so lb % 8 is the XCD
    int xr, st;
    tile_of(tm, lb, xr, st);
    auto epi = PhaseEpi<PH, POS>::make(S);
    if (epi.skip()) continue;  // uniform over the block: one system
    epi.begin();
    double2 acc = {0.0, 0.0};
    // the system's views and halo tables copied out of global memory once, so the slice
    // loop keeps them in registers as the single-system kernel keeps its arguments
    const HaloView hv = S.hv;
    const int ts = F64 ? S.ts64 : S.ts32;
    const int nxr = S.nxr;
    auto run = [&](const auto v, const auto& in, auto tag) {
      using T = decltype(tag);
      Spin<T>* tile = reinterpret_cast<Spin<T>*>(smem_raw);
      with_gauge(v, [&](auto tg) {
        stencil_lds_block<T, NS, DAG, decltype(tg)::value>(v, in, epi, ts, xr, nxr, st, hv, tile, acc);
      });
    };
    if constexpr (PH == kK1)
      run(S.vh, SrcPUpdateH{S.r, S.ring[(POS + kXGroup - 1) % kXGroup], S.sc}.bind(), float{});
    if constexpr (PH == kK2) run(S.vh, SrcHalf{S.w}, float{});
    if constexpr (PH == kU2) run(S.v64, SrcPlain<double>{S.x}, double{});
    if constexpr (PH == kU3) run(S.v64, SrcPlain<double>{S.t}, double{});
    if constexpr (decltype(epi)::kReduce) store_partial(acc, partials + blk0, unsigned(lb));
  }
}

// one block per system: sum that system's partials in block order and run its finaliser
template <int PH, int POS>
__global__ __launch_bounds__(kBlock) void k_batch_finalize(const Sys* __restrict__ sys, const double2* partials) {
  const Sys& S = sys[blockIdx.x];
  constexpr bool F64 = PH == kU2 || PH == kU3;
  const int blk0 = F64 ? S.blk64 : S.blk32;
  const int nb = F64 ? S.nblk64 : S.nblk32;
  const auto epi = PhaseEpi<PH, POS>::make(S);
  if (epi.skip()) return;
  double2 s = {0.0, 0.0};
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    const double2 p = partials[blk0 + i];
    s.x += p.x;
    s.y += p.y;
  }
  const double2 tot = block_sum2(s);
  if (threadIdx.x == 0) epi(tot);
}

// Streaming phases over each system's points: 0 start (x = 0, x_inner = 0, zero p_old,
// |b|^2 and the solver scalars), 1 r = b / scale, 2 flush when an update is due, 3 rescale
// after an update, 4 final flush.  A capped grid strides over the virtual blocks (each
// system's nsblk blocks), as for the reliable-update stencils.
template <int PH>
__global__ __launch_bounds__(kBlock) void k_batch_stream(const Sys* __restrict__ sys, const int* first,
                                                         int nsys, int vtotal, double2* partials,
                                                         double tol2, double delta2, int maxiter) {
  for (int vb = blockIdx.x; vb < vtotal; vb += gridDim.x) {
    const int k = sys_of(first, nsys, vb);
    const Sys& S = sys[k];
    CGScalars* sc = S.sc;
    const int lb = vb - S.sblk;
    const size_t stride = size_t(S.nsblk) * blockDim.x;
    const size_t i0 = size_t(lb) * blockDim.x + threadIdx.x;
    if constexpr (PH == 0) {
      double2 acc = {0.0, 0.0};
      for (size_t i = i0; i < S.n; i += stride) {
        const Spin<double> bb = S.b[i];
        Spin<double> z;
        z.a = {0.0, 0.0};
        z.b = {0.0, 0.0};
        S.x[i] = z;
        Spin<float> zf;
        zf.a = {0.f, 0.f};
        zf.b = {0.f, 0.f};
        S.xin[i] = zf;
        S.ring[kXGroup - 1][i] = to_half(zf);
        acc.x += norm2(bb);
      }
      const double ncomp = 4.0 * double(S.n);
      grid_reduce_part(acc, partials + S.sblk, S.counter, unsigned(lb), unsigned(S.nsblk),
                       [=](double2 t) { half_start(sc, t.x, ncomp, tol2, delta2, maxiter); });
    } else if constexpr (PH == 1) {
      if (sc->done != 0) continue;
      const double inv = 1.0 / sc->scale;
      for (size_t i = i0; i < S.n; i += stride) half_store_b(S.b, S.r, inv, i);
    } else if constexpr (PH == 2 || PH == 4) {
      if (PH == 2 && (sc->done != 0 || sc->need_ru == 0)) continue;
      const double s = sc->scale;
      float ap[kXGroup - 1];
      PendRing pend;
      for (int j = 0; j < kXGroup - 1; j++) {
        ap[j] = float(sc->alpha_pend[j]);
        pend.p[j] = S.ring[j];
      }
      for (size_t i = i0; i < S.n; i += stride) half_flush_point(S.x, S.xin, pend, ap, s, i);
    } else {
      if (sc->done != 0 || sc->rescale == 0) continue;
      const float f = float(sc->rfac);
      for (size_t i = i0; i < S.n; i += stride) half_rescale_point(S.rt, S.r, f, i);
    }
  }
}

}  // namespace

namespace {
std::vector<CGResult> batch_launches(const std::vector<Dirac*>& Ds, const std::vector<Spin<double>*>& xs,
                                     const std::vector<const Spin<double>*>& bs, const CGOptions& o);
}  // namespace

// Systems whose own K1 grid covers the GPU (at least kLargeBlocks blocks of 256 at their
// time tile) run the single-system kernels, which run at lower register pressure than the
// batched kernels and in the temporal gauge, in rounds of one chunk per system on the
// systems' own streams (cg_solve_half_many: no GPU idle time on host polls, and systems
// that do not fill the GPU overlap); the others, too small for a launch of their own to
// occupy the GPU, share the batched launches.  On the 64 systems of n in {32..128},
// Lt in {16..64} the rounds were 12% faster than solving one system after another, the
// batched launches 2% slower.
std::vector<CGResult> batch_cg(const std::vector<Dirac*>& Ds, const std::vector<Spin<double>*>& xs,
                               const std::vector<const Spin<double>*>& bs, const CGOptions& o) {
  const int ns = int(Ds.size());
  if (ns == 0 || xs.size() != Ds.size() || bs.size() != Ds.size())
    throw std::invalid_argument("batch: sizes differ");
  for (Dirac* D : Ds)
    if (!D->has_gauge()) throw std::invalid_argument("batch: a system has no gauge field");
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<int> big, small;
  for (int k = 0; k < ns; k++) {
    const Dirac& D = *Ds[k];
    const int ts = std::max(1, std::min(D.tile<float>(false).ts, D.Lt()));
    const long blocks = long((D.V() + kBlock - 1) / kBlock) * ((D.Lt() + ts - 1) / ts);
    (blocks >= kLargeBlocks ? big : small).push_back(k);
  }
  std::vector<CGResult> out(ns);
  CGOptions os = o;
  os.mixed = os.half = true;
  if (!big.empty()) {
    std::vector<Dirac*> d;
    std::vector<Spin<double>*> x;
    std::vector<const Spin<double>*> b;
    for (int k : big) {
      d.push_back(Ds[k]);
      x.push_back(xs[k]);
      b.push_back(bs[k]);
    }
    const std::vector<CGResult> r = cg_solve_half_many(d, x, b, os);
    for (size_t j = 0; j < big.size(); j++) out[big[j]] = r[j];
  }
  if (!small.empty()) {
    std::vector<Dirac*> d;
    std::vector<Spin<double>*> x;
    std::vector<const Spin<double>*> b;
    for (int k : small) {
      d.push_back(Ds[k]);
      x.push_back(xs[k]);
      b.push_back(bs[k]);
    }
    const std::vector<CGResult> r = batch_launches(d, x, b, o);
    for (size_t j = 0; j < small.size(); j++) out[small[j]] = r[j];
  }
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (auto& r : out) r.seconds = secs;
  return out;
}

namespace {
std::vector<CGResult> batch_launches(const std::vector<Dirac*>& Ds, const std::vector<Spin<double>*>& xs,
                                     const std::vector<const Spin<double>*>& bs, const CGOptions& o) {
  const auto t0 = std::chrono::steady_clock::now();
  const int ns = int(Ds.size());
  if (ns == 0 || xs.size() != Ds.size() || bs.size() != Ds.size())
    throw std::invalid_argument("batch: sizes differ");
  for (Dirac* D : Ds)
    if (!D->has_gauge()) throw std::invalid_argument("batch: a system has no gauge field");
  hipStream_t st = Ds[0]->stream();
  // Systems are grouped by neighbour count (one kernel instantiation per group, so a
  // trivalent system does not run at the registers of the six-neighbour stencil) and,
  // within a group, ordered by descending time tile, so the longest blocks are dispatched
  // first and the grid's tail is short.
  auto ts_of = [](const Dirac& D, bool f64) {
    return std::max(1, std::min(f64 ? D.tile<double>(false).ts : D.tile<float>(false).ts, D.Lt()));
  };
  std::vector<int> order(ns);
  for (int k = 0; k < ns; k++) order[k] = k;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    const int na = Ds[a]->tables().max_nn, nb = Ds[b]->tables().max_nn;
    if (na != nb) return na < nb;
    return ts_of(*Ds[a], false) > ts_of(*Ds[b], false);
  });
  std::vector<void*> owned;
  auto dalloc = [&](size_t bytes) {
    void* p = nullptr;
    QED3_HIP(hipMalloc(&p, bytes));
    owned.push_back(p);
    return p;
  };
  CGScalars* scal = static_cast<CGScalars*>(dalloc(ns * sizeof(CGScalars)));
  unsigned int* ctr = static_cast<unsigned int*>(dalloc(ns * sizeof(unsigned int)));
  struct Group {
    int nn, first_sys = 0, nsys = 0, tot32 = 0, tot64 = 0, tots = 0;
  };
  std::vector<Group> groups;
  std::vector<Sys> sys(ns);
  std::vector<int> f32(ns), f64(ns), fs(ns);
  size_t smem32 = 0, smem64 = 0;
  int maxtot = 1;
  for (int j = 0; j < ns; j++) {
    Dirac& D = *Ds[order[j]];
    const int nn = D.tables().max_nn;
    if (groups.empty() || groups.back().nn != nn) {
      Group g;
      g.nn = nn;
      g.first_sys = j;
      groups.push_back(g);
    }
    Group& g = groups.back();
    g.nsys++;
    Sys& S = sys[j];
    S.vh = D.view_half_phases();
    S.v64 = D.view<double>();
    S.hv = D.halo(kBlock);
    S.ns = nn;
    S.n = D.points();
    S.inv_ncomp = 1.0 / (4.0 * double(S.n));
    S.nxr = (D.V() + kBlock - 1) / kBlock;
    S.nxr_pad = (S.nxr + 7) / 8 * 8;
    S.ts32 = ts_of(D, false);
    S.ts64 = ts_of(D, true);
    S.nst32 = (D.Lt() + S.ts32 - 1) / S.ts32;
    S.nst64 = (D.Lt() + S.ts64 - 1) / S.ts64;
    S.blk32 = f32[j] = g.tot32;
    S.nblk32 = S.nxr_pad * S.nst32;
    g.tot32 += S.nblk32;
    S.blk64 = f64[j] = g.tot64;
    S.nblk64 = S.nxr_pad * S.nst64;
    g.tot64 += S.nblk64;
    S.sblk = fs[j] = g.tots;
    S.nsblk = int(std::min<size_t>((S.n + kBlock - 1) / kBlock, 2048));
    g.tots += S.nsblk;
    maxtot = std::max({maxtot, g.tot32, g.tot64, g.tots});
    smem32 = std::max(smem32, lds_bytes<float>(kBlock, S.hv.hmax));
    smem64 = std::max(smem64, lds_bytes<double>(kBlock, S.hv.hmax));
    const size_t hb = S.n * sizeof(SpinH), fb = S.n * sizeof(Spin<float>);
    S.r = static_cast<SpinH*>(dalloc(hb));
    S.w = static_cast<SpinH*>(dalloc(hb));
    for (int i = 0; i < kXGroup; i++) S.ring[i] = static_cast<SpinH*>(dalloc(hb));
    S.xin = static_cast<Spin<float>*>(dalloc(fb));
    S.rt = static_cast<Spin<float>*>(dalloc(fb));
    S.t = static_cast<Spin<double>*>(dalloc(S.n * sizeof(Spin<double>)));
    S.x = xs[order[j]];
    S.b = bs[order[j]];
    S.sc = scal + j;
    S.counter = ctr + j;
  }
  if (size_t(maxtot) > kReduceCapacity) throw std::invalid_argument("batch: grid exceeds the reduction scratch");
  QED3_HIP(hipMemsetAsync(ctr, 0, ns * sizeof(unsigned int), st));
  Sys* dsys = static_cast<Sys*>(dalloc(ns * sizeof(Sys)));
  int* d32 = static_cast<int*>(dalloc(ns * sizeof(int)));
  int* d64 = static_cast<int*>(dalloc(ns * sizeof(int)));
  int* dss = static_cast<int*>(dalloc(ns * sizeof(int)));
  // all launches are in order on one stream, so the groups share the partials buffer
  double2* part = static_cast<double2*>(dalloc(size_t(maxtot) * sizeof(double2)));
  std::vector<double2*> partials(groups.size(), part);
  QED3_HIP(hipMemcpyAsync(dsys, sys.data(), ns * sizeof(Sys), hipMemcpyHostToDevice, st));
  QED3_HIP(hipMemcpyAsync(d32, f32.data(), ns * sizeof(int), hipMemcpyHostToDevice, st));
  QED3_HIP(hipMemcpyAsync(d64, f64.data(), ns * sizeof(int), hipMemcpyHostToDevice, st));
  QED3_HIP(hipMemcpyAsync(dss, fs.data(), ns * sizeof(int), hipMemcpyHostToDevice, st));
  const double tol2 = o.tol * o.tol, delta2 = o.delta * o.delta;
  // capped grids for the phases that are idle in most iterations (8 blocks per CU)
  const int cap = 2048;
  auto stream = [&](auto kern) {
    for (size_t g = 0; g < groups.size(); g++) {
      const Group& G = groups[g];
      hipLaunchKernelGGL(kern, dim3(std::min(G.tots, cap)), dim3(kBlock), 0, st, dsys + G.first_sys,
                         dss + G.first_sys, G.nsys, G.tots, partials[g], tol2, delta2, o.maxiter);
    }
  };
  stream(k_batch_stream<0>);
  stream(k_batch_stream<1>);
  QED3_HIP(hipGetLastError());
  auto group_phase = [&](auto pos_tag, auto nn_tag, size_t g) {
    constexpr int POS = decltype(pos_tag)::value;
    constexpr int NN = decltype(nn_tag)::value;
    const Group& G = groups[g];
    Sys* gs = dsys + G.first_sys;
    hipLaunchKernelGGL((k_batch_stencil<kK1, POS, NN>), dim3(G.tot32), dim3(kBlock), smem32, st, gs,
                       d32 + G.first_sys, G.nsys, G.tot32, partials[g]);
    hipLaunchKernelGGL((k_batch_finalize<kK1, POS>), dim3(G.nsys), dim3(kBlock), 0, st, gs, partials[g]);
    hipLaunchKernelGGL((k_batch_stencil<kK2, POS, NN>), dim3(G.tot32), dim3(kBlock), smem32, st, gs,
                       d32 + G.first_sys, G.nsys, G.tot32, partials[g]);
    hipLaunchKernelGGL((k_batch_finalize<kK2, POS>), dim3(G.nsys), dim3(kBlock), 0, st, gs, partials[g]);
  };
  auto group_update = [&](auto nn_tag, size_t g) {
    constexpr int NN = decltype(nn_tag)::value;
    const Group& G = groups[g];
    Sys* gs = dsys + G.first_sys;
    hipLaunchKernelGGL((k_batch_stencil<kU2, 0, NN>), dim3(std::min(G.tot64, cap)), dim3(kBlock), smem64, st,
                       gs, d64 + G.first_sys, G.nsys, G.tot64, partials[g]);
    hipLaunchKernelGGL((k_batch_stencil<kU3, 0, NN>), dim3(std::min(G.tot64, cap)), dim3(kBlock), smem64, st,
                       gs, d64 + G.first_sys, G.nsys, G.tot64, partials[g]);
    hipLaunchKernelGGL((k_batch_finalize<kU3, 0>), dim3(G.nsys), dim3(kBlock), 0, st, gs, partials[g]);
  };
  auto phase = [&](auto pos_tag) {
    for (size_t g = 0; g < groups.size(); g++) {
      if (groups[g].nn == 6)
        group_phase(pos_tag, std::integral_constant<int, 6>{}, g);
      else
        group_phase(pos_tag, std::integral_constant<int, 3>{}, g);
    }
    stream(k_batch_stream<2>);
    for (size_t g = 0; g < groups.size(); g++) {
      if (groups[g].nn == 6)
        group_update(std::integral_constant<int, 6>{}, g);
      else
        group_update(std::integral_constant<int, 3>{}, g);
    }
    stream(k_batch_stream<3>);
  };
  static_assert(kXGroup == 4, "one phase instantiation per group position");
  auto one_iter = [&](int i) {
    switch (i % kXGroup) {
      case 0: phase(std::integral_constant<int, 0>{}); break;
      case 1: phase(std::integral_constant<int, 1>{}); break;
      case 2: phase(std::integral_constant<int, 2>{}); break;
      default: phase(std::integral_constant<int, 3>{}); break;
    }
  };
  int chunk = o.chunk > 0 ? o.chunk : 8;
  chunk = (chunk + kXGroup - 1) / kXGroup * kXGroup;  // every replay starts a new group
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (o.use_graph && std::getenv("QED3_NO_GRAPH") == nullptr) {
    QED3_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < chunk; i++) one_iter(i);
    QED3_HIP(hipStreamEndCapture(st, &graph));
    QED3_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  }
  std::vector<CGScalars> hs(ns);
  const int max_chunks = o.maxiter / chunk + 4;
  for (int c = 0; c < max_chunks; c++) {
    if (exec)
      QED3_HIP(hipGraphLaunch(exec, st));
    else
      for (int i = 0; i < chunk; i++) one_iter(i);
    QED3_HIP(hipMemcpyAsync(hs.data(), scal, ns * sizeof(CGScalars), hipMemcpyDeviceToHost, st));
    QED3_HIP(hipStreamSynchronize(st));
    bool all = true;
    for (auto& s : hs) all = all && s.done != 0;
    if (all) break;
  }
  stream(k_batch_stream<4>);
  QED3_HIP(hipMemcpyAsync(hs.data(), scal, ns * sizeof(CGScalars), hipMemcpyDeviceToHost, st));
  QED3_HIP(hipStreamSynchronize(st));
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  for (void* p : owned) (void)hipFree(p);
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::vector<CGResult> out(ns);
  for (int j = 0; j < ns; j++) {
    const int k = order[j];
    out[k].iterations = hs[j].iter;
    out[k].reliable_updates = hs[j].n_ru;
    out[k].status = hs[j].done == 1 ? 0 : (hs[j].done == 2 ? 4 : 3);
    out[k].true_resid = true_residual(*Ds[k], xs[k], bs[k]);
    Ds[k]->release_scratch();  // many contexts: do not keep a residual's temporaries on each
    if (out[k].status == 0 && !(out[k].true_resid <= o.tol * 1.0000001)) out[k].status = 3;
    out[k].seconds = secs;
  }
  return out;
}
}  // namespace

}  // namespace qed3
