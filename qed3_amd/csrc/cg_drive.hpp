// Host-side driver pieces shared by the device-driven CG solvers (cg.hip, pcg.hip):
// grid sizes of the streaming kernels, the reported true residual, per-kernel event
// timing and the chunked graph replay with one host poll per chunk.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <functional>
#include <stdexcept>
#include <utility>
#include <vector>

#include "cg.hpp"
#include "cg_epi.hpp"

namespace qed3 {
namespace cgdrive {

inline size_t stream_blocks(size_t n) { return std::min<size_t>((n + 255) / 256, 8192); }
// Grid of the per-iteration kernels that only work after a reliable update (flush,
// rescale) and otherwise exit at once: a smaller grid makes the early exit cheaper and
// still has eight waves per SIMD for the updates themselves.
inline size_t gated_blocks(size_t n) { return std::min<size_t>((n + 255) / 256, 2048); }

// The FP64 true residual of the returned x.  A solve that converged did so in a reliable
// update, which computed |b - D^dag D x|^2 in FP64 for this same x (every pending inner
// term was flushed before it), so that value is reported; otherwise it is recomputed.
inline double final_residual(Dirac& D, const Spin<double>* x, const Spin<double>* b, const CGScalars& h) {
  if (h.done == 1 && h.n_ru > 0 && h.bnorm2 > 0) return std::sqrt(h.true_r2 / h.bnorm2);
  return true_residual(D, x, b);
}

// Event pairs for KernelTimes: pair(k) hands out a fresh (start, end) pair tagged with
// kernel class k; finish() averages the pairs of the first `iters` iterations.
class EventTimer {
 public:
  explicit EventTimer(KernelTimes* t) : t_(t) {}
  ~EventTimer() {
    for (auto& e : ev_) (void)hipEventDestroy(e.first[0]), (void)hipEventDestroy(e.first[1]);
  }
  const hipEvent_t* pair(int cls, int iter) {
    if (!t_) return nullptr;
    if (cls < 0 || cls >= KernelTimes::kClasses) throw std::logic_error("bad kernel class");
    ev_.emplace_back();
    auto& e = ev_.back();
    QED3_HIP(hipEventCreate(&e.first[0]));
    QED3_HIP(hipEventCreate(&e.first[1]));
    e.second = {cls, iter};
    return e.first.data();
  }
  void finish(int iters) {
    if (!t_) return;
    for (auto& e : ev_) {
      if (e.second.second >= iters) continue;  // launched after convergence: early exit
      float ms = 0;
      QED3_HIP(hipEventElapsedTime(&ms, e.first[0], e.first[1]));
      t_->ms[e.second.first] += ms;
      t_->count[e.second.first] += 1;
    }
    for (int k = 0; k < KernelTimes::kClasses; k++)
      if (t_->count[k]) t_->ms[k] /= t_->count[k];
  }

 private:
  KernelTimes* t_;
  std::vector<std::pair<std::array<hipEvent_t, 2>, std::pair<int, int>>> ev_;
};

// One solve's iterations in chunks: one_iter(i, k), i = 0 .. chunk-1 (the iteration's
// position in the chunk), k its index in the solve (-1 while a graph is captured), then
// tail(); the chunk is captured once as a HIP graph unless disabled.  launch() queues one
// chunk and a copy of the device scalars into hsc; after the stream has drained them,
// finished() tells whether the device marked the solve done (or the chunk budget ran out).
// Several loops on different streams can be launched before any host wait.
class ChunkLoop {
 public:
  ChunkLoop(std::function<void(int, int)> one_iter, std::function<void()> tail, const CGOptions& o, size_t n,
            hipStream_t st, const CGScalars* sc, CGScalars* hsc)
      : one_iter_(std::move(one_iter)), tail_(std::move(tail)), st_(st), sc_(sc), hsc_(hsc) {
    chunk_ = o.chunk > 0 ? o.chunk : (n > (1u << 20) ? 8 : 16);
    chunk_ = (chunk_ + cgepi::kXGroup - 1) / cgepi::kXGroup * cgepi::kXGroup;  // every replay starts a new group
    // every chunk runs at least one iteration (a reliable update ends each chunk), so the
    // device's iteration cap ends the solve within this many chunks
    max_chunks_ = o.maxiter + 4;
    if (o.use_graph && !o.times && std::getenv("QED3_NO_GRAPH") == nullptr) {
      QED3_HIP(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
      for (int i = 0; i < chunk_; i++) one_iter_(i, -1);
      tail_();
      QED3_HIP(hipStreamEndCapture(st_, &graph_));
      QED3_HIP(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
    }
  }
  ~ChunkLoop() {
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
  }
  ChunkLoop(const ChunkLoop&) = delete;
  ChunkLoop& operator=(const ChunkLoop&) = delete;
  hipStream_t stream() const { return st_; }
  void launch() {
    if (exec_) {
      QED3_HIP(hipGraphLaunch(exec_, st_));
    } else {
      for (int i = 0; i < chunk_; i++) one_iter_(i, launched_ * chunk_ + i);
      tail_();
    }
    QED3_HIP(hipGetLastError());
    QED3_HIP(hipMemcpyAsync(hsc_, sc_, sizeof(CGScalars), hipMemcpyDeviceToHost, st_));
    launched_++;
  }
  bool finished() const { return launched_ > 0 && (hsc_->done != 0 || launched_ >= max_chunks_); }

 private:
  std::function<void(int, int)> one_iter_;
  std::function<void()> tail_;
  hipStream_t st_;
  const CGScalars* sc_;
  CGScalars* hsc_;
  int chunk_ = 8, max_chunks_ = 1, launched_ = 0;
  hipGraphExec_t exec_ = nullptr;
  hipGraph_t graph_ = nullptr;
};

// One solve's chunks with one host poll per chunk, until the device marks the solve done.
// On return hsc holds the final scalars.
template <class Iter, class Tail>
void drive_chunks(const Iter& one_iter, const Tail& tail, const CGOptions& o, size_t n, hipStream_t st,
                  const CGScalars* sc, CGScalars* hsc) {
  ChunkLoop loop(one_iter, tail, o, n, st, sc, hsc);
  do {
    loop.launch();
    QED3_HIP(hipStreamSynchronize(st));
  } while (!loop.finished());
}

// Kernels of the half-storage start and flush, shared by the half-storage and the
// preconditioned solvers (one copy per translation unit).
namespace {
using cgepi::half_flush_point;
using cgepi::kXGroup;
using cgepi::PendRing;
// Half-storage start: x = 0, p_old = 0, |b|^2; the scale makes the stored residual
// unit-rms per component (k_store_b0 then stores r = b / scale).
__global__ __launch_bounds__(256) void k_init_h(const Spin<double>* __restrict__ b, Spin<double>* x64,
                                                Spin<float>* xin, SpinH* p, size_t n, GridReduce gr) {
  double2 acc = {0.0, 0.0};
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x) {
    const Spin<double> bb = b[i];
    Spin<float> zf;
    zf.a = {0.f, 0.f};
    zf.b = {0.f, 0.f};
    xin[i] = zf;
    p[i] = to_half(zf);
    Spin<double> z;
    z.a = {0.0, 0.0};
    z.b = {0.0, 0.0};
    x64[i] = z;
    acc.x += norm2(bb);
  }
  store_partial(acc, gr.partials, blockIdx.x);
}

// Half-storage flush: x64 += scale (x_inner + sum_j alpha_pend[j] ring[j]), x_inner = 0,
// taking the x terms the current group's iterations left pending (alpha_pend[j] = 0 if none)
__global__ __launch_bounds__(256) void k_flush_h(Spin<double>* x64, Spin<float>* xin, PendRing ring,
                                                 size_t n, const CGScalars* sc, int always,
                                                 const phase16_t* rot) {
  if (!always && (sc->done != 0 || sc->need_ru == 0)) return;
  const double s = sc->scale;
  float ap[kXGroup - 1];
  for (int j = 0; j < kXGroup - 1; j++) ap[j] = float(sc->alpha_pend[j]);
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x)
    half_flush_point(x64, xin, ring, ap, s, i, rot);
}

}  // namespace

}  // namespace cgdrive
}  // namespace qed3
