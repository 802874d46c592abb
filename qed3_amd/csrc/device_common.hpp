// Device-side types and helpers shared by the gfx950 kernels: 2-component complex
// spinors, complex arithmetic, wave64 + grid reductions with a last-block finaliser.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>

namespace qed3 {

#define QED3_HIP(call)                                                                 \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      throw ::qed3::HipError(std::string(#call) + ": " + hipGetErrorString(e_) + " @" + \
                             __FILE__ + ":" + std::to_string(__LINE__));               \
  } while (0)

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

template <typename T>
using cplx_t = typename std::conditional<std::is_same<T, double>::value, double2, float2>::type;
template <typename T>
using real4_t = typename std::conditional<std::is_same<T, double>::value, double4, float4>::type;

// One lattice point: the two spin components at (s, x).
template <typename T>
struct alignas(4 * sizeof(T)) Spin {
  cplx_t<T> a, b;
};

template <typename T>
__device__ __forceinline__ cplx_t<T> cmul(cplx_t<T> u, cplx_t<T> v) {
  return {u.x * v.x - u.y * v.y, u.x * v.y + u.y * v.x};
}
template <typename T>
__device__ __forceinline__ cplx_t<T> cconjmul(cplx_t<T> u, cplx_t<T> v) {  // conj(u) v
  return {u.x * v.x + u.y * v.y, u.x * v.y - u.y * v.x};
}
template <typename T>
__device__ __forceinline__ cplx_t<T> cadd(cplx_t<T> u, cplx_t<T> v) {
  return {u.x + v.x, u.y + v.y};
}
template <typename T>
__device__ __forceinline__ cplx_t<T> cscale(T s, cplx_t<T> v) {
  return {s * v.x, s * v.y};
}

// Spin loads/stores as whole 16 B / 32 B vectors.
template <typename T>
__device__ __forceinline__ Spin<T> ld(const Spin<T>* p) {
  return *p;
}
template <typename T>
__device__ __forceinline__ void st(Spin<T>* p, const Spin<T>& v) {
  *p = v;
}

// One lattice point stored as four IEEE binary16 values (8 B), computed on in FP32.  The
// half-storage CG keeps its inner fields in this form, scaled to unit rms per component.
typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef float float4v_t __attribute__((ext_vector_type(4)));
struct alignas(8) SpinH {
  half4_t v;
};

__device__ __forceinline__ Spin<float> from_half(const SpinH& h) {
  const float4v_t f = __builtin_convertvector(h.v, float4v_t);
  Spin<float> s;
  s.a = {f.x, f.y};
  s.b = {f.z, f.w};
  return s;
}
// round to nearest even; values beyond the binary16 range become +-inf
__device__ __forceinline__ SpinH to_half(const Spin<float>& s) {
  const float4v_t f = {s.a.x, s.a.y, s.b.x, s.b.y};
  SpinH h;
  h.v = __builtin_convertvector(f, half4_t);
  return h;
}
__device__ __forceinline__ Spin<float> ldh(const SpinH* p) { return from_half(*p); }
// Streaming store of a binary16 spinor (the CG's p, w and r: written once per kernel and
// read by the next one), marked non-temporal so it does not displace the halo and
// neighbour spinors that the stencil re-reads from L2.
__device__ __forceinline__ void sth_nt(SpinH* p, const SpinH& h) {
  unsigned long long v;
  __builtin_memcpy(&v, &h, 8);
  __builtin_nontemporal_store(v, reinterpret_cast<unsigned long long*>(p));
}

// Link phases exp(i theta): stored as complex in the FP64 and FP32 tables.
__device__ __forceinline__ double2 ldu(const double2* p) { return *p; }
__device__ __forceinline__ float2 ldu(const float2* p) { return *p; }
// Link phase of the half-storage solver as a 16-bit fixed-point angle
// (theta = 2 pi v / 65536, 2 B per link, uniform resolution 9.6e-5 rad); the hardware sine
// and cosine take the angle in revolutions.
struct phase16_t {
  short v;
};
__device__ __forceinline__ float2 ldu(const phase16_t* p) {
  const float rev = float(p->v) * (1.0f / 65536.0f);
  return {__builtin_amdgcn_cosf(rev), __builtin_amdgcn_sinf(rev)};
}

// Link phase as a binary16 (cos, sin) pair (4 B per link): read without a sine or cosine,
// for kernels whose arithmetic on the phase is itself binary16
typedef _Float16 hvec2_t __attribute__((ext_vector_type(2)));
struct phasepair_t {
  hvec2_t v;
};
__device__ __forceinline__ float2 ldu(const phasepair_t* p) {
  const hvec2_t h = p->v;
  return {float(h.x), float(h.y)};
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block-wide sum of two doubles; result valid in thread 0.
__device__ __forceinline__ double2 block_sum2(double2 v) {
  __shared__ double2 sh[16];
  v.x = wave_sum(v.x);
  v.y = wave_sum(v.y);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double2 r = {0.0, 0.0};
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; i++) {
      r.x += sh[i].x;
      r.y += sh[i].y;
    }
  }
  return r;
}

// Scratch for single-pass grid reductions: one partial per block plus a ticket counter.
struct GridReduce {
  double2* partials;
  unsigned int* counter;
};

// Every block of the grid must call this exactly once (uniformly).  The last block to
// arrive sums the partials in block order (deterministic) and calls fin(total) from
// thread 0, then re-arms the counter for the next launch on the stream.
template <class Fin>
__device__ __forceinline__ void grid_reduce(double2 v, GridReduce gr, const Fin& fin) {
  __shared__ bool is_last;
  const double2 b = block_sum2(v);
  const unsigned nblocks = gridDim.x * gridDim.y;
  if (threadIdx.x == 0) {
    const unsigned bid = blockIdx.y * gridDim.x + blockIdx.x;
    __hip_atomic_store(&gr.partials[bid].x, b.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&gr.partials[bid].y, b.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    const unsigned t = atomicAdd(gr.counter, 1u);
    is_last = (t == nblocks - 1);
  }
  __syncthreads();
  if (is_last) {
    __threadfence();
    double2 s = {0.0, 0.0};
    for (unsigned i = threadIdx.x; i < nblocks; i += blockDim.x) {
      s.x += __hip_atomic_load(&gr.partials[i].x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s.y += __hip_atomic_load(&gr.partials[i].y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const double2 tot = block_sum2(s);
    if (threadIdx.x == 0) {
      fin(tot);
      __hip_atomic_store(gr.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// Two-launch reduction for the hot loops: every block stores its partial with a plain
// store; a one-block finaliser kernel launched next on the stream sums the partials in
// block order and runs the epilogue's finaliser.  The kernel boundary orders the
// stores, so no block issues a device-scope fence (on a multi-XCD part such a fence
// writes back the XCD's L2).
__device__ __forceinline__ void store_partial(double2 v, double2* partials, unsigned slot) {
  const double2 b = block_sum2(v);
  if (threadIdx.x == 0) partials[slot] = b;
}

template <class Fin>
__global__ __launch_bounds__(256) void k_finalize(const double2* __restrict__ partials, unsigned n,
                                                  Fin fin) {
  if (fin.skip()) return;
  double2 s = {0.0, 0.0};
  for (unsigned i = threadIdx.x; i < n; i += blockDim.x) {
    const double2 p = partials[i];
    s.x += p.x;
    s.y += p.y;
  }
  const double2 tot = block_sum2(s);
  if (threadIdx.x == 0) fin(tot);
}

// Same as grid_reduce for a sub-grid: the caller's block is number lb of nb blocks that
// share `partials` (nb slots) and `counter`.
template <class Fin>
__device__ __forceinline__ void grid_reduce_part(double2 v, double2* partials, unsigned int* counter,
                                                 unsigned lb, unsigned nb, const Fin& fin) {
  __shared__ bool is_last_p;
  const double2 b = block_sum2(v);
  if (threadIdx.x == 0) {
    __hip_atomic_store(&partials[lb].x, b.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&partials[lb].y, b.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    const unsigned t = atomicAdd(counter, 1u);
    is_last_p = (t == nb - 1);
  }
  __syncthreads();
  if (is_last_p) {
    __threadfence();
    double2 s = {0.0, 0.0};
    for (unsigned i = threadIdx.x; i < nb; i += blockDim.x) {
      s.x += __hip_atomic_load(&partials[i].x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s.y += __hip_atomic_load(&partials[i].y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const double2 tot = block_sum2(s);
    if (threadIdx.x == 0) {
      fin(tot);
      __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

}  // namespace qed3
