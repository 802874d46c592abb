// qed3_cg_noalloc: the FP64 CG of qed3_cg for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// Its work fields live in `scratch`, a file-scope device array of kNoAllocBytes that every
// GPU the library runs on holds until the process ends.  In host code the array's name is a
// host-side stand-in, not a device address: the device address is looked up once per device
// with hipGetSymbolAddress (with that device current) and only that address is handed to
// kernels.  Every call on a GPU shares that GPU's array and overwrites it (the solver
// initialises every field it reads); calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <array>
#include <cmath>
#include <mutex>
#include <string>

#include "../../include/qed3.h"
#include "cg.hpp"
#include "dirac.hpp"

namespace qed3 {
namespace {

constexpr size_t kNoAllocBytes = size_t(64) << 20;  // 64 MiB per GPU
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
std::array<void*, kMaxDevices> g_ptr{};

// device address of this device's scratch (looked up on the first call on the device)
void* scratch_ptr(std::string* err) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    *err = "no current HIP device";
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return nullptr;
    }
    g_ptr[dev] = p;
  }
  return g_ptr[dev];
}

}  // namespace

// used by the C entry point below; errors are returned as text
int cg_noalloc(Dirac& D, void* d_x, const void* d_b, double tol, int maxiter, CGResult* res, std::string* err) {
  const size_t need = cg_workspace_bytes(D.points());
  if (need > kNoAllocBytes) {
    size_t pmax = kNoAllocBytes / 192;  // six FP64 fields per point, less the alignment slack
    while (cg_workspace_bytes(pmax) > kNoAllocBytes) --pmax;
    *err = "qed3_cg_noalloc: the lattice (" + std::to_string(D.points()) + " points, V*Lt) needs " +
           std::to_string(need) + " bytes of work memory, more than the " + std::to_string(kNoAllocBytes) +
           " bytes the library holds per GPU (at most " + std::to_string(pmax) + " points); use qed3_cg";
    return QED3_ERR_ARG;
  }
  void* ws = scratch_ptr(err);
  if (!ws) return QED3_ERR_HIP;
  CGOptions o;
  o.tol = tol;
  o.maxiter = maxiter;
  *res = cg_solve_f64_workspace(D, static_cast<Spin<double>*>(d_x), static_cast<const Spin<double>*>(d_b), o, ws,
                                kNoAllocBytes);
  return QED3_OK;
}

}  // namespace qed3
