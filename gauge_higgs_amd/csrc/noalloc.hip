// ghx_rng_gauss_noalloc: ghx_rng_gauss for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc). The draws are staged in a file-scope device array instead
// of a buffer allocated per call; the entry point allocates nothing.
//
// Each GPU holds its own copy of `scratch` (kNoAllocMaxDraws doubles, 8 MiB)
// until the process ends. Host code never passes the array's name to a kernel
// or a copy: on the first call on a device, with that device current, its
// address is looked up once with hipGetSymbolAddress and kept per device.
// Every call on one GPU shares that GPU's array and leaves the last call's
// draws in it; calls that overlap on one GPU (from two host threads) race.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "../../include/ghx.h"
#include "rng_kernels.hpp"

namespace {

constexpr int64_t kNoAllocMaxDraws = int64_t(1) << 20;
static __device__ __align__(256) double scratch[kNoAllocMaxDraws];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};

// Device address of this device's `scratch`, looked up on its first use.
int scratch_ptr(double** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return GHX_E_HIP;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return GHX_E_HIP;
    g_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_ptr[dev];
  return GHX_OK;
}

}  // namespace

extern "C" int ghx_rng_gauss_noalloc(uint64_t seed, uint64_t key0, int64_t n, double* out) {
  if (n < 0 || n > kNoAllocMaxDraws || (n > 0 && !out)) return GHX_E_INVALID;
  if (n == 0) return GHX_OK;
  double* d = nullptr;
  if (int rc = scratch_ptr(&d)) return rc;
  // every draw of the staging range is written before it is copied: nothing
  // in the array needs clearing between calls
  hipLaunchKernelGGL(ghx::rng_gauss_kernel, dim3(unsigned((n + ghx::kRngBlock - 1) / ghx::kRngBlock)),
                     dim3(ghx::kRngBlock), 0, nullptr, (unsigned long long)seed, (unsigned long long)key0, (long long)n, d);
  if (hipGetLastError() != hipSuccess) return GHX_E_HIP;
  if (hipMemcpy(out, d, size_t(n) * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return GHX_E_HIP;
  return GHX_OK;
}
