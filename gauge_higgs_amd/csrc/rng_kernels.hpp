// The device's Gaussian stream as a kernel: out[i] = Rng(seed).gauss(key0 + i)
// (rng.hpp:42-46), one draw per thread, through the same draw_gauss the
// momentum refresh and the hot starts use. Internal linkage: each source file
// that launches it registers its own copy with the HIP runtime.
#pragma once
#include "device_math.hpp"

namespace ghx {
constexpr int kRngBlock = 256;
static __global__ __launch_bounds__(kRngBlock) void rng_gauss_kernel(unsigned long long seed, unsigned long long key0, long long n,
                                                                   double* out) {
  const long long i = (long long)blockIdx.x * kRngBlock + threadIdx.x;
  if (i < n) out[i] = draw_gauss(seed, key0 + (unsigned long long)i);
}
}  // namespace ghx
