// CPU check of gauss_cr.hpp (the device RNG's log and cos) on the Box-Muller
// inputs of the reference RNG (rng.hpp:35-46): uniforms u1, u2 from
// splitmix64 sub-keys 2k+1, 2k+2 of key k, argument x = fl(2 pi u2).
//
//   gauss_check SEED K0 N
//   gauss_check sqrt N
//
// prints one line of counts over keys K0..K0+N-1:
//   n  log_vs_quad  cos_vs_quad  glibc_log_vs_quad  glibc_cos_vs_quad
//   gauss_vs_glibc  gauss_vs_quad
// where "quad" is the libquadmath value rounded to double (the correctly
// rounded result up to quad precision), "gauss" is sqrt(-2 log u1) cos x with
// this header's log and cos, and the glibc columns use the host libm.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>

#include <quadmath.h>

#include "gauss_cr.hpp"

static inline uint64_t sm64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}
static inline double uni(uint64_t seed, uint64_t k) {
  const uint64_t h = sm64(seed ^ sm64(k));
  return (double(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// sqrt_fix from approximations one ulp low, exact and one ulp high, on
// random doubles over 2^-100..2^100 and on powers of two: prints the count of
// results that differ from the correctly rounded std::sqrt.
static int check_sqrt(long n) {
  long bad = 0;
  uint64_t s = 0x243F6A8885A308D3ULL;
  for (long i = 0; i < n; ++i) {
    s = sm64(s);
    double x = std::ldexp(1.0 + double(s >> 11) * 0x1p-53, int(s % 200) - 100);
    if (i % 7 == 0) x = std::ldexp(1.0, int(s % 200) - 100);
    const double t = std::sqrt(x);
    for (double c : {std::nextafter(t, 0.0), t, std::nextafter(t, 1e300)}) bad += ghx::gcr::sqrt_fix(x, c) != t;
  }
  std::printf("%ld %ld\n", 3 * n, bad);
  return 0;
}

int main(int argc, char** argv) {
  if (argc == 3 && std::string(argv[1]) == "sqrt") return check_sqrt(std::atol(argv[2]));
  if (argc < 4) return 2;
  const uint64_t seed = std::strtoull(argv[1], nullptr, 0), k0 = std::strtoull(argv[2], nullptr, 0);
  const long n = std::atol(argv[3]);
  long c_log = 0, c_cos = 0, g_log = 0, g_cos = 0, c_gauss_glibc = 0, c_gauss_quad = 0;
#pragma omp parallel for reduction(+ : c_log, c_cos, g_log, g_cos, c_gauss_glibc, c_gauss_quad) schedule(static)
  for (long i = 0; i < n; ++i) {
    const uint64_t k = k0 + uint64_t(i);
    const double u1 = uni(seed, sm64(k * 2 + 1)), u2 = uni(seed, sm64(k * 2 + 2));
    const double x = 2.0 * 3.14159265358979323846 * u2;
    const double lq = double(logq(__float128(u1))), cq = double(cosq(__float128(x)));
    const double lo = ghx::gcr::log_cr(u1), co = ghx::gcr::cos_cr(x);
    const double lg = std::log(u1), cg = std::cos(x);
    c_log += lo != lq;
    c_cos += co != cq;
    g_log += lg != lq;
    g_cos += cg != cq;
    const double go = std::sqrt(-2.0 * lo) * co;
    c_gauss_glibc += go != std::sqrt(-2.0 * lg) * cg;
    c_gauss_quad += go != std::sqrt(-2.0 * lq) * cq;
  }
  std::printf("%ld %ld %ld %ld %ld %ld %ld\n", n, c_log, c_cos, g_log, g_cos, c_gauss_glibc, c_gauss_quad);
  return 0;
}
