// Correctly rounded log, cos and sqrt for the Box-Muller Gaussian of the counter-based
// RNG (rng.hpp:42-46): g = sqrt(-2 log u1) cos(2 pi u2).
//
// log and cos evaluate in double-double arithmetic (Dekker's error-free sum
// and product, the product through fma) to about 2^-100 relative and round
// once, so they return the correctly rounded double except when the exact
// value lies within ~2^-100 of a rounding boundary (not seen in 10^9 draws).
// sqrt corrects the hardware-assisted approximation with an exact residual
// test, since the device's own double sqrt is not guaranteed to round correctly.
// Host glibc's log and cos (0.52 ulp) are not correctly rounded: they differ
// from these on about 1 draw in 600, and on those draws the reference's
// Gaussian and this one differ in the last place.
//
// log (Tang's table-driven method): x = 2^e m, m in [0.75, 1.5), c = j/128 the
// nearest table point, f = m - c (exact), u = f/(2c + f);
//   log x = e ln 2 + log c + 2 atanh u,   |u| <= 2^-8.5,
// with 2 atanh u = 2u (1 + w/3 + w^2/5 + ... + w^6/13), w = u^2.
// cos (Cody-Waite reduction): k = rint(x 2/pi), r = x - k pi/2 with pi/2 in
// three parts (the first product exact through fma), then r = a + t with
// a = j/64 from a table of sin a, cos a and |t| <= 1/128:
//   sin(a + t) = sin a cos t + cos a sin t,  cos(a + t) = cos a cos t - sin a sin t,
// and the Taylor series of sin t and cos t.
// The constants come from gen_gauss_tables.py (gauss_tables.hpp).
//
// Host and device: the CPU tests compile this header with g++ and compare it
// with libquadmath and glibc; the device RNG (device_math.hpp) calls it.
#pragma once
#include <cstdint>

#include "gauss_tables.hpp"

#if defined(__HIPCC__)
#define GHX_HD __host__ __device__ inline
#else
#define GHX_HD inline
#endif
// The error-free sums and products rely on every operation rounding on its
// own: fusing a*b + c into an fma (hipcc's default contraction for device code)
// changes the rounding errors they capture. Each function below opens with
// this, so no operation in it is contracted.
#if defined(__clang__)
#define GCR_EXACT _Pragma("clang fp contract(off)")
#else
#define GCR_EXACT
#endif

namespace ghx {
namespace gcr {

struct dd {
  double h, l;
};

GHX_HD dd two_sum(double a, double b) {
  GCR_EXACT
  const double s = a + b;
  const double bb = s - a;
  return {s, (a - (s - bb)) + (b - bb)};
}
GHX_HD dd fast_two_sum(double a, double b) {  // |a| >= |b| or a == 0
  GCR_EXACT
  const double s = a + b;
  return {s, b - (s - a)};
}
GHX_HD dd two_prod(double a, double b) {
  GCR_EXACT
  const double p = a * b;
  return {p, __builtin_fma(a, b, -p)};
}
GHX_HD dd add(dd a, dd b) {
  GCR_EXACT
  dd s = two_sum(a.h, b.h);
  const dd t = two_sum(a.l, b.l);
  s.l += t.h;
  s = fast_two_sum(s.h, s.l);
  s.l += t.l;
  return fast_two_sum(s.h, s.l);
}
GHX_HD dd mul(dd a, dd b) {
  GCR_EXACT
  dd p = two_prod(a.h, b.h);
  p.l += a.h * b.l + a.l * b.h;
  return fast_two_sum(p.h, p.l);
}
GHX_HD dd mul_d(dd a, double b) {
  GCR_EXACT
  dd p = two_prod(a.h, b);
  p.l += a.l * b;
  return fast_two_sum(p.h, p.l);
}
GHX_HD dd div(dd a, dd b) {
  GCR_EXACT
  const double q1 = a.h / b.h;
  const dd r = add(a, mul_d(b, -q1));
  const double q2 = r.h / b.h;
  return fast_two_sum(q1, q2);
}
GHX_HD dd dd_of(const double* c) { return {c[0], c[1]}; }

GHX_HD uint64_t bits(double x) { return __builtin_bit_cast(uint64_t, x); }
GHX_HD double from_bits(uint64_t b) { return __builtin_bit_cast(double, b); }

// log of a positive normal double (the Box-Muller u1 is in [2^-54, 1]).
GHX_HD double log_cr(double x) {
  GCR_EXACT
  const uint64_t ix = bits(x);
  int e = int((ix >> 52) & 0x7ff) - 1023;
  double m = from_bits((ix & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);  // [1, 2)
  if (m >= 1.5) { m *= 0.5; ++e; }                                             // [0.75, 1.5)
  const int j = int(m * 128.0 + 0.5);                                          // 96..192
  const double c = double(j) * (1.0 / 128.0);
  const double f = m - c;  // exact (Sterbenz), |f| <= 1/256
  dd P = {0.0, 0.0};
  if (f != 0.0) {
    const dd u = div(dd{f, 0.0}, two_sum(2.0 * c, f));
    const dd w = mul(u, u);
    const double wh = w.h;
    const double L2 = 1.0 / 7.0 + wh * (1.0 / 9.0 + wh * (1.0 / 11.0 + wh * (1.0 / 13.0)));
    const dd L1 = add(dd_of(kFifth), mul_d(w, L2));
    const dd L0 = add(dd_of(kThird), mul(w, L1));
    const dd Q = add(dd{1.0, 0.0}, mul(w, L0));
    P = mul(u, Q);
    P.h *= 2.0;
    P.l *= 2.0;
  }
  const double ed = double(e);
  dd E = two_prod(ed, kLn2[0]);
  E = add(E, two_prod(ed, kLn2[1]));
  E.l += ed * kLn2[2];
  const dd R = add(add(E, dd_of(kLogC[j - 96])), P);
  return R.h;
}

// sin t and cos t, |t| <= 1/128 + 2^-50, as double-doubles.
GHX_HD void sincos_small(dd t, dd* s, dd* c) {
  GCR_EXACT
  const dd z = mul(t, t);
  const double zh = z.h;
  // sin t = t (1 - z/6 + z^2/120 - z^3/5040 + z^4/9! - z^5/11!)
  const double T2 = -1.0 / 5040.0 + zh * (1.0 / 362880.0 - zh * (1.0 / 39916800.0));
  const dd T1 = add(dd_of(kInv120), mul_d(z, T2));
  const dd T0 = add(dd{-kSixth[0], -kSixth[1]}, mul(z, T1));
  *s = mul(t, add(dd{1.0, 0.0}, mul(z, T0)));
  // cos t = 1 - z/2 + z^2/24 - z^3/720 + z^4/8! - z^5/10! + z^6/12!
  const double C2 = -1.0 / 720.0 + zh * (1.0 / 40320.0 - zh * (1.0 / 3628800.0 - zh * (1.0 / 479001600.0)));
  const dd C1 = add(dd_of(kInv24), mul_d(z, C2));
  const dd C0 = add(dd{-0.5, 0.0}, mul(z, C1));
  *c = add(dd{1.0, 0.0}, mul(z, C0));
}

// cos of a double with |x| < 2^20 (the Box-Muller argument is in [0, 2 pi)).
GHX_HD double cos_cr(double x) {
  GCR_EXACT
  x = x < 0.0 ? -x : x;
  const double kd = __builtin_rint(x * kTwoOverPi);
  const int q = int(kd) & 3;
  const double r1 = __builtin_fma(-kd, kPio2[0], x);  // exact: k small, result on x's grid
  const dd t2 = two_prod(kd, kPio2[1]);
  dd r = two_sum(r1, -t2.h);
  r.l += -t2.l - kd * kPio2[2];
  r = fast_two_sum(r.h, r.l);
  const bool neg = r.h < 0.0;
  if (neg) { r.h = -r.h; r.l = -r.l; }
  const int j = int(r.h * 64.0 + 0.5);  // 0..51
  const double a = double(j) * (1.0 / 64.0);
  const dd t = two_sum(r.h - a, r.l);  // r.h - a exact (Sterbenz for j >= 1)
  dd st, ct;
  sincos_small(t, &st, &ct);
  const dd Sa = dd_of(&kSinCos[j][0]), Ca = dd_of(&kSinCos[j][2]);
  dd v;
  if (q == 0 || q == 2) {  // +-cos r
    v = add(mul(Ca, ct), mul(Sa, dd{-st.h, -st.l}));
    if (q == 2) v.h = -v.h;
  } else {  // -+sin r; sin(-|r|) = -sin|r|
    v = add(mul(Sa, ct), mul(Ca, st));
    if ((q == 1) != neg) v.h = -v.h;
  }
  return v.h;
}

// Correctly rounded sqrt from an approximation c within one ulp of it (the
// hardware-assisted sqrt): with r = x - c^2, exact through fma, c is too small
// when x lies above (c + u/2)^2, i.e. r > c u (u the spacing above c), and too
// large when x lies at or below (c - v/2)^2, i.e. r <= -c v (v the spacing
// below c, u/2 at a power of two). Every quantity compared is exact: r, c u and
// c v are multiples of u^2 / 2, while the neglected squares of half spacings are
// smaller than that unit.
GHX_HD double sqrt_fix(double x, double c) {
  GCR_EXACT
  if (!(c > 0.0)) return c;
  const double r = __builtin_fma(-c, c, x);
  const double up = from_bits(bits(c) + 1) - c, dn = c - from_bits(bits(c) - 1);
  if (r > c * up) return c + up;
  if (r <= -(c * dn)) return c - dn;
  return c;
}
GHX_HD double sqrt_cr(double x) { return sqrt_fix(x, __builtin_sqrt(x)); }

}  // namespace gcr
}  // namespace ghx
