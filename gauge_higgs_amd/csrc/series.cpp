// Time-series analysis of a Monte Carlo observable (host only): the Stats
// accumulator of observables.hpp:104-145 (mean, population variance, Binder
// cumulant, block-jackknife error) and the autocorrelation module of
// autocorr.hpp:55-212 (biased autocovariance, normalised autocorrelation,
// Sokal's self-consistent window M >= 6 tau_int(M) with lags capped at 2000,
// N_eff, naive / autocorrelation-aware / blocking errors). Definitions and
// guard values (empty, length-1 and constant series) are the reference's.
#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/ghx.h"

namespace {

constexpr double kSokalC = 6.0;
constexpr int kMaxLagCap = 2000;
constexpr double kDegenRelTol = 1e-24;  // C(0) <= tol <x^2> counts as no fluctuation

double mean(const double* x, long n) {
  if (n <= 0) return 0.0;
  double s = 0.0;
  for (long i = 0; i < n; ++i) s += x[i];
  return s / double(n);
}
double meansq(const double* x, long n) {
  if (n <= 0) return 0.0;
  double s = 0.0;
  for (long i = 0; i < n; ++i) s += x[i] * x[i];
  return s / double(n);
}
bool degenerate(double c0, double msq) { return c0 <= kDegenRelTol * msq + 1e-300; }

// C(t) = (1/n) sum_{s+t<n} (x_s - m)(x_{s+t} - m), t = 0..L (L = n-1 when maxlag < 0).
std::vector<double> autocov(const double* x, long n, long maxlag) {
  if (n <= 0) return {};
  const long L = (maxlag < 0 || maxlag > n - 1) ? n - 1 : maxlag;
  const double m = mean(x, n);
  std::vector<double> dx(x, x + n);
  for (double& v : dx) v -= m;
  std::vector<double> C(size_t(L) + 1);
  for (long t = 0; t <= L; ++t) {
    double s = 0.0;
    for (long i = 0; i + t < n; ++i) s += dx[size_t(i)] * dx[size_t(i + t)];
    C[size_t(t)] = s / double(n);
  }
  return C;
}

std::vector<double> autocorr(const double* x, long n, long maxlag) {
  std::vector<double> C = autocov(x, n, maxlag);
  if (C.empty()) return C;
  const double c0 = C[0];
  if (degenerate(c0, meansq(x, n))) {  // no memory by convention
    std::fill(C.begin(), C.end(), 0.0);
    C[0] = 1.0;
    return C;
  }
  for (double& v : C) v /= c0;
  return C;
}

double tau_int(const double* x, long n, int* window) {
  if (n < 2) { *window = 0; return 0.5; }
  const std::vector<double> rho = autocorr(x, n, std::min<long>(n - 1, kMaxLagCap));
  if (rho.size() < 2 || rho[0] <= 0.0) { *window = 0; return 0.5; }
  const int Mmax = int(rho.size()) - 1;
  double tau = 0.5;
  int W = Mmax;
  for (int M = 1; M <= Mmax; ++M) {
    tau += rho[size_t(M)];
    if (double(M) >= kSokalC * tau) { W = M; break; }
  }
  *window = W;
  return tau < 0.5 ? 0.5 : tau;
}

double population_var(const double* x, long n) {
  const double m = mean(x, n);
  double s2 = 0.0;
  for (long i = 0; i < n; ++i) s2 += x[i] * x[i];
  s2 = n > 0 ? s2 / double(n) : 0.0;
  return s2 - m * m;
}

double binder(const double* x, long n) {
  if (n <= 0) return 0.0;
  const double m = mean(x, n);
  double m2 = 0.0, m4 = 0.0;
  for (long i = 0; i < n; ++i) {
    const double d = x[i] - m;
    m2 += d * d;
    m4 += d * d * d * d;
  }
  m2 /= double(n);
  m4 /= double(n);
  return m2 > 0.0 ? 1.0 - m4 / (3.0 * m2 * m2) : 0.0;
}

// Standard error of the means of nblocks consecutive blocks of bs samples
// (sample variance of the block means over nblocks).
double block_means_error(const double* x, long nblocks, long bs) {
  std::vector<double> bm(size_t(nblocks), 0.0);
  for (long b = 0; b < nblocks; ++b) {
    double s = 0.0;
    for (long i = b * bs; i < (b + 1) * bs; ++i) s += x[i];
    bm[size_t(b)] = s / double(bs);
  }
  double mb = 0.0;
  for (double v : bm) mb += v;
  mb /= double(nblocks);
  double v = 0.0;
  for (double b : bm) v += (b - mb) * (b - mb);
  v /= double(nblocks - 1);
  return std::sqrt(v / double(nblocks));
}

double binned_error(const double* x, long n, int nbins = 16) {
  if (n < 2L * nbins) nbins = int(std::max<long>(2, n / 2));
  if (n < 2) return 0.0;
  return block_means_error(x, nbins, n / nbins);
}

double blocking_error(const double* x, long n, long bs) {
  if (bs < 1) bs = 1;
  const long nb = n / bs;
  if (nb < 2) return 0.0;
  return block_means_error(x, nb, bs);
}

}  // namespace

extern "C" int ghx_series_analysis(const double* x, long n, int block, double* out) {
  if (!out || n < 0 || (n > 0 && !x)) return GHX_E_INVALID;
  int W = 0;
  const double tau = tau_int(x, n, &W);
  const double c0 = n >= 2 ? autocov(x, n, 0)[0] : 0.0;
  out[0] = mean(x, n);
  out[1] = population_var(x, n);
  out[2] = binder(x, n);
  out[3] = binned_error(x, n);
  out[4] = tau;
  out[5] = double(W);
  out[6] = n > 0 ? double(n) / (2.0 * tau) : 0.0;
  out[7] = n >= 2 ? std::sqrt(c0 / double(n)) : 0.0;
  out[8] = (n < 2 || degenerate(c0, meansq(x, n))) ? 0.0 : std::sqrt(2.0 * tau * c0 / double(n));
  out[9] = blocking_error(x, n, block);
  return GHX_OK;
}

extern "C" int ghx_autocorr_function(const double* x, long n, int maxlag, double* rho, long* nlag) {
  if (!nlag || n < 0 || (n > 0 && (!x || !rho))) return GHX_E_INVALID;
  const std::vector<double> r = autocorr(x, n, maxlag);
  std::copy(r.begin(), r.end(), rho);
  *nlag = long(r.size());
  return GHX_OK;
}
