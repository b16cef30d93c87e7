// Shared pieces of the drop-in CLIs: argument helpers, the reference's Stats
// accumulator (observables.hpp:104-145) and error handling around the C ABI.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ghx.h"

namespace cli {

struct Stats {  // observables.hpp:104-145
  std::vector<double> x;
  void add(double v) { x.push_back(v); }
  int n() const { return int(x.size()); }
  double mean() const { double s = 0; for (double v : x) s += v; return x.empty() ? 0.0 : s / x.size(); }
  double var() const {
    const double m = mean();
    double s2 = 0;
    for (double v : x) s2 += v * v;
    s2 = x.empty() ? 0.0 : s2 / x.size();
    return s2 - m * m;
  }
  double susceptibility(double V) const { return V * var(); }
  double binder() const {
    const double m = mean();
    double m2 = 0, m4 = 0;
    for (double v : x) { const double d = v - m; m2 += d * d; m4 += d * d * d * d; }
    if (x.empty()) return 0.0;
    m2 /= x.size(); m4 /= x.size();
    return m2 > 0 ? 1.0 - m4 / (3.0 * m2 * m2) : 0.0;
  }
  double binned_error(int nbins = 16) const {
    const int N = n();
    if (N < 2 * nbins) nbins = std::max(2, N / 2);
    if (N < 2) return 0.0;
    const int bs = N / nbins;
    std::vector<double> bm;
    for (int b = 0; b < nbins; ++b) { double s = 0; for (int i = b * bs; i < (b + 1) * bs; ++i) s += x[i]; bm.push_back(s / bs); }
    double mb = 0; for (double v : bm) mb += v; mb /= bm.size();
    double v = 0; for (double b : bm) v += (b - mb) * (b - mb);
    v /= (bm.size() - 1);
    return std::sqrt(v / bm.size());
  }
};

// Autocorrelation analysis of a Monte Carlo series (autocorr.hpp:53-184):
// biased autocovariance, Sokal's self-consistent window M >= 6 tau_int(M)
// with the lag capped at 2000, N_eff = N / (2 tau_int) and the error
// sqrt(2 tau_int Var / N).
struct Autocorr {
  static constexpr double kSokal = 6.0;
  static constexpr int kLagCap = 2000;
  static constexpr double kDegen = 1e-24;  // C(0) below kDegen <x^2>: no fluctuations
  static double mean(const std::vector<double>& x) {
    double s = 0;
    for (double v : x) s += v;
    return x.empty() ? 0.0 : s / x.size();
  }
  static double meansq(const std::vector<double>& x) {
    double s = 0;
    for (double v : x) s += v * v;
    return x.empty() ? 0.0 : s / x.size();
  }
  static bool degenerate(double c0, double msq) { return c0 <= kDegen * msq + 1e-300; }
  static std::vector<double> autocov(const std::vector<double>& x, int maxlag) {
    const long N = long(x.size());
    if (N <= 0) return {};
    const long L = (maxlag < 0 || maxlag > N - 1) ? N - 1 : maxlag;
    const double m = mean(x);
    std::vector<double> C(L + 1);
    for (long t = 0; t <= L; ++t) {
      double s = 0;
      for (long i = 0; i + t < N; ++i) s += (x[i] - m) * (x[i + t] - m);
      C[t] = s / double(N);
    }
    return C;
  }
  static double tau_int(const std::vector<double>& x, int* window = nullptr) {
    const int N = int(x.size());
    if (N < 2) { if (window) *window = 0; return 0.5; }
    std::vector<double> rho = autocov(x, std::min(N - 1, kLagCap));
    if (degenerate(rho[0], meansq(x))) { if (window) *window = 0; return 0.5; }
    const double c0 = rho[0];
    for (double& v : rho) v /= c0;
    const int Mmax = int(rho.size()) - 1;
    double tau = 0.5;
    int W = Mmax;
    for (int M = 1; M <= Mmax; ++M) {
      tau += rho[M];
      if (double(M) >= kSokal * tau) { W = M; break; }
    }
    if (window) *window = W;
    return tau < 0.5 ? 0.5 : tau;
  }
  static double neff(const std::vector<double>& x) { return x.empty() ? 0.0 : x.size() / (2.0 * tau_int(x)); }
  static double error(const std::vector<double>& x) {
    if (x.size() < 2) return 0.0;
    const double var = autocov(x, 0)[0];
    if (degenerate(var, meansq(x))) return 0.0;
    return std::sqrt(2.0 * tau_int(x) * var / double(x.size()));
  }
};

// Options that follow the reference's positional arguments.
struct Flags {
  int N = 2, D = 4, integrator = GHX_OMELYAN, precision = GHX_FP64;
  bool graph = true;
};
inline int split_flags(int argc, char** argv, std::vector<char*>& pos, Flags& f) {
  pos.push_back(argv[0]);
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a.rfind("--", 0) != 0) { pos.push_back(argv[i]); continue; }
    const auto eq = a.find('=');
    const std::string k = a.substr(2, eq == std::string::npos ? std::string::npos : eq - 2);
    const std::string v = eq == std::string::npos ? "" : a.substr(eq + 1);
    if (k == "N") f.N = std::atoi(v.c_str());
    else if (k == "D") f.D = std::atoi(v.c_str());
    else if (k == "integrator") f.integrator = (v == "leapfrog") ? GHX_LEAPFROG : GHX_OMELYAN;
    else if (k == "precision") f.precision = (v == "fp32") ? GHX_FP32MD : GHX_FP64;
    else if (k == "no-graph") f.graph = false;
    else { std::fprintf(stderr, "unknown option %s\n", a.c_str()); return 1; }
  }
  return 0;
}

#define GHX_CHECK(ctx, call)                                                              \
  do {                                                                                    \
    const int rc_ = (call);                                                               \
    if (rc_ != GHX_OK) {                                                                  \
      std::fprintf(stderr, "ghx error %d in %s: %s\n", rc_, #call, ghx_last_error(ctx));  \
      std::exit(1);                                                                       \
    }                                                                                     \
  } while (0)

}  // namespace cli
