// Shared pieces of the drop-in CLIs: argument helpers, the reference's Stats
// accumulator (observables.hpp:104-145) over the library's series analysis and
// error handling around the C ABI.
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ghx.h"

namespace cli {

// The reference's Stats accumulator (observables.hpp:104-145) and
// autocorrelation analysis (autocorr.hpp:55-212), evaluated by
// ghx_series_analysis in the library.
struct Stats {
  std::vector<double> x;
  void add(double v) { x.push_back(v); }
  int n() const { return int(x.size()); }
  std::array<double, 10> analysis(int block = 1) const {
    std::array<double, 10> o{};
    ghx_series_analysis(x.data(), long(x.size()), block, o.data());
    return o;
  }
  double mean() const { return analysis()[0]; }
  double var() const { return analysis()[1]; }
  double susceptibility(double V) const { return V * var(); }
  double binder() const { return analysis()[2]; }
  double binned_error() const { return analysis()[3]; }
  double tau_int(int* window = nullptr) const {
    const auto o = analysis();
    if (window) *window = int(o[5]);
    return o[4];
  }
  double neff() const { return analysis()[6]; }
  double autocorr_error() const { return analysis()[8]; }
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
