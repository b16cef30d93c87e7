// Driver for the AddressSanitizer / UndefinedBehaviorSanitizer build of the
// host code (tests/test_sanitizers.py): the library's host-only parts (the
// representation builder, the symmetric-power engine's host build and program,
// the multi-invariant potential tables, the series analysis) and the FP64
// oracle (a few trajectories, forces, observables and measurements per rep).
// Exits 0 when every step ran; the sanitizers abort on the first finding.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "potential.hpp"
#include "symrep.hpp"

extern "C" {
int ghx_series_analysis(const double* x, long n, int block, double* out);
int ghx_autocorr_function(const double* x, long n, int maxlag, double* rho, long* nlag);
void* orc_create(int D, int N, const int* L, const char* rep_spec, std::uint64_t seed);
void orc_destroy(void* h);
int orc_rep_dim(void* h);
void orc_set_couplings(void* h, double beta, double kappa, double lambda, double tau, int nmd, int omelyan, int reunit);
void orc_hot(void* h, double sigma);
void orc_phi_gaussian(void* h, std::uint64_t stream, double sigma);
void orc_forces(void* h);
void orc_observables(void* h, double* out6);
int orc_trajectory(void* h, double* out3);
double orc_wilson_loop(void* h, const char* probe, int Ra, int Rb);
double orc_polyakov_rep(void* h, const char* probe, int tdir);
}

int main() {
  // host-only library code
  for (const auto& [N, spec] : std::vector<std::pair<int, std::string>>{{2, "fund"}, {2, "adj"}, {3, "2"}, {3, "2,1"}, {3, "6,1"}, {2, "6"}}) {
    const ghx::RepInfo r = ghx::build_rep(N, spec);
    const ghx::SymBuild s = ghx::build_sym(r);
    if (s.ok) (void)ghx::sym_program(s);
    if (r.d <= 8) {
      const std::vector<double> lam = ghx::casimir_channels(r);
      const std::vector<double> f(lam.size(), 1.0);
      (void)ghx::hermitian_superop(r.d, ghx::multi_superop(r, lam, f));
    }
    std::printf("rep %d:%s d=%d ok\n", N, spec.c_str(), r.d);
  }
  std::vector<double> x(5000), out(10), rho(5000);
  double v = 0.0;
  for (size_t i = 0; i < x.size(); ++i) { v = 0.9 * v + std::sin(0.37 * double(i * i % 101)); x[i] = v; }
  long nl = 0;
  if (ghx_series_analysis(x.data(), long(x.size()), 4, out.data()) || ghx_autocorr_function(x.data(), long(x.size()), -1, rho.data(), &nl))
    return 1;
  (void)ghx_series_analysis(x.data(), 1, 1, out.data());
  (void)ghx_series_analysis(x.data(), 0, 1, out.data());
  std::printf("series ok (tau_int %.3f)\n", out[4]);
  // the oracle
  for (const auto& [N, spec] : std::vector<std::pair<int, std::string>>{{2, ""}, {2, "fund"}, {2, "adj"}, {3, "2"}, {3, "1,1"}}) {
    const int L[4] = {3, 2, 3, 2};
    void* h = orc_create(4, N, L, spec.c_str(), 1);
    if (!h) return 2;
    orc_set_couplings(h, N == 2 ? 2.3 : 5.7, 0.2, 0.5, 1.0, 4, 1, 1);
    orc_hot(h, 0.8);
    if (!spec.empty()) orc_phi_gaussian(h, 12345, 0.3);
    orc_forces(h);
    double o6[6], o3[3];
    for (int t = 0; t < 2; ++t) (void)orc_trajectory(h, o3);
    orc_observables(h, o6);
    (void)orc_wilson_loop(h, "fund", 1, 2);
    (void)orc_polyakov_rep(h, "adj", 3);
    std::printf("oracle %d:%s plaq %.6f ok\n", N, spec.c_str(), o6[0]);
    orc_destroy(h);
  }
  return 0;
}
