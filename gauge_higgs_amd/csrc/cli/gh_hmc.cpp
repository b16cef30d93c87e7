// Drop-in for the reference pure-gauge driver (main.cpp:11-51):
//   gh_hmc [L=8] [beta=2.3] [ntherm=50] [nmeas=100] [nmd=20] [tau=1.0] [seed=1]
//          [--N=2] [--D=4] [--integrator=omelyan|leapfrog] [--precision=fp64|fp32] [--no-graph]
// Same argv meaning and stdout lines; the HMC runs on the GPU through the ghx C ABI.
#include "cli_common.hpp"

int main(int argc, char** argv) {
  std::vector<char*> a;
  cli::Flags fl;
  if (cli::split_flags(argc, argv, a, fl)) return 1;
  const int ac = int(a.size());
  auto argf = [&](int i, double def) { return i < ac ? std::atof(a[i]) : def; };
  auto argi = [&](int i, long def) { return i < ac ? std::atol(a[i]) : def; };
  const int Lext = int(argi(1, 8));
  const double beta = argf(2, 2.3);
  const int ntherm = int(argi(3, 50)), nmeas = int(argi(4, 100)), nmd = int(argi(5, 20));
  const double tau = argf(6, 1.0);
  const uint64_t seed = uint64_t(argi(7, 1));

  std::vector<int> L(fl.D, Lext);
  ghx_ctx* ctx = nullptr;
  if (ghx_create(&ctx, fl.D, fl.N, 1, L.data(), fl.precision) != GHX_OK) { std::fprintf(stderr, "ghx_create failed\n"); return 1; }
  GHX_CHECK(ctx, ghx_set_rep(ctx, "", nullptr));
  GHX_CHECK(ctx, ghx_set_integrator(ctx, nmd, fl.integrator, 1));
  GHX_CHECK(ctx, ghx_set_couplings(ctx, 0, beta, 0.0, 0.0, tau));
  GHX_CHECK(ctx, ghx_set_seed(ctx, 0, seed));
  long long vol = 1;
  for (int m = 0; m < fl.D; ++m) vol *= Lext;
  std::printf("# Pure-gauge HMC: D=%d  SU(%d)  L=%d^%d  vol=%lld  beta=%.4f  tau=%.3f  nmd=%d\n", fl.D, fl.N, Lext, fl.D, vol, beta, tau, nmd);
  GHX_CHECK(ctx, ghx_init_fields(ctx, 0, GHX_HOT, 0.8, 0, 0, 0.0));

  GHX_CHECK(ctx, ghx_trajectories(ctx, ntherm, fl.graph ? 1 : 0));
  uint64_t tc = 0, acc = 0;
  double dH = 0;
  GHX_CHECK(ctx, ghx_get_counters(ctx, 0, &tc, &acc, &dH));
  std::printf("# thermalized %d trajectories; acceptance %.3f\n", ntherm, tc ? double(acc) / tc : 0.0);

  GHX_CHECK(ctx, ghx_set_counters(ctx, 0, 0, 0));  // main.cpp:36
  double sP = 0, sP2 = 0, sExp = 0, obs[6];
  for (int t = 0; t < nmeas; ++t) {
    GHX_CHECK(ctx, ghx_trajectories(ctx, 1, fl.graph ? 1 : 0));
    GHX_CHECK(ctx, ghx_get_counters(ctx, 0, &tc, &acc, &dH));
    GHX_CHECK(ctx, ghx_observables(ctx, obs));
    sP += obs[0]; sP2 += obs[0] * obs[0]; sExp += std::exp(-dH);
  }
  const double mP = sP / nmeas, varP = sP2 / nmeas - mP * mP;
  double mu = 0;
  GHX_CHECK(ctx, ghx_max_unitarity(ctx, 0, &mu));
  std::printf("# measured %d trajectories\n", nmeas);
  std::printf("plaquette       = %.6f +/- %.6f\n", mP, std::sqrt(std::max(0.0, varP) / nmeas));
  std::printf("acceptance      = %.4f\n", tc ? double(acc) / tc : 0.0);
  std::printf("<exp(-dH)>      = %.5f  (should be ~1)\n", sExp / nmeas);
  std::printf("max ||U^dU-I||  = %.3e\n", mu);
  ghx_destroy(ctx);
  return 0;
}
