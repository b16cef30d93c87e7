// Drop-in for the reference gauge+Higgs driver (hmc_higgs.cpp:36-100):
//   hmc_higgs <rep> <L> <beta> <kappa> <lambda> [ntherm=80] [nmeas=200] [nmd=24] [tau=1.0] [seed=1]
//             [--N=2] [--D=4] [--integrator=omelyan|leapfrog] [--precision=fp64|fp32] [--no-graph]
// <rep> = fund | adj | Young rows "r1,r2,..." with optional ":real".
#include <cstdlib>

#include "cli_common.hpp"

int main(int argc, char** argv) {
  std::vector<char*> a;
  cli::Flags fl;
  if (cli::split_flags(argc, argv, a, fl)) return 1;
  const int ac = int(a.size());
  if (ac < 6) {
    std::fprintf(stderr,
                 "usage: %s <rep> <L> <beta> <kappa> <lambda> [ntherm=80] [nmeas=200] [nmd=24] [tau=1.0] [seed=1]\n"
                 "  <rep> = fund | adj | <Young rows e.g. 2,1> [add ':real' for a real-scalar rep, e.g. 2,1:real]\n",
                 a[0]);
    return 1;
  }
  auto argf = [&](int i, double def) { return i < ac ? std::atof(a[i]) : def; };
  auto argi = [&](int i, long def) { return i < ac ? std::atol(a[i]) : def; };
  const char* repname = a[1];
  const int Lext = int(argi(2, 8));
  const double beta = argf(3, 2.3), kappa = argf(4, 0.2), lambda = argf(5, 0.5);
  const int ntherm = int(argi(6, 80)), nmeas = int(argi(7, 200)), nmd = int(argi(8, 24));
  const double tau = argf(9, 1.0);
  const uint64_t seed = uint64_t(argi(10, 1));

  std::vector<int> L(fl.D, Lext);
  ghx_ctx* ctx = nullptr;
  if (ghx_create(&ctx, fl.D, fl.N, 1, L.data(), fl.precision) != GHX_OK) { std::fprintf(stderr, "ghx_create failed\n"); return 1; }
  int d = 0;
  const int rc = ghx_set_rep(ctx, repname, &d);
  if (rc == GHX_E_REP) { std::fprintf(stderr, "rep error: %s\n", ghx_last_error(ctx)); return 1; }
  GHX_CHECK(ctx, rc);
  int nality = 0, real = 0;
  long dimT = 0;
  GHX_CHECK(ctx, ghx_rep_info(ctx, &d, &nality, &real, &dimT));
  GHX_CHECK(ctx, ghx_set_integrator(ctx, nmd, fl.integrator, 1));
  GHX_CHECK(ctx, ghx_set_couplings(ctx, 0, beta, kappa, lambda, tau));
  GHX_CHECK(ctx, ghx_set_seed(ctx, 0, seed));
  long long vol = 1;
  for (int m = 0; m < fl.D; ++m) vol *= Lext;
  std::string name = repname;
  if (name == "fund") name = "fundamental";
  else if (name == "adj") name = "adjoint";
  else { const auto c = name.find(':'); name = "general[" + name.substr(0, c) + "]"; }
  std::printf("# Gauge+Higgs HMC: D=%d SU(%d) rep=%s d=%d N-ality=%d L=%d^%d vol=%lld\n", fl.D, fl.N, name.c_str(), d, nality, Lext,
              fl.D, vol);
  std::printf("# beta=%.4f kappa=%.4f lambda=%.4f tau=%.3f nmd=%d  (kappa_c,free=1/(2D)=%.4f)\n", beta, kappa, lambda, tau, nmd,
              1.0 / (2.0 * fl.D));
  // Start modes of the reference's string / multi-invariant drivers (gh_string.cpp:106-125,
  // hmc_higgs_multi.cpp:81-90): GH_COLD -> unit links and phi = (1, 0, ...); GH_FROZEN ->
  // |phi_x| = 1 frozen-length scalar, normalised before thermalising. Default: hot start.
  if (std::getenv("GH_COLD")) GHX_CHECK(ctx, ghx_init_fields(ctx, 0, GHX_COLD, 0.0, 0, 0, 1.0));
  else GHX_CHECK(ctx, ghx_init_fields(ctx, 0, GHX_HOT, 0.8, 1, 12345, 0.3));  // hmc_higgs.cpp:72-73
  if (std::getenv("GH_FROZEN")) {
    GHX_CHECK(ctx, ghx_set_frozen_phi(ctx, 1));
    GHX_CHECK(ctx, ghx_normalize_phi(ctx));
  }
  GHX_CHECK(ctx, ghx_trajectories(ctx, ntherm, fl.graph ? 1 : 0));

  GHX_CHECK(ctx, ghx_set_counters(ctx, 0, 0, 0));  // hmc_higgs.cpp:76
  cli::Stats plaq, Lphi, Llink, poly;
  double sExp = 0, obs[6], dH = 0;
  uint64_t tc = 0, acc = 0;
  for (int t = 0; t < nmeas; ++t) {
    GHX_CHECK(ctx, ghx_trajectories(ctx, 1, fl.graph ? 1 : 0));
    GHX_CHECK(ctx, ghx_get_counters(ctx, 0, &tc, &acc, &dH));
    GHX_CHECK(ctx, ghx_observables(ctx, obs));
    plaq.add(obs[0]); Llink.add(obs[3]); Lphi.add(obs[4]); poly.add(obs[5]);
    sExp += std::exp(-dH);
  }
  double mu = 0;
  GHX_CHECK(ctx, ghx_max_unitarity(ctx, 0, &mu));
  const double V = double(vol);
  std::printf("plaquette   = %.6f +/- %.6f\n", plaq.mean(), plaq.binned_error());
  std::printf("L_phi       = %.6f +/- %.6f   (<phi^dag phi>)\n", Lphi.mean(), Lphi.binned_error());
  std::printf("L_link      = %.6f +/- %.6f   (gauge-inv hopping energy; transition locator)\n", Llink.mean(), Llink.binned_error());
  std::printf("Polyakov    = %.6f +/- %.6f\n", poly.mean(), poly.binned_error());
  std::printf("chi_link    = %.5f   (V*Var(L_link), susceptibility)\n", Llink.susceptibility(V));
  std::printf("Binder_link = %.5f   (U4 of L_link)\n", Llink.binder());
  std::printf("chi_phi     = %.5f   (V*Var(L_phi))\n", Lphi.susceptibility(V));
  std::printf("acceptance  = %.4f\n", tc ? double(acc) / tc : 0.0);
  std::printf("<exp(-dH)>  = %.5f\n", sExp / nmeas);
  std::printf("max||UdU-I||= %.2e\n", mu);
  ghx_destroy(ctx);
  return 0;
}
