// Drop-in for the reference multi-invariant-potential driver (hmc_higgs_multi.cpp:40-125):
//   hmc_higgs_multi <rep>                      print the channel Casimirs and the coupling count
//   hmc_higgs_multi <rep> <L> <beta> <kappa> <mu2> <couplings|auto> [ntherm=80] [nmeas=200] [nmd=24]
//                   [tau=1.0] [seed=1] [ts_outfile]   [--N=3] [--D=4] [--integrator=...] [--precision=...] [--no-graph]
// V = -mu2 phi^dag phi + sum_c f_c V_c; GH_FROZEN (|phi_x| = 1) and GH_COLD start modes as the reference.
#include <cstdlib>
#include <fstream>
#include <sstream>

#include "cli_common.hpp"

int main(int argc, char** argv) {
  std::vector<char*> a;
  cli::Flags fl;
  fl.N = 3;  // the reference driver is SU(3) (kN = 3)
  if (cli::split_flags(argc, argv, a, fl)) return 1;
  const int ac = int(a.size());
  if (ac < 2) {
    std::fprintf(stderr,
                 "usage: %s <rep> <L> <beta> <kappa> <mu2> <couplings|auto> [ntherm=80 nmeas=200 nmd=24 tau=1 seed=1 ts_outfile]\n"
                 "  run with just <rep> to print the channel C2 values and the required #couplings.\n",
                 a[0]);
    return 1;
  }
  const char* repname = a[1];
  int nch = 0;
  const int qrc = ghx_casimir_channels(fl.N, repname, nullptr, &nch);
  if (qrc == GHX_E_REP) { std::fprintf(stderr, "rep error: invalid representation '%s'\n", repname); return 1; }
  if (qrc != GHX_OK) { std::fprintf(stderr, "error: no channels for '%s'\n", repname); return 1; }
  std::vector<double> lam(nch);
  ghx_casimir_channels(fl.N, repname, lam.data(), &nch);
  int d = 0, nality = 0, real = 0;
  long dimT = 0;
  ghx_rep_query(fl.N, repname, &d, &nality, &real, &dimT, nullptr, nullptr);
  std::string name = repname;
  if (name == "fund") name = "fundamental";
  else if (name == "adj") name = "adjoint";
  else { const auto c = name.find(':'); name = "general[" + name.substr(0, c) + "]"; }
  std::printf("# rep=%s d=%d  %d quartic channels (C2): ", name.c_str(), d, nch);
  for (double c : lam) std::printf("%.4g ", c);
  std::printf("\n");
  if (ac < 7) { std::printf("# provide %d couplings f_c (comma list) as arg 6 (or 'auto').\n", nch); return 0; }

  auto argf = [&](int i, double def) { return i < ac ? std::atof(a[i]) : def; };
  auto argi = [&](int i, long def) { return i < ac ? std::atol(a[i]) : def; };
  const int Lext = int(argi(2, 8));
  const double beta = argf(3, 2.3), kappa = argf(4, 0.2), mu2 = argf(5, 1.0);
  const int ntherm = int(argi(7, 80)), nmeas = int(argi(8, 200)), nmd = int(argi(9, 24));
  const double tau = argf(10, 1.0);
  const uint64_t seed = uint64_t(argi(11, 1));
  const std::string ts_out = ac > 12 ? std::string(a[12]) : std::string();

  std::vector<double> f;
  if (std::string(a[6]) == "auto") f.assign(nch, 1.0);
  else { std::stringstream ss(a[6]); std::string t; while (std::getline(ss, t, ',')) if (!t.empty()) f.push_back(std::atof(t.c_str())); }
  if (int(f.size()) != nch) { std::fprintf(stderr, "error: need %d couplings, got %zu\n", nch, f.size()); return 1; }

  std::vector<int> L(fl.D, Lext);
  ghx_ctx* ctx = nullptr;
  if (ghx_create(&ctx, fl.D, fl.N, 1, L.data(), fl.precision) != GHX_OK) { std::fprintf(stderr, "ghx_create failed\n"); return 1; }
  GHX_CHECK(ctx, ghx_set_rep(ctx, repname, &d));
  GHX_CHECK(ctx, ghx_set_integrator(ctx, nmd, fl.integrator, 1));
  GHX_CHECK(ctx, ghx_set_couplings(ctx, 0, beta, kappa, 0.0, tau));
  GHX_CHECK(ctx, ghx_set_seed(ctx, 0, seed));
  GHX_CHECK(ctx, ghx_set_multi_potential(ctx, f.data(), nch, mu2));
  const bool frozen = std::getenv("GH_FROZEN") != nullptr;
  std::printf("# D=%d SU(%d) L=%d^%d  beta=%.3f kappa=%.3f mu2=%.3f nmd=%d  potential=multi-invariant%s\n", fl.D, fl.N, Lext, fl.D,
              beta, kappa, mu2, nmd, frozen ? "  [FROZEN |phi|=1]" : "");
  if (std::getenv("GH_COLD")) GHX_CHECK(ctx, ghx_init_fields(ctx, 0, GHX_COLD, 0.0, 0, 0, 1.0));
  else GHX_CHECK(ctx, ghx_init_fields(ctx, 0, GHX_HOT, 0.8, 1, 12345, 0.3));
  if (frozen) {
    GHX_CHECK(ctx, ghx_set_frozen_phi(ctx, 1));
    GHX_CHECK(ctx, ghx_normalize_phi(ctx));
  }
  GHX_CHECK(ctx, ghx_trajectories(ctx, ntherm, fl.graph ? 1 : 0));

  GHX_CHECK(ctx, ghx_set_counters(ctx, 0, 0, 0));
  cli::Stats plaq, Lphi, Llink;
  std::vector<double> pl_series;
  double sExp = 0, obs[6], dH = 0;
  uint64_t tc = 0, acc = 0;
  for (int t = 0; t < nmeas; ++t) {
    GHX_CHECK(ctx, ghx_trajectories(ctx, 1, fl.graph ? 1 : 0));
    GHX_CHECK(ctx, ghx_get_counters(ctx, 0, &tc, &acc, &dH));
    GHX_CHECK(ctx, ghx_observables(ctx, obs));
    plaq.add(obs[0]); Lphi.add(obs[4]); Llink.add(obs[3]);
    pl_series.push_back(obs[0]);
    sExp += std::exp(-dH);
  }
  if (!ts_out.empty()) {
    std::ofstream fo(ts_out);
    fo << "# traj plaq L_link  (rep=" << name << " beta=" << beta << " kappa=" << kappa << " mu2=" << mu2 << " nmd=" << nmd << ")\n";
    for (int t = 0; t < nmeas; ++t) fo << t << ' ' << pl_series[t] << ' ' << Llink.x[t] << '\n';
  }
  long long vol = 1;
  for (int m = 0; m < fl.D; ++m) vol *= Lext;
  int W = 0;
  const double tau_ll = Llink.tau_int(&W);
  std::printf("plaquette   = %.6f +/- %.6f\n", plaq.mean(), plaq.binned_error());
  std::printf("L_phi       = %.6f +/- %.6f\n", Lphi.mean(), Lphi.binned_error());
  std::printf("L_link      = %.6f +/- %.6f   chi_link=%.4f\n", Llink.mean(), Llink.autocorr_error(), Llink.susceptibility(double(vol)));
  std::printf("autocorr    = tau_int(L_link)=%.2f window=%d  N_eff=%.1f / %d meas\n", tau_ll, W, Llink.neff(), nmeas);
  std::printf("acceptance  = %.4f   <exp(-dH)>=%.5f\n", tc ? double(acc) / tc : 0.0, sExp / nmeas);
  ghx_destroy(ctx);
  return 0;
}
