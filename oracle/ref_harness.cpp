// Golden-vector harness: drives the REFERENCE headers (compiled in place from
// the reference's src/ tree by oracle/Makefile, target `ref`) and writes raw
// little-endian binaries plus a manifest that tests/golden/make_golden.py packs
// into fixtures. Test infrastructure only — nothing in the product links it.
//
//   ref_harness dump <outdir>            all golden cases
//   ref_harness bench <L> <steps>        time kick+drift MD steps of SU(3)+{2}
#include "hmc/hmc.hpp"
#include "hmc/gauge_higgs_hmc.hpp"
#include "measure/observables.hpp"
#include "measure/autocorr.hpp"
#include "action/scalar_invariants.hpp"
#include "rep/rep_fundamental.hpp"
#include "rep/rep_adjoint.hpp"
#include "rep/rep_general.hpp"
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <fstream>
#include <map>
#include <memory>
#include <random>
#include <sstream>
#include <string>
#include <vector>

using namespace gh;

static std::string g_dir;
static std::ostringstream g_manifest;
static bool g_first = true;

static void put_scalar(const std::string& key, double v) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.17g", v);
  g_manifest << (g_first ? "" : ",\n") << "  \"" << key << "\": " << buf;
  g_first = false;
}
static void put_array(const std::string& key, const std::vector<double>& v) {
  std::ofstream f(g_dir + "/" + key + ".f64", std::ios::binary);
  f.write(reinterpret_cast<const char*>(v.data()), std::streamsize(v.size() * sizeof(double)));
}
static void put_u64(const std::string& key, const std::vector<std::uint64_t>& v) {
  std::ofstream f(g_dir + "/" + key + ".u64", std::ios::binary);
  f.write(reinterpret_cast<const char*>(v.data()), std::streamsize(v.size() * sizeof(std::uint64_t)));
}

template <int D, int N>
static std::vector<double> flat_U(const GaugeField<D, N>& U) {
  std::vector<double> o;
  o.reserve(U.u.size() * N * N * 2);
  for (const auto& m : U.u) for (const auto& z : m.a) { o.push_back(z.real()); o.push_back(z.imag()); }
  return o;
}
template <int D, int N>
static std::vector<double> flat_P(const LinkMom<D, N>& P) {
  std::vector<double> o;
  for (const auto& v : P.p) for (double x : v) o.push_back(x);
  return o;
}
template <int D>
static std::vector<double> flat_phi(const CVecField<D>& f) {
  std::vector<double> o;
  for (const auto& z : f.data) { o.push_back(z.real()); o.push_back(z.imag()); }
  return o;
}

static void dump_rng() {
  std::vector<std::uint64_t> x, mx, keys;
  for (std::uint64_t i = 0; i < 16; ++i) { x.push_back(i); mx.push_back(splitmix64(i)); }
  for (std::uint64_t v : {0xDEADBEEFULL, 0xFFFFFFFFFFFFFFFFULL, 0x123456789ULL}) { x.push_back(v); mx.push_back(splitmix64(v)); }
  put_u64("rng_mix_in", x);
  put_u64("rng_mix_out", mx);
  // keys of the shapes the HMC uses
  std::vector<std::uint64_t> kin;
  for (std::uint64_t t = 0; t < 4; ++t) {
    keys.push_back(Rng::key(0xA1, t)); keys.push_back(Rng::key(0xC3, t)); keys.push_back(Rng::key(0xD4, t));
    keys.push_back(Rng::key(0xB2, t)); keys.push_back(Rng::key(0xE5, t));
  }
  for (std::uint64_t s = 0; s < 8; ++s) for (std::uint64_t mu = 0; mu < 4; ++mu) keys.push_back(Rng::key(1, s, mu, 2, 0));
  put_u64("rng_keys", keys);
  // uniforms and gaussians on a sweep of keys for two seeds
  std::vector<double> uni, gau;
  std::vector<std::uint64_t> gk;
  for (std::uint64_t seed : {1ULL, 12345ULL}) {
    Rng r(seed);
    for (std::uint64_t i = 0; i < 4096; ++i) {
      const std::uint64_t k = Rng::key(0xC3, 7, i, i % 4, i % 3);
      if (seed == 1) gk.push_back(k);
      uni.push_back(r.uniform(k));
      gau.push_back(r.gauss(k));
    }
  }
  put_u64("rng_gauss_keys", gk);
  put_array("rng_uniform", uni);
  put_array("rng_gauss", gau);
}

template <int N>
static void dump_expi(const char* tag, std::uint64_t seed) {
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> g(0.0, 1.3);
  std::vector<double> H, E, Pin, R;
  for (int t = 0; t < 16; ++t) {
    AlgVec<N> v{};
    for (int a = 0; a < n_gen<N>(); ++a) v[a] = g(rng) * (1 + t % 4);
    const Cmat<N> Hm = alg_to_mat<N>(v);
    const Cmat<N> Em = expi<N>(Hm);
    for (const auto& z : Hm.a) { H.push_back(z.real()); H.push_back(z.imag()); }
    for (const auto& z : Em.a) { E.push_back(z.real()); E.push_back(z.imag()); }
    Cmat<N> Pm = Em;  // push off the group, then reunitarize
    for (auto& z : Pm.a) z *= Complex(1.0 + 1e-3 * g(rng), 1e-3 * g(rng));
    for (const auto& z : Pm.a) { Pin.push_back(z.real()); Pin.push_back(z.imag()); }
    reunitarize<N>(Pm);
    for (const auto& z : Pm.a) { R.push_back(z.real()); R.push_back(z.imag()); }
  }
  put_array(std::string(tag) + "_H", H);
  put_array(std::string(tag) + "_expi", E);
  put_array(std::string(tag) + "_reunit_in", Pin);
  put_array(std::string(tag) + "_reunit_out", R);
}

// Pure-gauge case (GaugeHMC): hot start, force, action, one trajectory per integrator.
template <int D, int N>
static void dump_pure(const std::string& tag, int L, double beta, int nmd) {
  std::array<int, D> ext{};
  ext.fill(L);
  for (int integ = 0; integ < 2; ++integ) {
    GaugeHMC<D, N> h(ext, 1);
    h.beta = beta; h.tau = 1.0; h.nmd = nmd;
    h.integ = integ ? Integrator::Omelyan2MN : Integrator::Leapfrog;
    h.U.hot(h.rng, 0.8);
    const std::string t = tag + (integ ? "_omel" : "_lf");
    if (integ == 0) {
      put_array(tag + "_U0", flat_U(h.U));
      LinkMom<D, N> F(h.lat); F.zero(); add_gauge_force<D, N>(h.U, beta, F);
      put_array(tag + "_F", flat_P(F));
      put_scalar(tag + "_Sg", gauge_action<D, N>(h.U, beta));
      put_scalar(tag + "_plaq", avg_plaquette<D, N>(h.U));
      put_scalar(tag + "_poly", polyakov_loop<D, N>(h.U));
      LinkMom<D, N> P0(h.lat); P0.refresh(h.rng, Rng::key(0xA1, 0));
      put_array(tag + "_P0", flat_P(P0));
    }
    std::vector<double> dH, acc;
    for (int k = 0; k < 3; ++k) { acc.push_back(h.trajectory()); dH.push_back(h.last_dH); }
    put_array(t + "_dH", dH);
    put_array(t + "_acc", acc);
    put_array(t + "_U3", flat_U(h.U));
    put_scalar(t + "_plaq3", avg_plaquette<D, N>(h.U));
  }
}

// Gauge+Higgs case: fields, forces, actions, one kick/drift, trajectories.
template <int D, int N>
static void dump_higgs(const std::string& tag, const Representation<N>& rep, int L, double beta,
                       double kappa, double lambda, int nmd, bool with_traj, bool frozen = false) {
  std::array<int, D> ext{};
  ext.fill(L);
  GaugeHiggsHMC<D, N> h(ext, rep, 1);
  h.beta = beta; h.kappa = kappa; h.lambda = lambda; h.tau = 1.0; h.nmd = nmd;
  h.U.hot(h.rng, 0.8);
  h.phi.gaussian(h.rng, 12345, rep.real, 0.3);
  if (frozen) { h.frozen_phi = true; h.normalize_phi(); }  // |phi_x| = 1 scalar
  put_array(tag + "_U0", flat_U(h.U));
  put_array(tag + "_phi0", flat_phi(h.phi));
  // forces on the start configuration (as in kick: zero, gauge, matter; scalar force)
  h.Flink.zero();
  add_gauge_force<D, N>(h.U, beta, h.Flink);
  add_matter_link_force<D, N>(h.phi, h.U, rep, kappa, h.Flink);
  scalar_force<D, N>(h.phi, h.U, rep, kappa, lambda, h.Fphi);
  put_array(tag + "_Flink", flat_P(h.Flink));
  put_array(tag + "_Fphi", flat_phi(h.Fphi));
  put_scalar(tag + "_Sg", gauge_action<D, N>(h.U, beta));
  put_scalar(tag + "_SH", scalar_action<D, N>(h.phi, h.U, rep, kappa, lambda));
  put_scalar(tag + "_plaq", avg_plaquette<D, N>(h.U));
  put_scalar(tag + "_Llink", link_energy<D, N>(h.phi, h.U, rep));
  put_scalar(tag + "_Lphi", higgs_length<D>(h.phi));
  put_scalar(tag + "_poly", polyakov_loop<D, N>(h.U));
  // momenta of trajectory 0
  h.refresh_momenta();
  put_array(tag + "_P0", flat_P(h.P));
  put_array(tag + "_pi0", flat_phi(h.pi));
  put_scalar(tag + "_H0", h.hamiltonian());
  // one kick then one drift
  const auto Usave = h.U.u; const auto phisave = h.phi.data; const auto Psave = h.P.p; const auto pisave = h.pi.data;
  h.kick(0.05);
  put_array(tag + "_P_kick", flat_P(h.P));
  put_array(tag + "_pi_kick", flat_phi(h.pi));
  h.drift(0.05);
  put_array(tag + "_U_drift", flat_U(h.U));
  put_array(tag + "_phi_drift", flat_phi(h.phi));
  if (frozen) put_array(tag + "_pi_drift", flat_phi(h.pi));  // the geodesic drift also turns pi
  h.U.u = Usave; h.phi.data = phisave; h.P.p = Psave; h.pi.data = pisave;
  if (!with_traj) return;
  for (int integ = 0; integ < 2; ++integ) {
    h.U.u = Usave; h.phi.data = phisave; h.traj_count = 0; h.accept_count = 0;
    h.integ = integ ? Integrator::Omelyan2MN : Integrator::Leapfrog;
    const std::string t = tag + (integ ? "_omel" : "_lf");
    std::vector<double> dH, acc;
    for (int k = 0; k < 2; ++k) { acc.push_back(h.trajectory()); dH.push_back(h.last_dH); }
    put_array(t + "_dH", dH);
    put_array(t + "_acc", acc);
    put_array(t + "_U2", flat_U(h.U));
    put_array(t + "_phi2", flat_phi(h.phi));
    put_scalar(t + "_plaq2", avg_plaquette<D, N>(h.U));
    put_scalar(t + "_Llink2", link_energy<D, N>(h.phi, h.U, rep));
    put_scalar(t + "_Lphi2", higgs_length<D>(h.phi));
  }
}

// Measurement layer (observables.hpp:36-101) on a thermalised pure-gauge
// field of an anisotropic lattice: Wilson loops and Polyakov loops traced in
// several probe irreps. Keys: <tag>_W_<probe>_<Ra>x<Rb>, <tag>_P_<probe>_t<dir>.
template <int D, int N>
static void dump_loops(const std::string& tag, const std::array<int, D>& ext, double beta, int ntraj,
                       const std::vector<std::pair<std::string, const Representation<N>*>>& probes,
                       const std::vector<std::pair<int, int>>& sizes, const std::vector<std::string>& small_only) {
  GaugeHMC<D, N> h(ext, 3);
  h.beta = beta; h.nmd = 12;
  h.U.hot(h.rng, 0.35);
  for (int t = 0; t < ntraj; ++t) h.trajectory();
  put_array(tag + "_U", flat_U(h.U));
  for (const auto& sz : sizes)
    put_scalar(tag + "_Wfund_" + std::to_string(sz.first) + "x" + std::to_string(sz.second),
               wilson_loop_fund<D, N>(h.U, sz.first, sz.second));
  for (const auto& pr : probes) {
    const bool small = std::find(small_only.begin(), small_only.end(), pr.first) != small_only.end();
    for (const auto& sz : sizes) {
      if (small && sz.first * sz.second > 4) continue;
      put_scalar(tag + "_W_" + pr.first + "_" + std::to_string(sz.first) + "x" + std::to_string(sz.second),
                 wilson_loop_rep<D, N>(h.U, *pr.second, sz.first, sz.second));
    }
    for (int t = 0; t < D; ++t) put_scalar(tag + "_P_" + pr.first + "_t" + std::to_string(t), polyakov_loop_rep<D, N>(h.U, *pr.second, t));
  }
  // one rectangle product for the path order (x = 5, mu = 0, nu = D-1, 2 x 3)
  const Cmat<N> W = wilson_rect<D, N>(h.U, 5, 0, D - 1, 2, 3);
  std::vector<double> w;
  for (const auto& z : W.a) { w.push_back(z.real()); w.push_back(z.imag()); }
  put_array(tag + "_rect", w);
}

// Time-series analysis (observables.hpp:104-145 Stats, autocorr.hpp:55-212) of
// AR(1) series x_{t+1} = a x_t + sqrt(1-a^2) eta_t and a constant series.
static void dump_series(const std::string& tag, double a, int n, std::uint64_t seed) {
  std::vector<Real> x(static_cast<std::size_t>(n));
  if (a >= 0.0) {
    std::mt19937_64 rng(seed);
    std::normal_distribution<Real> g(0.0, 1.0);
    const Real s = std::sqrt(1.0 - a * a);
    x[0] = g(rng);
    for (int t = 1; t < n; ++t) x[t] = a * x[t - 1] + s * g(rng);
  } else {
    std::fill(x.begin(), x.end(), 3.7);
  }
  put_array(tag + "_x", x);
  Stats st;
  for (Real v : x) st.add(v);
  int W = 0;
  const Real tau = tau_int(x, &W);
  put_array(tag + "_out", {st.mean(), st.var(), st.binder(), st.binned_error(), tau, double(W),
                           effective_sample_size(x), naive_error(x), autocorr_error(x)});
  std::vector<double> be;
  for (int bs : {1, 3, 16, 50}) be.push_back(blocking_error(x, bs));
  put_array(tag + "_blocking", be);
  put_array(tag + "_rho", autocorr_function(x, 12));
}

// Multi-invariant quartic potential (scalar_invariants.hpp): channels, V and
// dV/dphi^* on random phi, then the HMC with this potential plugged in.
template <int D, int N>
static void dump_multi(const std::string& tag, const Representation<N>& rep, int L, double beta, double kappa,
                       double mu2, int nmd) {
  CasimirChannels<N> ch(rep);
  put_scalar(tag + "_nch", ch.n_channels());
  put_array(tag + "_lambda", ch.lambda);
  std::vector<double> f;
  for (int c = 0; c < ch.n_channels(); ++c) f.push_back(0.2 + 0.15 * c);
  put_array(tag + "_f", f);
  MultiInvariantPotential<N> pot(ch, f, mu2);
  std::mt19937_64 rng(41);
  std::normal_distribution<double> g(0.0, 0.7);
  std::vector<double> px, val, dv, dvu;
  for (int t = 0; t < 3; ++t) {
    DVec x(rep.d);
    for (int k = 0; k < rep.d; ++k) x(k) = Complex(g(rng), rep.real ? 0.0 : g(rng));
    for (int k = 0; k < rep.d; ++k) { px.push_back(x(k).real()); px.push_back(x(k).imag()); }
    val.push_back(pot.value(x));
    const DVec a = pot.dV_dphibar(x), b = ch.dV_dphibar(x, f, mu2);
    for (int k = 0; k < rep.d; ++k) { dv.push_back(a(k).real()); dv.push_back(a(k).imag()); dvu.push_back(b(k).real()); dvu.push_back(b(k).imag()); }
  }
  put_array(tag + "_px", px); put_array(tag + "_V", val); put_array(tag + "_dV", dv); put_array(tag + "_dV_uncached", dvu);
  std::array<int, D> ext{};
  ext.fill(L);
  GaugeHiggsHMC<D, N> h(ext, rep, 1);
  h.beta = beta; h.kappa = kappa; h.tau = 1.0; h.nmd = nmd; h.potential = &pot;
  h.U.hot(h.rng, 0.8);
  h.phi.gaussian(h.rng, 12345, rep.real, 0.3);
  h.Flink.zero();
  add_gauge_force<D, N>(h.U, beta, h.Flink);
  add_matter_link_force<D, N>(h.phi, h.U, rep, kappa, h.Flink);
  scalar_force<D, N>(h.phi, h.U, rep, kappa, pot, h.Fphi);
  put_array(tag + "_Fphi", flat_phi(h.Fphi));
  put_scalar(tag + "_SH", scalar_action<D, N>(h.phi, h.U, rep, kappa, pot));
  h.refresh_momenta();
  put_scalar(tag + "_H0", h.hamiltonian());
  const auto Usave = h.U.u; const auto phisave = h.phi.data;
  for (int integ = 0; integ < 2; ++integ) {
    h.U.u = Usave; h.phi.data = phisave; h.traj_count = 0; h.accept_count = 0;
    h.integ = integ ? Integrator::Omelyan2MN : Integrator::Leapfrog;
    const std::string t = tag + (integ ? "_omel" : "_lf");
    std::vector<double> dH, acc;
    for (int k = 0; k < 2; ++k) { acc.push_back(h.trajectory()); dH.push_back(h.last_dH); }
    put_array(t + "_dH", dH);
    put_array(t + "_acc", acc);
    put_array(t + "_U2", flat_U(h.U));
    put_array(t + "_phi2", flat_phi(h.phi));
  }
}

template <int N>
static void dump_rep(const std::string& tag, const GeneralRep<N>& rep, std::uint64_t seed) {
  put_scalar(tag + "_d", rep.d);
  put_scalar(tag + "_dimT", double(rep.dimT));
  put_scalar(tag + "_nality", rep.nality);
  std::vector<double> Wr;
  for (int l = 0; l < rep.d; ++l) for (long i = 0; i < rep.dimT; ++i) Wr.push_back(rep.W[l][i].real());
  double wim = 0.0;
  for (int l = 0; l < rep.d; ++l) for (long i = 0; i < rep.dimT; ++i) wim = std::max(wim, std::fabs(rep.W[l][i].imag()));
  put_scalar(tag + "_W_max_imag", wim);
  put_array(tag + "_W", Wr);
  std::vector<double> T;
  for (int a = 0; a < n_gen<N>(); ++a) for (const auto& z : rep.T[a].a) { T.push_back(z.real()); T.push_back(z.imag()); }
  put_array(tag + "_T", T);
  // applies on random links/vectors
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> g(0.0, 0.8);
  std::vector<double> Us, px, py, rot, rotd, hop;
  for (int t = 0; t < 4; ++t) {
    AlgVec<N> v{};
    for (int a = 0; a < n_gen<N>(); ++a) v[a] = g(rng);
    const Cmat<N> U = expi<N>(alg_to_mat<N>(v));
    DVec x(rep.d), y(rep.d);
    for (int k = 0; k < rep.d; ++k) { x(k) = Complex(g(rng), g(rng)); y(k) = Complex(g(rng), g(rng)); }
    for (const auto& z : U.a) { Us.push_back(z.real()); Us.push_back(z.imag()); }
    for (int k = 0; k < rep.d; ++k) { px.push_back(x(k).real()); px.push_back(x(k).imag()); py.push_back(y(k).real()); py.push_back(y(k).imag()); }
    const DVec r1 = rep.rotate(U, y), r2 = rep.rotate_dag(U, y);
    for (int k = 0; k < rep.d; ++k) { rot.push_back(r1(k).real()); rot.push_back(r1(k).imag()); rotd.push_back(r2(k).real()); rotd.push_back(r2(k).imag()); }
    const AlgVec<N> gg = rep.hop_link_g(U, x, y);
    for (double q : gg) hop.push_back(q);
  }
  put_array(tag + "_app_U", Us); put_array(tag + "_app_px", px); put_array(tag + "_app_py", py);
  put_array(tag + "_app_rot", rot); put_array(tag + "_app_rotdag", rotd); put_array(tag + "_app_hop", hop);
}

static int dump(const char* dir) {
  g_dir = dir;
  dump_rng();
  dump_expi<2>("expi2", 7);
  dump_expi<3>("expi3", 8);
  dump_pure<4, 2>("su2_pure", 4, 2.3, 20);
  dump_pure<4, 3>("su3_pure", 4, 5.7, 12);
  { FundamentalRep<2> f; dump_higgs<4, 2>("su2_fund", f, 4, 2.3, 0.265, 0.5, 20, true); }
  { FundamentalRep<2> f; dump_higgs<4, 2>("su2_fund_frozen", f, 4, 2.3, 0.6, 0.5, 20, true, true); }
  { AdjointRep<2> a; dump_higgs<3, 2>("su2_adj_frozen", a, 4, 2.3, 0.5, 0.5, 12, true, true); }
  { GeneralRep<2> s4({4}); dump_higgs<3, 2>("su2_4_frozen", s4, 3, 2.3, 0.5, 0.5, 10, true, true); }
  { GeneralRep<3> s21({2, 1}); dump_higgs<3, 3>("su3_21_frozen", s21, 3, 5.7, 0.3, 0.5, 10, true, true); }
  { FundamentalRep<2> f; dump_multi<4, 2>("multi_su2_fund", f, 4, 2.3, 0.2, 0.4, 12); }
  { AdjointRep<2> a; dump_multi<3, 2>("multi_su2_adj", a, 4, 2.3, 0.2, 0.4, 12); }
  { GeneralRep<3> s2({2}); dump_multi<4, 3>("multi_su3_2", s2, 4, 5.7, 0.1, 0.5, 12); }
  { GeneralRep<3> s21({2, 1}); dump_multi<3, 3>("multi_su3_21", s21, 3, 5.7, 0.1, 0.5, 10); }
  { GeneralRep<2> s12({12}); dump_multi<3, 2>("multi_su2_12", s12, 3, 2.3, 0.1, 0.4, 10); }
  { AdjointRep<2> a; dump_higgs<3, 2>("su2_adj", a, 4, 2.3, 0.3, 0.5, 12, true); }
  { FundamentalRep<3> f; dump_higgs<3, 3>("su3_fund", f, 4, 5.7, 0.2, 0.5, 12, true); }
  {
    GeneralRep<3> s({2});
    dump_rep<3>("rep_su3_2", s, 31);
    dump_higgs<4, 3>("su3_sym2", s, 4, 5.7, 0.1, 0.5, 20, true);
  }
  {
    GeneralRep<3> s({1, 1});
    dump_rep<3>("rep_su3_11", s, 32);
  }
  {
    GeneralRep<2> s({3});
    dump_rep<2>("rep_su2_3", s, 33);
  }
  {
    GeneralRep<3> s({6, 1});
    dump_rep<3>("rep_su3_61", s, 34);
    dump_higgs<4, 3>("su3_61", s, 3, 5.7, 0.05, 0.5, 16, false);
  }
  {
    const std::vector<std::pair<int, int>> sizes = {{1, 1}, {1, 2}, {2, 1}, {2, 2}, {1, 3}, {3, 2}};
    FundamentalRep<3> f; AdjointRep<3> a;
    GeneralRep<3> s2({2}), s11({1, 1}), s21({2, 1}), s3({3}), s22({2, 2}), s61({6, 1});
    dump_loops<4, 3>("loops_su3", {4, 3, 4, 5}, 5.7, 3,
                     {{"fund", &f}, {"adj", &a}, {"2", &s2}, {"11", &s11}, {"21", &s21}, {"3", &s3}, {"22", &s22}, {"61", &s61}},
                     sizes, {"61"});
    FundamentalRep<2> f2; AdjointRep<2> a2;
    GeneralRep<2> t2({2}), t3({3}), t4({4});
    dump_loops<3, 2>("loops_su2", {5, 4, 3}, 2.3, 3, {{"fund", &f2}, {"adj", &a2}, {"2", &t2}, {"3", &t3}, {"4", &t4}}, sizes, {});
  }
  dump_series("series_white", 0.0, 1000, 0xA1u);
  dump_series("series_ar05", 0.5, 4000, 0xB2u);
  dump_series("series_ar09", 0.9, 8000, 0xC3u);
  dump_series("series_ar099", 0.99, 3000, 0xD4u);
  dump_series("series_const", -1.0, 1000, 0);
  std::ofstream m(std::string(dir) + "/manifest.json");
  m << "{\n" << g_manifest.str() << "\n}\n";
  return 0;
}

// CPU baseline of the reference itself: SU(3)+{2}, D=4, L^4, `steps` kick+drift pairs.
static int bench(int L, int steps) {
  GeneralRep<3> rep({2});
  std::array<int, 4> ext{};
  ext.fill(L);
  GaugeHiggsHMC<4, 3> h(ext, rep, 1);
  h.beta = 5.7; h.kappa = 0.1; h.lambda = 0.5; h.tau = 1.0; h.nmd = 20;
  h.U.hot(h.rng, 0.8);
  h.phi.gaussian(h.rng, 12345, rep.real, 1.0);
  h.refresh_momenta();
  const double eps = 1.0 / 20;
  h.kick(0.5 * eps);  // warm
  const auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < steps; ++i) { h.drift(eps); h.kick(eps); }
  const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::printf("{\"L\": %d, \"steps\": %d, \"seconds\": %.6f, \"site_updates_per_s\": %.6e}\n", L, steps, sec,
              double(h.lat.vol) * steps / sec);
  return 0;
}

int main(int argc, char** argv) {
  if (argc >= 3 && std::string(argv[1]) == "dump") return dump(argv[2]);
  if (argc >= 4 && std::string(argv[1]) == "bench") return bench(std::atoi(argv[2]), std::atoi(argv[3]));
  std::fprintf(stderr, "usage: %s dump <outdir> | bench <L> <steps>\n", argv[0]);
  return 2;
}
