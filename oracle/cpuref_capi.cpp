// ORACLE — test infrastructure only (see cpuref.hpp). C ABI over the FP64 CPU
// restatement so pytest (ctypes) and bench.py's cpu_baseline leg can drive it.
// Field buffers use the reference's own AoS layouts:
//   U   [s*D+mu][i][j] complex (re,im)      (fields.hpp:13-16, linalg.hpp:17)
//   P,F [s*D+mu][a] real                     (fields.hpp:47-50)
//   phi [s*d+k] complex                      (scalar_field.hpp:12-18)
#include <quadmath.h>
#include "cpuref.hpp"
#include <chrono>
#include <cstring>
#include <memory>

using namespace orc;

namespace {

struct Base {
  virtual ~Base() = default;
  virtual int ncol() const = 0;
  virtual int rep_d() const = 0;
  virtual int rep_real() const = 0;
  virtual long rep_dimT() const = 0;
  virtual int rep_nality() const = 0;
  virtual std::int64_t vol() const = 0;
  virtual void set_couplings(double b, double k, double l, double t, int nmd, int omel, int reunit) = 0;
  virtual void hot(double sigma) = 0;
  virtual void cold() = 0;
  virtual void phi_gaussian(std::uint64_t stream, double sigma) = 0;
  virtual void phi_cold(double c) = 0;
  virtual void refresh() = 0;
  virtual void get(int which, double* out) const = 0;
  virtual void set(int which, const double* in) = 0;
  virtual void forces() = 0;
  virtual void observables(double* out) const = 0;
  virtual void ham_parts(double* out) const = 0;
  virtual void kick(double eps) = 0;
  virtual void drift(double eps) = 0;
  virtual void md_evolve() = 0;
  virtual int trajectory(double* out) = 0;
  virtual void reunitarize() = 0;
  virtual double max_unitarity() const = 0;
  virtual std::uint64_t& traj_count() = 0;
  virtual std::uint64_t& accept_count() = 0;
  virtual void rep_T(double* out) const = 0;
  virtual void rep_W(double* out) const = 0;
  virtual void rep_apply(const double* U, const double* phi, int dag, double* out) const = 0;
  virtual void rep_hop_g(const double* U, const double* px, const double* py, double* out) const = 0;
  virtual void rep_matrix(const double* U, double* out) const = 0;
  virtual void set_frozen(int on) = 0;
  virtual void normalize_phi() = 0;
  virtual int set_multi(const double* f, int nf, double mu2, double* lambda_out) = 0;
  virtual int multi_eval(const double* phi, double* V, double* dV) const = 0;
  virtual double wilson_loop(const char* probe, int Ra, int Rb) const = 0;
  virtual double polyakov_rep(const char* probe, int tdir) const = 0;
  virtual void wilson_rect(long x, int mu, int nu, int Ra, int Rb, double* out) const = 0;
};

template <int N>
struct Impl : Base {
  std::unique_ptr<Rep<N>> rep;
  Engine<N> e;
  Impl(int D, const int* L, std::unique_ptr<Rep<N>> r, std::uint64_t seed)
      : rep(std::move(r)), e(D, L, rep.get(), seed) {}
  int ncol() const override { return N; }
  int rep_d() const override { return rep ? rep->d : 0; }
  int rep_real() const override { return rep ? int(rep->real) : 0; }
  long rep_dimT() const override { return rep ? rep->dimT : 0; }
  int rep_nality() const override { return rep ? rep->nality : 0; }
  std::int64_t vol() const override { return e.g.vol; }
  void set_couplings(double b, double k, double l, double t, int nmd, int omel, int reunit) override {
    e.beta = b; e.kappa = k; e.lambda = l; e.tau = t; e.nmd = nmd; e.omelyan = omel != 0; e.reunit_each_traj = reunit != 0;
  }
  void hot(double sigma) override { e.hot(sigma); }
  void cold() override { e.cold(); }
  void phi_gaussian(std::uint64_t stream, double sigma) override { e.gaussian_field(e.phi, stream, sigma); }
  void phi_cold(double c) override { e.phi_cold(c); }
  void refresh() override { e.refresh_momenta(); }
  // which: 0=U 1=P 2=phi 3=pi 4=F 5=Fphi
  void get(int which, double* out) const override {
    switch (which) {
      case 0: for (std::size_t l = 0; l < e.U.size(); ++l) for (int i = 0; i < N * N; ++i) { out[(l * N * N + i) * 2] = e.U[l].a[i].real(); out[(l * N * N + i) * 2 + 1] = e.U[l].a[i].imag(); } break;
      case 1: case 4: { const auto& v = which == 1 ? e.P : e.F; for (std::size_t l = 0; l < v.size(); ++l) for (int a = 0; a < ngen<N>(); ++a) out[l * ngen<N>() + a] = v[l][a]; } break;
      default: { const auto& v = which == 2 ? e.phi : which == 3 ? e.pi : e.Fphi; for (std::size_t i = 0; i < v.size(); ++i) { out[2 * i] = v[i].real(); out[2 * i + 1] = v[i].imag(); } }
    }
  }
  void set(int which, const double* in) override {
    switch (which) {
      case 0: for (std::size_t l = 0; l < e.U.size(); ++l) for (int i = 0; i < N * N; ++i) e.U[l].a[i] = cd(in[(l * N * N + i) * 2], in[(l * N * N + i) * 2 + 1]); break;
      case 1: for (std::size_t l = 0; l < e.P.size(); ++l) for (int a = 0; a < ngen<N>(); ++a) e.P[l][a] = in[l * ngen<N>() + a]; break;
      case 2: case 3: { auto& v = which == 2 ? e.phi : e.pi; for (std::size_t i = 0; i < v.size(); ++i) v[i] = cd(in[2 * i], in[2 * i + 1]); } break;
      default: break;
    }
  }
  void forces() override { e.forces(); }
  // plaquette, S_g, S_H, L_link, L_phi, polyakov(last dir)
  void observables(double* out) const override {
    out[0] = e.plaquette();
    out[1] = e.gauge_action();
    out[2] = rep ? e.scalar_action() : 0.0;
    out[3] = rep ? e.link_energy() : 0.0;
    out[4] = rep ? e.higgs_length() : 0.0;
    out[5] = e.polyakov(e.D - 1);
  }
  void ham_parts(double* out) const override {
    out[0] = e.kinetic_P();
    out[1] = rep ? e.kinetic_pi() : 0.0;
    out[2] = e.gauge_action();
    out[3] = rep ? e.scalar_action() : 0.0;
    out[4] = e.hamiltonian();
  }
  void kick(double eps) override { e.kick(eps); }
  void drift(double eps) override { e.drift(eps); }
  void md_evolve() override { e.md_evolve(); }
  int trajectory(double* out) override {
    double Hi = 0, Hf = 0;
    const bool acc = e.trajectory(&Hi, &Hf);
    if (out) { out[0] = e.last_dH; out[1] = Hi; out[2] = Hf; }
    return acc ? 1 : 0;
  }
  void reunitarize() override { e.reunitarize_all(); }
  void set_frozen(int on) override { e.frozen = on != 0; }
  std::unique_ptr<Channels<N>> ch;
  // nf < 0: only report the channel count and eigenvalues
  int set_multi(const double* f, int nf, double mu2, double* lambda_out) override {
    if (!rep) return -1;
    if (!ch) ch = std::make_unique<Channels<N>>(*rep);
    if (lambda_out) for (int c = 0; c < ch->n_channels(); ++c) lambda_out[c] = ch->lambda[c];
    if (nf < 0) return ch->n_channels();
    if (nf != ch->n_channels()) return -1;
    e.multi = ch.get();
    e.multi_f.assign(f, f + nf);
    e.multi_mu2 = mu2;
    return ch->n_channels();
  }
  int multi_eval(const double* phi, double* V, double* dV) const override {
    if (!e.multi) return -1;
    DV x(rep->d);
    for (int k = 0; k < rep->d; ++k) x[k] = cd(phi[2 * k], phi[2 * k + 1]);
    *V = e.pot_value(x);
    const DV g = e.pot_grad(x);
    for (int k = 0; k < rep->d; ++k) { dV[2 * k] = g[k].real(); dV[2 * k + 1] = g[k].imag(); }
    return 0;
  }
  // probe spec "fund" | "adj" | Young rows (as make_rep); "fund" uses the plain trace
  double wilson_loop(const char* probe, int Ra, int Rb) const override {
    if (std::string(probe) == "fund") return e.wilson_loop(nullptr, Ra, Rb);
    const Rep<N> r = make_rep<N>(probe);
    return e.wilson_loop(&r, Ra, Rb);
  }
  double polyakov_rep(const char* probe, int tdir) const override {
    if (std::string(probe) == "fund") return e.polyakov_rep(nullptr, tdir);
    const Rep<N> r = make_rep<N>(probe);
    return e.polyakov_rep(&r, tdir);
  }
  void wilson_rect(long x, int mu, int nu, int Ra, int Rb, double* out) const override {
    const Mat<N> W = e.wilson_rect(x, mu, nu, Ra, Rb);
    for (int i = 0; i < N * N; ++i) { out[2 * i] = W.a[i].real(); out[2 * i + 1] = W.a[i].imag(); }
  }
  void normalize_phi() override { e.normalize_phi(); }
  double max_unitarity() const override { return e.max_unitarity_violation(); }
  std::uint64_t& traj_count() override { return e.traj_count; }
  std::uint64_t& accept_count() override { return e.accept_count; }
  void rep_T(double* out) const override {
    const int d = rep->d;
    for (int a = 0; a < ngen<N>(); ++a)
      for (int i = 0; i < d * d; ++i) { out[(std::size_t(a) * d * d + i) * 2] = rep->T[a].a[i].real(); out[(std::size_t(a) * d * d + i) * 2 + 1] = rep->T[a].a[i].imag(); }
  }
  void rep_W(double* out) const override {  // real parts of W[l][i] (W is real)
    for (int l = 0; l < rep->d; ++l) for (long i = 0; i < rep->dimT; ++i) out[std::size_t(l) * rep->dimT + i] = rep->W[l][i].real();
  }
  static Mat<N> load_U(const double* U) { Mat<N> m; for (int i = 0; i < N * N; ++i) m.a[i] = cd(U[2 * i], U[2 * i + 1]); return m; }
  DV load_v(const double* p) const { DV v(rep->d); for (int k = 0; k < rep->d; ++k) v[k] = cd(p[2 * k], p[2 * k + 1]); return v; }
  void rep_apply(const double* U, const double* phi, int dag, double* out) const override {
    const DV r = dag ? rep->rotate_dag(load_U(U), load_v(phi)) : rep->rotate(load_U(U), load_v(phi));
    for (int k = 0; k < rep->d; ++k) { out[2 * k] = r[k].real(); out[2 * k + 1] = r[k].imag(); }
  }
  void rep_hop_g(const double* U, const double* px, const double* py, double* out) const override {
    const Alg<N> gg = rep->hop_g(load_U(U), load_v(px), load_v(py));
    for (int a = 0; a < ngen<N>(); ++a) out[a] = gg[a];
  }
  void rep_matrix(const double* U, double* out) const override {
    const DM M = rep->rep_matrix(load_U(U));
    for (std::size_t i = 0; i < M.a.size(); ++i) { out[2 * i] = M.a[i].real(); out[2 * i + 1] = M.a[i].imag(); }
  }
};

thread_local std::string g_err;

template <int N>
Base* make(int D, const int* L, const char* spec, std::uint64_t seed) {
  std::unique_ptr<Rep<N>> r;
  if (spec && spec[0]) r = std::make_unique<Rep<N>>(make_rep<N>(spec));
  return new Impl<N>(D, L, std::move(r), seed);
}

}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

// rep_spec "" => pure gauge (GaugeHMC semantics); else fund|adj|rows[:real].
void* orc_create(int D, int N, const int* L, const char* rep_spec, std::uint64_t seed) {
  try {
    if (N == 2) return make<2>(D, L, rep_spec, seed);
    if (N == 3) return make<3>(D, L, rep_spec, seed);
    g_err = "N must be 2 or 3";
  } catch (const std::exception& ex) { g_err = ex.what(); }
  return nullptr;
}
void orc_destroy(void* h) { delete static_cast<Base*>(h); }
int orc_rep_dim(void* h) { return static_cast<Base*>(h)->rep_d(); }
int orc_rep_real(void* h) { return static_cast<Base*>(h)->rep_real(); }
long orc_rep_dimT(void* h) { return static_cast<Base*>(h)->rep_dimT(); }
int orc_rep_nality(void* h) { return static_cast<Base*>(h)->rep_nality(); }
std::int64_t orc_vol(void* h) { return static_cast<Base*>(h)->vol(); }
void orc_set_couplings(void* h, double beta, double kappa, double lambda, double tau, int nmd, int omelyan, int reunit) {
  static_cast<Base*>(h)->set_couplings(beta, kappa, lambda, tau, nmd, omelyan, reunit);
}
void orc_hot(void* h, double sigma) { static_cast<Base*>(h)->hot(sigma); }
void orc_cold(void* h) { static_cast<Base*>(h)->cold(); }
void orc_phi_gaussian(void* h, std::uint64_t stream, double sigma) { static_cast<Base*>(h)->phi_gaussian(stream, sigma); }
void orc_phi_cold(void* h, double c) { static_cast<Base*>(h)->phi_cold(c); }
// frozen-length scalar (gauge_higgs_hmc.hpp:38, 61-103)
void orc_set_frozen(void* h, int on) { static_cast<Base*>(h)->set_frozen(on); }
void orc_normalize_phi(void* h) { static_cast<Base*>(h)->normalize_phi(); }
// multi-invariant potential (scalar_invariants.hpp): returns the channel count (or -1)
int orc_set_multi(void* h, const double* f, int nf, double mu2, double* lambda_out) {
  return static_cast<Base*>(h)->set_multi(f, nf, mu2, lambda_out);
}
int orc_multi_eval(void* h, const double* phi, double* V, double* dV) { return static_cast<Base*>(h)->multi_eval(phi, V, dV); }
// measurement layer (observables.hpp:36-101); NaN with orc_last_error set on a bad probe spec
double orc_wilson_loop(void* h, const char* probe, int Ra, int Rb) {
  try { return static_cast<Base*>(h)->wilson_loop(probe, Ra, Rb); } catch (const std::exception& ex) { g_err = ex.what(); }
  return std::nan("");
}
double orc_polyakov_rep(void* h, const char* probe, int tdir) {
  try { return static_cast<Base*>(h)->polyakov_rep(probe, tdir); } catch (const std::exception& ex) { g_err = ex.what(); }
  return std::nan("");
}
void orc_wilson_rect(void* h, long x, int mu, int nu, int Ra, int Rb, double* out) {
  static_cast<Base*>(h)->wilson_rect(x, mu, nu, Ra, Rb, out);
}
void orc_refresh(void* h) { static_cast<Base*>(h)->refresh(); }
void orc_get(void* h, int which, double* out) { static_cast<Base*>(h)->get(which, out); }
void orc_set(void* h, int which, const double* in) { static_cast<Base*>(h)->set(which, in); }
void orc_forces(void* h) { static_cast<Base*>(h)->forces(); }
void orc_observables(void* h, double* out6) { static_cast<Base*>(h)->observables(out6); }
void orc_ham_parts(void* h, double* out5) { static_cast<Base*>(h)->ham_parts(out5); }
void orc_kick(void* h, double eps) { static_cast<Base*>(h)->kick(eps); }
void orc_drift(void* h, double eps) { static_cast<Base*>(h)->drift(eps); }
void orc_md_evolve(void* h) { static_cast<Base*>(h)->md_evolve(); }
int orc_trajectory(void* h, double* out3) { return static_cast<Base*>(h)->trajectory(out3); }
void orc_reunitarize(void* h) { static_cast<Base*>(h)->reunitarize(); }
double orc_max_unitarity(void* h) { return static_cast<Base*>(h)->max_unitarity(); }
std::uint64_t orc_get_traj_count(void* h) { return static_cast<Base*>(h)->traj_count(); }
void orc_set_counts(void* h, std::uint64_t traj, std::uint64_t acc) { static_cast<Base*>(h)->traj_count() = traj; static_cast<Base*>(h)->accept_count() = acc; }
std::uint64_t orc_get_accept_count(void* h) { return static_cast<Base*>(h)->accept_count(); }
void orc_rep_T(void* h, double* out) { static_cast<Base*>(h)->rep_T(out); }
void orc_rep_W(void* h, double* out) { static_cast<Base*>(h)->rep_W(out); }
void orc_rep_apply(void* h, const double* U, const double* phi, int dag, double* out) { static_cast<Base*>(h)->rep_apply(U, phi, dag, out); }
void orc_rep_hop_g(void* h, const double* U, const double* px, const double* py, double* out) { static_cast<Base*>(h)->rep_hop_g(U, px, py, out); }
void orc_rep_matrix(void* h, const double* U, double* out) { static_cast<Base*>(h)->rep_matrix(U, out); }

std::uint64_t orc_splitmix64(std::uint64_t x) { return mix64(x); }
std::uint64_t orc_key(std::uint64_t a, std::uint64_t b, std::uint64_t c, std::uint64_t d, std::uint64_t e) { return stream_key(a, b, c, d, e); }
double orc_uniform(std::uint64_t seed, std::uint64_t k) { return draw_uniform(seed, k); }
double orc_gauss(std::uint64_t seed, std::uint64_t k) { return draw_gauss(seed, k); }
void orc_gauss_many(std::uint64_t seed, const std::uint64_t* keys, long n, double* out) {
  for (long i = 0; i < n; ++i) out[i] = draw_gauss(seed, keys[i]);
}
// The Box-Muller draw with log and cos evaluated in quad precision
// (libquadmath) and rounded to double: the correctly rounded draw, against
// which a disagreement between the device and glibc is attributed.
static double gauss_quad(std::uint64_t seed, std::uint64_t k) {
  const double u1 = draw_uniform(seed, mix64(k * 2 + 1));
  const double u2 = draw_uniform(seed, mix64(k * 2 + 2));
  const double l = double(logq(__float128(u1))), c = double(cosq(__float128(2.0 * kPi * u2)));
  return std::sqrt(-2.0 * l) * c;
}
void orc_gauss_quad_many(std::uint64_t seed, const std::uint64_t* keys, long n, double* out) {
  for (long i = 0; i < n; ++i) out[i] = gauss_quad(seed, keys[i]);
}
// keys key0 .. key0+n-1 (OpenMP): the glibc draw and the correctly rounded draw
void orc_gauss_range(std::uint64_t seed, std::uint64_t key0, long n, double* glibc, double* quad) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) {
    if (glibc) glibc[i] = draw_gauss(seed, key0 + std::uint64_t(i));
    if (quad) quad[i] = gauss_quad(seed, key0 + std::uint64_t(i));
  }
}
// exp(iH) for an N x N Hermitian H (reference algorithm: Taylor-18 or SU(2) closed form)
void orc_expi(int N, const double* H, double* out) {
  if (N == 2) { Mat<2> m; for (int i = 0; i < 4; ++i) m.a[i] = cd(H[2 * i], H[2 * i + 1]); const Mat<2> r = expi<2>(m); for (int i = 0; i < 4; ++i) { out[2 * i] = r.a[i].real(); out[2 * i + 1] = r.a[i].imag(); } }
  else { Mat<3> m; for (int i = 0; i < 9; ++i) m.a[i] = cd(H[2 * i], H[2 * i + 1]); const Mat<3> r = expi<3>(m); for (int i = 0; i < 9; ++i) { out[2 * i] = r.a[i].real(); out[2 * i + 1] = r.a[i].imag(); } }
}
void orc_generators(int N, double* out) {
  if (N == 2) { for (int a = 0; a < 3; ++a) for (int i = 0; i < 4; ++i) { out[(a * 4 + i) * 2] = gens<2>()[a].a[i].real(); out[(a * 4 + i) * 2 + 1] = gens<2>()[a].a[i].imag(); } }
  else { for (int a = 0; a < 8; ++a) for (int i = 0; i < 9; ++i) { out[(a * 9 + i) * 2] = gens<3>()[a].a[i].real(); out[(a * 9 + i) * 2 + 1] = gens<3>()[a].a[i].imag(); } }
}

// Wall-clock of k calls of md_evolve (seconds), for the CPU baseline.
double orc_time_md(void* h, int k) {
  const auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < k; ++i) static_cast<Base*>(h)->md_evolve();
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}
double orc_time_kick_drift(void* h, int k, double eps) {
  const auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < k; ++i) { static_cast<Base*>(h)->kick(eps); static_cast<Base*>(h)->drift(eps); }
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // extern "C"
