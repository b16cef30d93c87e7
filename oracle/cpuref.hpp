// ORACLE — test infrastructure only. FP64 CPU restatement of the reference
// gauge+Higgs HMC (hlammiv/gauge_higgs). Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load it; the product library never does.
//
// The arithmetic of every routine below follows the reference expression for
// expression (std::complex<double>, same loop orders, same zero skips). Integer
// RNG streams, uniforms, Gaussians and start fields then agree bit for bit with
// the reference compiled by the same g++; forces, actions and MD fields agree to
// a few ulps (g++ contracts FMAs per call site). Each routine cites the
// reference file:line it restates.
//
// Parity status per function (how each is pinned) is listed in oracle/README.md.
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

using cd = std::complex<double>;
constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double kOmelyanLambda = 0.1931833275037836;  // config.hpp:25

// ---------------------------------------------------------------------------
// Counter-based RNG — rng.hpp:11-46
// ---------------------------------------------------------------------------
inline std::uint64_t mix64(std::uint64_t x) {  // rng.hpp:11-16 (SplitMix64 finaliser)
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}
// rng.hpp:24-31: five key fields chained through the mixer.
inline std::uint64_t stream_key(std::uint64_t a, std::uint64_t b = 0, std::uint64_t c = 0,
                                std::uint64_t d = 0, std::uint64_t e = 0) {
  const std::uint64_t g = 0x9E3779B97F4A7C15ULL;
  std::uint64_t h = mix64(a);
  for (std::uint64_t f : {b, c, d, e}) h = mix64(h ^ (f + g));
  return h;
}
// rng.hpp:35-39: 53-bit uniform in (0,1).
inline double draw_uniform(std::uint64_t seed, std::uint64_t k) {
  const std::uint64_t h = mix64(seed ^ mix64(k));
  return ((h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
// rng.hpp:42-46: Box-Muller from two sub-keyed uniforms (glibc log/cos/sqrt).
inline double draw_gauss(std::uint64_t seed, std::uint64_t k) {
  const double u1 = draw_uniform(seed, mix64(k * 2 + 1));
  const double u2 = draw_uniform(seed, mix64(k * 2 + 2));
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
}

// ---------------------------------------------------------------------------
// Lattice geometry — geometry.hpp:12-53 (direction 0 fastest, periodic)
// ---------------------------------------------------------------------------
struct Geom {
  int D = 0;
  std::vector<int> L;
  std::vector<std::int64_t> stride, fwd, bwd;
  std::int64_t vol = 0;
  Geom() = default;
  Geom(int dim, const int* ext) : D(dim), L(ext, ext + dim), stride(dim) {
    vol = 1;
    for (int m = 0; m < D; ++m) { stride[m] = vol; vol *= L[m]; }
    fwd.resize(vol * D);
    bwd.resize(vol * D);
    std::vector<int> x(D);
    for (std::int64_t s = 0; s < vol; ++s) {
      std::int64_t r = s;
      for (int m = 0; m < D; ++m) { x[m] = int(r % L[m]); r /= L[m]; }
      for (int m = 0; m < D; ++m) {
        const std::int64_t up = (x[m] + 1 == L[m]) ? -std::int64_t(x[m]) : 1;
        const std::int64_t dn = (x[m] == 0) ? std::int64_t(L[m] - 1) : -1;
        fwd[s * D + m] = s + up * stride[m];
        bwd[s * D + m] = s + dn * stride[m];
      }
    }
  }
  std::int64_t up(std::int64_t s, int m) const { return fwd[s * D + m]; }
  std::int64_t dn(std::int64_t s, int m) const { return bwd[s * D + m]; }
  std::int64_t n_plaq() const { return vol * (std::int64_t(D) * (D - 1) / 2); }  // geometry.hpp:51
  int coord(std::int64_t s, int m) const { return int((s / stride[m]) % L[m]); }
};

// ---------------------------------------------------------------------------
// N x N complex matrices — linalg.hpp:17-125 (row-major a[i*N+j])
// ---------------------------------------------------------------------------
template <int N>
struct Mat {
  std::array<cd, N * N> a{};
  cd& operator()(int i, int j) { return a[i * N + j]; }
  const cd& operator()(int i, int j) const { return a[i * N + j]; }
  static Mat eye() { Mat m; for (int i = 0; i < N; ++i) m(i, i) = cd(1.0, 0.0); return m; }
  Mat& operator+=(const Mat& o) { for (std::size_t i = 0; i < a.size(); ++i) a[i] += o.a[i]; return *this; }
  Mat& operator-=(const Mat& o) { for (std::size_t i = 0; i < a.size(); ++i) a[i] -= o.a[i]; return *this; }
  Mat& operator*=(cd s) { for (auto& z : a) z *= s; return *this; }
  Mat dag() const { Mat r; for (int i = 0; i < N; ++i) for (int j = 0; j < N; ++j) r(i, j) = std::conj((*this)(j, i)); return r; }
  cd tr() const { cd t(0.0, 0.0); for (int i = 0; i < N; ++i) t += (*this)(i, i); return t; }
  double fnorm() const { double s = 0.0; for (const auto& z : a) s += std::norm(z); return std::sqrt(s); }
};
template <int N>
Mat<N> operator*(const Mat<N>& A, const Mat<N>& B) {  // linalg.hpp:66-75 (i,k,j order, zero skip)
  Mat<N> C;
  for (int i = 0; i < N; ++i)
    for (int k = 0; k < N; ++k) {
      const cd aik = A(i, k);
      if (aik == cd(0.0, 0.0)) continue;
      for (int j = 0; j < N; ++j) C(i, j) += aik * B(k, j);
    }
  return C;
}
template <int N> Mat<N> operator*(Mat<N> A, cd s) { A *= s; return A; }
template <int N> Mat<N> operator+(Mat<N> A, const Mat<N>& B) { A += B; return A; }
template <int N> Mat<N> operator-(Mat<N> A, const Mat<N>& B) { A -= B; return A; }

template <int N>
cd tr_prod(const Mat<N>& A, const Mat<N>& B) {  // linalg.hpp:91-96
  cd s(0.0, 0.0);
  for (int i = 0; i < N; ++i) for (int k = 0; k < N; ++k) s += A(i, k) * B(k, i);
  return s;
}
template <int N>
double re_tr_prod(const Mat<N>& A, const Mat<N>& B) {  // linalg.hpp:83-88
  double s = 0.0;
  for (int i = 0; i < N; ++i) for (int k = 0; k < N; ++k) s += (A(i, k) * B(k, i)).real();
  return s;
}
template <int N>
cd det(const Mat<N>& M) {  // linalg.hpp:100-125: elimination with partial pivoting
  std::array<cd, N * N> m = M.a;
  auto at = [&](int i, int j) -> cd& { return m[i * N + j]; };
  cd d(1.0, 0.0);
  for (int col = 0; col < N; ++col) {
    int piv = col;
    double best = std::abs(at(col, col));
    for (int r = col + 1; r < N; ++r) { const double v = std::abs(at(r, col)); if (v > best) { best = v; piv = r; } }
    if (best == 0.0) return cd(0.0, 0.0);
    if (piv != col) { for (int j = 0; j < N; ++j) std::swap(at(piv, j), at(col, j)); d = -d; }
    d *= at(col, col);
    const cd inv = cd(1.0, 0.0) / at(col, col);
    for (int r = col + 1; r < N; ++r) {
      const cd f = at(r, col) * inv;
      if (f == cd(0.0, 0.0)) continue;
      for (int j = col; j < N; ++j) at(r, j) -= f * at(col, j);
    }
  }
  return d;
}

// ---------------------------------------------------------------------------
// su(N) generators, Bertlmann-Krammer order — algebra.hpp:24-50
// ---------------------------------------------------------------------------
template <int N>
constexpr int ngen() { return N * N - 1; }
template <int N>
using Alg = std::array<double, N * N - 1>;

template <int N>
const std::vector<Mat<N>>& gens() {
  static const std::vector<Mat<N>> T = [] {
    std::vector<Mat<N>> g(ngen<N>());
    int a = 0;
    for (int j = 0; j < N; ++j) for (int k = j + 1; k < N; ++k) {  // symmetric pairs
      g[a](j, k) += cd(0.5, 0.0); g[a](k, j) += cd(0.5, 0.0); ++a;
    }
    for (int j = 0; j < N; ++j) for (int k = j + 1; k < N; ++k) {  // antisymmetric pairs
      g[a](j, k) += cd(0.0, -0.5); g[a](k, j) += cd(0.0, 0.5); ++a;
    }
    for (int l = 1; l <= N - 1; ++l) {  // diagonal
      const double c = std::sqrt(2.0 / (double(l) * (l + 1)));
      for (int j = 0; j < l; ++j) g[a](j, j) += cd(0.5 * c, 0.0);
      g[a](l, l) += cd(-0.5 * c * l, 0.0);
      ++a;
    }
    return g;
  }();
  return T;
}
template <int N>
Mat<N> alg_to_mat(const Alg<N>& v) {  // algebra.hpp:60-71
  const auto& T = gens<N>();
  Mat<N> M;
  for (int a = 0; a < ngen<N>(); ++a) {
    const double va = v[a];
    if (va == 0.0) continue;
    for (int i = 0; i < N; ++i) for (int j = 0; j < N; ++j) M(i, j) += va * T[a](i, j);
  }
  return M;
}
template <int N>
Alg<N> mat_to_alg(const Mat<N>& M) {  // algebra.hpp:76-81
  Alg<N> v{};
  for (int a = 0; a < ngen<N>(); ++a) v[a] = 2.0 * re_tr_prod(gens<N>()[a], M);
  return v;
}
template <int N>
std::vector<double> structure_f() {  // algebra.hpp:103-120: f^{abc} = 2 Im Tr([Ta,Tb]Tc)
  const auto& T = gens<N>();
  const int n = ngen<N>();
  std::vector<double> f(std::size_t(n) * n * n, 0.0);
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b) {
      const Mat<N> cm = T[a] * T[b] - T[b] * T[a];
      for (int c = 0; c < n; ++c) f[(std::size_t(a) * n + b) * n + c] = 2.0 * tr_prod(cm, T[c]).imag();
    }
  return f;
}

// exp(iH): scaled Taylor-18 then squaring — sun.hpp:12-30
template <int N>
Mat<N> expi(const Mat<N>& H) {
  const double nrm = H.fnorm();
  int s = 0;
  double scaled = nrm;
  while (scaled > 0.5) { scaled *= 0.5; ++s; }
  const double inv2s = std::pow(0.5, s);
  Mat<N> A = H * cd(0.0, inv2s);
  Mat<N> term = Mat<N>::eye(), result = Mat<N>::eye();
  for (int k = 1; k <= 18; ++k) {
    term = term * A;
    term *= cd(1.0 / k, 0.0);
    result += term;
  }
  for (int i = 0; i < s; ++i) result = result * result;
  return result;
}
// SU(2) closed form — sun.hpp:34-46
template <>
inline Mat<2> expi<2>(const Mat<2>& H) {
  const double r2 = 0.5 * tr_prod(H, H).real();
  const double r = std::sqrt(std::max(r2, 0.0));
  double c, sinc;
  if (r < 1e-8) { c = 1.0 - 0.5 * r2; sinc = 1.0 - r2 / 6.0; }
  else { c = std::cos(r); sinc = std::sin(r) / r; }
  Mat<2> U = Mat<2>::eye();
  U *= cd(c, 0.0);
  Mat<2> iH = H * cd(0.0, sinc);
  U += iH;
  return U;
}
// Modified Gram-Schmidt on rows, then det phase removed from row 0 — sun.hpp:49-72
template <int N>
void reunitarize(Mat<N>& U) {
  for (int i = 0; i < N; ++i) {
    for (int k = 0; k < i; ++k) {
      cd proj(0.0, 0.0);
      for (int j = 0; j < N; ++j) proj += std::conj(U(k, j)) * U(i, j);
      for (int j = 0; j < N; ++j) U(i, j) -= proj * U(k, j);
    }
    double nrm = 0.0;
    for (int j = 0; j < N; ++j) nrm += std::norm(U(i, j));
    nrm = std::sqrt(nrm);
    const cd inv(1.0 / nrm, 0.0);
    for (int j = 0; j < N; ++j) U(i, j) *= inv;
  }
  const cd d = det(U);
  const double ad = std::abs(d);
  if (ad > 0.0) {
    const cd ph = d / ad;
    const cd inv = std::conj(ph);
    for (int j = 0; j < N; ++j) U(0, j) *= inv;
  }
}
template <int N>
double unitarity_violation(const Mat<N>& U) {  // sun.hpp:75-78
  Mat<N> d = U.dag() * U;
  for (int i = 0; i < N; ++i) d(i, i) -= cd(1.0, 0.0);
  return d.fnorm();
}

// ---------------------------------------------------------------------------
// Dynamic complex matrix/vector — linalg.hpp:159-307
// ---------------------------------------------------------------------------
struct DM {
  int r = 0, c = 0;
  std::vector<cd> a;
  DM() = default;
  DM(int rr, int cc) : r(rr), c(cc), a(std::size_t(rr) * cc, cd(0, 0)) {}
  cd& operator()(int i, int j) { return a[std::size_t(i) * c + j]; }
  const cd& operator()(int i, int j) const { return a[std::size_t(i) * c + j]; }
  DM dag() const { DM o(c, r); for (int i = 0; i < r; ++i) for (int j = 0; j < c; ++j) o(j, i) = std::conj((*this)(i, j)); return o; }
  cd tr() const { cd t(0, 0); for (int i = 0; i < std::min(r, c); ++i) t += (*this)(i, i); return t; }
};
using DV = std::vector<cd>;
inline DM operator*(const DM& A, const DM& B) {  // linalg.hpp:175-185
  DM C(A.r, B.c);
  for (int i = 0; i < A.r; ++i)
    for (int k = 0; k < A.c; ++k) {
      const cd aik = A(i, k);
      if (aik == cd(0, 0)) continue;
      for (int j = 0; j < B.c; ++j) C(i, j) += aik * B(k, j);
    }
  return C;
}
inline DM operator-(const DM& A, const DM& B) { DM C(A.r, A.c); for (std::size_t i = 0; i < A.a.size(); ++i) C.a[i] = A.a[i] - B.a[i]; return C; }
inline DM operator+(const DM& A, const DM& B) { DM C(A.r, A.c); for (std::size_t i = 0; i < A.a.size(); ++i) C.a[i] = A.a[i] + B.a[i]; return C; }
inline DV matvec(const DM& M, const DV& x) {  // linalg.hpp:302-307 (no zero skip)
  DV y(M.r, cd(0, 0));
  for (int i = 0; i < M.r; ++i) for (int j = 0; j < M.c; ++j) y[i] += M(i, j) * x[j];
  return y;
}
inline cd vdot(const DV& x, const DV& y) {  // x^dag y, linalg.hpp:297-301
  cd s(0, 0);
  for (std::size_t i = 0; i < x.size(); ++i) s += std::conj(x[i]) * y[i];
  return s;
}
inline double vnorm2(const DV& x) { double s = 0; for (auto& z : x) s += std::norm(z); return s; }
template <int N>
DM to_dm(const Mat<N>& M) { DM D(N, N); for (int i = 0; i < N; ++i) for (int j = 0; j < N; ++j) D(i, j) = M(i, j); return D; }

// ---------------------------------------------------------------------------
// Representations — representation.hpp:13-107, rep_fundamental.hpp,
// rep_adjoint.hpp, rep_general.hpp (tensor / Young-symmetrizer engine)
// ---------------------------------------------------------------------------
enum class RepKind { Fund, Adj, Tensor };

template <int N>
struct Rep {
  RepKind kind = RepKind::Fund;
  int d = 0, nality = 0;
  bool real = false;
  std::vector<DM> T;          // T^a_R, d x d
  // tensor engine
  std::vector<int> rows;
  int nbox = 0;
  long dimT = 0;
  std::vector<long> powN;
  std::vector<std::vector<cd>> W;  // orthonormal basis rows, length dimT
  std::vector<cd> Wflat;

  static Rep fund() {  // rep_fundamental.hpp:16-25
    Rep r; r.kind = RepKind::Fund; r.d = N; r.nality = 1; r.real = false;
    for (int a = 0; a < ngen<N>(); ++a) r.T.push_back(to_dm<N>(gens<N>()[a]));
    return r;
  }
  static Rep adj() {  // rep_adjoint.hpp:11-24
    Rep r; r.kind = RepKind::Adj;
    const int n = ngen<N>();
    r.d = n; r.nality = 0; r.real = true;
    const auto f = structure_f<N>();
    for (int a = 0; a < n; ++a) {
      DM M(n, n);
      for (int b = 0; b < n; ++b)
        for (int c = 0; c < n; ++c) M(b, c) = cd(0.0, -f[(std::size_t(a) * n + b) * n + c]);
      r.T.push_back(M);
    }
    return r;
  }
  // rep_general.hpp:54-84
  static Rep young(std::vector<int> rws, bool real_rep) {
    while (!rws.empty() && rws.back() == 0) rws.pop_back();
    if (rws.empty()) throw std::runtime_error("GeneralRep: empty diagram (singlet not supported)");
    for (std::size_t i = 0; i + 1 < rws.size(); ++i)
      if (rws[i] < rws[i + 1]) throw std::runtime_error("GeneralRep: rows must be non-increasing");
    if (int(rws.size()) > N - 1) throw std::runtime_error("GeneralRep: > N-1 rows (strip full columns of N)");
    Rep r; r.kind = RepKind::Tensor; r.rows = rws;
    r.nbox = std::accumulate(rws.begin(), rws.end(), 0);
    r.powN.assign(r.nbox + 1, 1);
    for (int k = 1; k <= r.nbox; ++k) r.powN[k] = r.powN[k - 1] * N;
    r.dimT = r.powN[r.nbox];
    if (r.dimT > 200000) throw std::runtime_error("GeneralRep: N^n exceeds cap; irrep too large for tensor build");
    if (rws.size() == 1) r.symmetric_basis(); else r.young_basis();
    r.tensor_generators();
    r.d = int(r.W.size());
    r.Wflat.resize(std::size_t(r.d) * r.dimT);
    for (int l = 0; l < r.d; ++l) std::copy(r.W[l].begin(), r.W[l].end(), r.Wflat.begin() + std::size_t(l) * r.dimT);
    r.nality = r.nbox % N;
    r.real = real_rep;
    return r;
  }

  // ---- D^(R)(U) as a matrix ----
  DM rep_matrix(const Mat<N>& U) const {
    if (kind == RepKind::Fund) return to_dm<N>(U);
    if (kind == RepKind::Adj) {  // rep_adjoint.hpp:27-40
      const int n = ngen<N>();
      const auto& g = gens<N>();
      const Mat<N> Ud = U.dag();
      DM M(n, n);
      for (int a = 0; a < n; ++a) {
        const Mat<N> TaU = g[a] * U;
        for (int b = 0; b < n; ++b) {
          const Mat<N> t = TaU * g[b] * Ud;
          M(a, b) = cd(2.0 * t.tr().real(), 0.0);
        }
      }
      return M;
    }
    DM D(d, d);  // rep_general.hpp:122-131
    for (int l = 0; l < d; ++l) {
      std::vector<cd> v = W[l];
      for (int s = 0; s < nbox; ++s) v = onebody(U, v, s);
      for (int k = 0; k < d; ++k) D(k, l) = inner(W[k], v);
    }
    return D;
  }
  // ---- hot-path applies ----
  DV rotate(const Mat<N>& U, const DV& phi) const {
    if (kind == RepKind::Tensor) return apply_tensor(U, phi, false);
    return matvec(rep_matrix(U), phi);  // representation.hpp:28
  }
  DV rotate_dag(const Mat<N>& U, const DV& phi) const {
    if (kind == RepKind::Tensor) return apply_tensor(U, phi, true);
    const DM M = rep_matrix(U);  // representation.hpp:29-32
    return matvec(M.dag(), phi);
  }
  // g^a = -2 Im phi_x^dag T^a_R D(U) phi_y — representation.hpp:36-46, rep_general.hpp:151-159
  Alg<N> hop_g(const Mat<N>& U, const DV& px, const DV& py) const {
    const DV Dy = (kind == RepKind::Tensor) ? apply_tensor(U, py, false) : matvec(rep_matrix(U), py);
    Alg<N> g{};
    for (int a = 0; a < ngen<N>(); ++a) {
      const DV t = matvec(T[a], Dy);
      const cd z = vdot(px, t);
      g[a] = -2.0 * z.imag();
    }
    return g;
  }

  // ---- tensor engine internals (rep_general.hpp:184-411) ----
  static cd inner(const std::vector<cd>& a, const std::vector<cd>& b) {
    cd s(0, 0);
    for (std::size_t i = 0; i < a.size(); ++i) s += std::conj(a[i]) * b[i];
    return s;
  }
  void onebody_into(const Mat<N>& M, const cd* src, cd* dst, int s) const {  // :184-194
    std::fill(dst, dst + dimT, cd(0, 0));
    const long blk = powN[nbox - 1 - s];
    for (long idx = 0; idx < dimT; ++idx) {
      const cd c = src[idx];
      if (c == cd(0, 0)) continue;
      const int digit = int((idx / blk) % N);
      const long base = idx - long(digit) * blk;
      for (int j = 0; j < N; ++j) dst[base + long(j) * blk] += M(j, digit) * c;
    }
  }
  std::vector<cd> onebody(const Mat<N>& M, const std::vector<cd>& v, int s) const {  // :241-252
    std::vector<cd> out(dimT, cd(0, 0));
    onebody_into(M, v.data(), out.data(), s);
    return out;
  }
  DV apply_tensor(const Mat<N>& U, const DV& phi, bool dg) const {  // :196-220
    std::vector<cd> sa(dimT, cd(0, 0)), sb(dimT, cd(0, 0));
    for (int l = 0; l < d; ++l) {
      const cd c = phi[l];
      if (c == cd(0, 0)) continue;
      const cd* w = &Wflat[std::size_t(l) * dimT];
      for (long i = 0; i < dimT; ++i) sa[i] += c * w[i];
    }
    const Mat<N> M = dg ? U.dag() : U;
    cd* cur = sa.data();
    cd* oth = sb.data();
    for (int s = 0; s < nbox; ++s) { onebody_into(M, cur, oth, s); std::swap(cur, oth); }
    DV out(d, cd(0, 0));
    for (int k = 0; k < d; ++k) {
      const cd* w = &Wflat[std::size_t(k) * dimT];
      cd acc(0, 0);
      for (long i = 0; i < dimT; ++i) acc += std::conj(w[i]) * cur[i];
      out[k] = acc;
    }
    return out;
  }
  long tup2idx(const std::vector<int>& t) const { long idx = 0; for (int j = 0; j < nbox; ++j) idx += long(t[j]) * powN[nbox - 1 - j]; return idx; }
  std::vector<int> idx2tup(long idx) const { std::vector<int> t(nbox); for (int j = 0; j < nbox; ++j) t[j] = int((idx / powN[nbox - 1 - j]) % N); return t; }
  std::vector<int> permute(const std::vector<int>& sg, const std::vector<int>& t) const {  // :266-270
    std::vector<int> u(nbox);
    for (int j = 0; j < nbox; ++j) u[sg[j]] = t[j];
    return u;
  }
  static int parity_sign(const std::vector<int>& p) {  // :271-282
    std::vector<char> seen(p.size(), 0);
    int sign = 1;
    for (std::size_t i = 0; i < p.size(); ++i) {
      if (seen[i]) continue;
      int len = 0; std::size_t j = i;
      while (!seen[j]) { seen[j] = 1; j = std::size_t(p[j]); ++len; }
      if (len % 2 == 0) sign = -sign;
    }
    return sign;
  }
  // Product of symmetric groups on disjoint blocks, enumerated in the reference order (:286-308).
  void block_group(const std::vector<std::vector<int>>& blocks, std::vector<std::vector<int>>& perms,
                   std::vector<int>& signs) const {
    std::vector<int> id(nbox);
    std::iota(id.begin(), id.end(), 0);
    perms = {id};
    for (const auto& blk : blocks) {
      std::vector<int> src = blk;
      std::sort(src.begin(), src.end());
      std::vector<std::vector<int>> orders;
      std::vector<int> cur = src;
      do { orders.push_back(cur); } while (std::next_permutation(cur.begin(), cur.end()));
      std::vector<std::vector<int>> next;
      next.reserve(perms.size() * orders.size());
      for (const auto& g : perms)
        for (const auto& o : orders) {
          std::vector<int> g2 = g;
          for (std::size_t k = 0; k < src.size(); ++k) g2[src[k]] = o[k];
          next.push_back(std::move(g2));
        }
      perms.swap(next);
    }
    signs.resize(perms.size());
    for (std::size_t i = 0; i < perms.size(); ++i) signs[i] = parity_sign(perms[i]);
  }
  std::vector<cd> symmetrizer(const std::vector<int>& t, const std::vector<std::vector<int>>& R,
                              const std::vector<std::vector<int>>& C, const std::vector<int>& Cs) const {  // :311-328
    std::vector<cd> v1(dimT, cd(0, 0));
    for (const auto& p : R) v1[tup2idx(permute(p, t))] += cd(1, 0);
    std::vector<cd> res(dimT, cd(0, 0));
    for (long idx = 0; idx < dimT; ++idx) {
      const cd c = v1[idx];
      if (c == cd(0, 0)) continue;
      const std::vector<int> ta = idx2tup(idx);
      for (std::size_t q = 0; q < C.size(); ++q) res[tup2idx(permute(C[q], ta))] += cd(Cs[q], 0) * c;
    }
    return res;
  }
  void symmetric_basis() {  // :338-362
    W.clear();
    std::vector<int> t(nbox, 0);
    while (true) {
      std::vector<cd> v(dimT, cd(0, 0));
      std::vector<int> perm = t;
      long count = 0;
      do { v[tup2idx(perm)] = cd(1, 0); ++count; } while (std::next_permutation(perm.begin(), perm.end()));
      const double inv = 1.0 / std::sqrt(double(count));
      for (auto& z : v) z *= cd(inv, 0);
      W.push_back(std::move(v));
      int j = nbox - 1;
      while (j >= 0 && t[j] == N - 1) --j;
      if (j < 0) break;
      ++t[j];
      for (int k = j + 1; k < nbox; ++k) t[k] = t[j];
    }
  }
  void young_basis() {  // :364-398
    std::vector<std::vector<int>> rowblk(rows.size());
    std::vector<std::vector<int>> colblk(rows[0]);
    int pos = 0;
    for (std::size_t i = 0; i < rows.size(); ++i)
      for (int j = 0; j < rows[i]; ++j) { rowblk[i].push_back(pos); colblk[j].push_back(pos); ++pos; }
    std::vector<std::vector<int>> R, C;
    std::vector<int> Rs, Cs;
    block_group(rowblk, R, Rs);
    block_group(colblk, C, Cs);
    W.clear();
    for (long tc = 0; tc < dimT; ++tc) {
      std::vector<cd> v = symmetrizer(idx2tup(tc), R, C, Cs);
      for (const auto& w : W) {
        const cd ov = inner(w, v);
        for (long i = 0; i < dimT; ++i) v[i] -= ov * w[i];
      }
      const double nrm = std::sqrt(std::max(0.0, inner(v, v).real()));
      if (nrm > 1e-9) {
        const cd inv(1.0 / nrm, 0);
        for (auto& z : v) z *= inv;
        W.push_back(std::move(v));
      }
    }
  }
  void tensor_generators() {  // :400-411 (derivation action restricted to the irrep)
    const int dd = int(W.size());
    T.assign(ngen<N>(), DM(dd, dd));
    for (int a = 0; a < ngen<N>(); ++a) {
      DM Ta(dd, dd);
      for (int l = 0; l < dd; ++l) {
        std::vector<cd> v(dimT, cd(0, 0));
        for (int s = 0; s < nbox; ++s) {  // derivation(a, W[l]), :255-263
          const std::vector<cd> tmp = onebody(gens<N>()[a], W[l], s);
          for (long i = 0; i < dimT; ++i) v[i] += tmp[i];
        }
        for (int k = 0; k < dd; ++k) Ta(k, l) = inner(W[k], v);
      }
      T[a] = Ta;
    }
  }
};

// Parse a rep spec the way hmc_higgs.cpp:23-34 does: fund | adj | r1,r2[:real]
template <int N>
Rep<N> make_rep(std::string spec) {
  if (spec == "fund") return Rep<N>::fund();
  if (spec == "adj") return Rep<N>::adj();
  bool real = false;
  const auto colon = spec.find(':');
  if (colon != std::string::npos) { real = (spec.substr(colon + 1) == "real"); spec = spec.substr(0, colon); }
  std::vector<int> rws;
  std::size_t p = 0;
  while (p <= spec.size()) {
    const std::size_t q = spec.find(',', p);
    const std::string tok = spec.substr(p, q == std::string::npos ? std::string::npos : q - p);
    if (!tok.empty()) rws.push_back(std::atoi(tok.c_str()));
    if (q == std::string::npos) break;
    p = q + 1;
  }
  if (rws.empty()) throw std::runtime_error("bad rep spec '" + spec + "'");
  return Rep<N>::young(rws, real);
}

// ---------------------------------------------------------------------------
// HMC state and sweeps — fields.hpp, scalar_field.hpp, gauge_wilson.hpp,
// scalar_higgs.hpp, hmc.hpp, gauge_higgs_hmc.hpp, observables.hpp
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Multi-invariant quartic potential — scalar_invariants.hpp:20-208.
// V = -mu2 phi^dag phi + sum_c f_c ||P_c(M)||^2, M = phi phi^dag, P_c the
// eigenprojector of Chat(M) = sum_a [T^a,[T^a,M]] on channel c (a Lagrange
// polynomial in Chat); dV/dphi^* = -mu2 phi + sum_c 2 f_c P_c(M) phi.
// The channel eigenvalues are found here by diagonalising Chat on the real
// space of Hermitian d x d matrices with cyclic Jacobi (the reference runs
// Lanczos from a random Hermitian start; both return the distinct C2's).
// ---------------------------------------------------------------------------
inline double re_tr_dag(const DM& A, const DM& B) {  // linalg.hpp:252-256
  double s = 0.0;
  for (std::size_t i = 0; i < A.a.size(); ++i) s += (std::conj(A.a[i]) * B.a[i]).real();
  return s;
}
inline std::vector<double> jacobi_eigenvalues(std::vector<double> A, int n) {
  auto at = [&](int i, int j) -> double& { return A[std::size_t(i) * n + j]; };
  for (int sweep = 0; sweep < 200; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < n; ++p) for (int q = p + 1; q < n; ++q) off += at(p, q) * at(p, q);
    if (off < 1e-30) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        if (std::fabs(at(p, q)) < 1e-300) continue;
        const double th = (at(q, q) - at(p, p)) / (2.0 * at(p, q));
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
        for (int k = 0; k < n; ++k) { const double a = at(k, p), b = at(k, q); at(k, p) = c * a - sn * b; at(k, q) = sn * a + c * b; }
        for (int k = 0; k < n; ++k) { const double a = at(p, k), b = at(q, k); at(p, k) = c * a - sn * b; at(q, k) = sn * a + c * b; }
      }
  }
  std::vector<double> ev(n);
  for (int i = 0; i < n; ++i) ev[i] = at(i, i);
  return ev;
}

template <int N>
struct Channels {
  const Rep<N>* rep;
  int d;
  std::vector<double> lambda;
  explicit Channels(const Rep<N>& R) : rep(&R), d(R.d) {
    // orthonormal basis of Hermitian d x d matrices (real inner product Re Tr(A^dag B))
    std::vector<DM> B;
    for (int i = 0; i < d; ++i) { DM e(d, d); e(i, i) = 1.0; B.push_back(e); }
    const double h = 1.0 / std::sqrt(2.0);
    for (int i = 0; i < d; ++i)
      for (int j = i + 1; j < d; ++j) {
        DM e(d, d); e(i, j) = h; e(j, i) = h; B.push_back(e);
        DM f(d, d); f(i, j) = cd(0, -h); f(j, i) = cd(0, h); B.push_back(f);
      }
    const int n = int(B.size());
    std::vector<double> C(std::size_t(n) * n);
    for (int q = 0; q < n; ++q) {
      const DM CB = casimir(B[q]);
      for (int p = 0; p < n; ++p) C[std::size_t(p) * n + q] = re_tr_dag(B[p], CB);
    }
    std::vector<double> ev = jacobi_eigenvalues(C, n);
    std::sort(ev.begin(), ev.end());
    for (double e : ev) if (lambda.empty() || std::fabs(e - lambda.back()) > 1e-4) lambda.push_back(e);  // as scalar_invariants.hpp:175-176
  }
  int n_channels() const { return int(lambda.size()); }
  DM casimir(const DM& M) const {  // scalar_invariants.hpp:31-39
    DM out(d, d);
    for (const DM& T : rep->T) {
      const DM c1 = T * M - M * T;
      out = out + (T * c1 - c1 * T);
    }
    return out;
  }
  DM proj(int c, const DM& M) const {  // scalar_invariants.hpp:42-50
    DM X = M;
    for (int cp = 0; cp < n_channels(); ++cp) {
      if (cp == c) continue;
      const DM CX = casimir(X);
      DM Y(d, d);
      for (std::size_t i = 0; i < Y.a.size(); ++i) Y.a[i] = (CX.a[i] - X.a[i] * lambda[cp]) * (1.0 / (lambda[c] - lambda[cp]));
      X = Y;
    }
    return X;
  }
  static DM outer(const DV& x) {
    DM M(int(x.size()), int(x.size()));
    for (std::size_t a = 0; a < x.size(); ++a)
      for (std::size_t b = 0; b < x.size(); ++b) M(int(a), int(b)) = x[a] * std::conj(x[b]);
    return M;
  }
  double value(const DV& x, const std::vector<double>& f, double mu2) const {  // scalar_invariants.hpp:62-67
    double v = -mu2 * vnorm2(x);
    const DM M = outer(x);
    for (int c = 0; c < n_channels(); ++c) if (f[c] != 0.0) v += f[c] * re_tr_dag(M, proj(c, M));
    return v;
  }
  DV grad(const DV& x, const std::vector<double>& f, double mu2) const {  // scalar_invariants.hpp:70-87
    const DM M = outer(x);
    DV g(x.size());
    for (std::size_t e = 0; e < x.size(); ++e) g[e] = cd(-mu2, 0) * x[e];
    for (int c = 0; c < n_channels(); ++c) {
      if (f[c] == 0.0) continue;
      const DV PMx = matvec(proj(c, M), x);
      for (std::size_t e = 0; e < x.size(); ++e) g[e] += cd(2.0 * f[c], 0) * PMx[e];
    }
    return g;
  }
};

template <int N>
struct Engine {
  Geom g;
  int D;
  const Rep<N>* rep = nullptr;  // null => pure-gauge GaugeHMC semantics (hmc.hpp)
  int d = 0;
  std::uint64_t seed;
  std::vector<Mat<N>> U;        // [s*D+mu]
  std::vector<Alg<N>> P, F;     // [s*D+mu]
  std::vector<cd> phi, pi, Fphi;  // [s*d+k]
  double beta = 2.3, kappa = 0.1, lambda = 0.5, tau = 1.0;
  int nmd = 20;
  bool omelyan = true, reunit_each_traj = true;
  bool frozen = false;  // |phi_x| = 1 scalar (gauge_higgs_hmc.hpp:38)
  std::uint64_t traj_count = 0, accept_count = 0;
  double last_dH = 0.0;

  Engine(int dim, const int* L, const Rep<N>* r, std::uint64_t sd)
      : g(dim, L), D(dim), rep(r), d(r ? r->d : 0), seed(sd),
        U(g.vol * dim), P(g.vol * dim), F(g.vol * dim),
        phi(g.vol * (r ? r->d : 0)), pi(phi.size()), Fphi(phi.size()) {}

  DV site_phi(std::int64_t s) const { return DV(phi.begin() + s * d, phi.begin() + (s + 1) * d); }

  // fields.hpp:25-34
  void hot(double sigma) {
#pragma omp parallel for schedule(static)
    for (std::int64_t s = 0; s < g.vol; ++s)
      for (int mu = 0; mu < D; ++mu) {
        Alg<N> v{};
        for (int a = 0; a < ngen<N>(); ++a) v[a] = sigma * draw_gauss(seed, stream_key(1, s, mu, a));
        U[s * D + mu] = expi<N>(alg_to_mat<N>(v));
      }
  }
  void cold() { for (auto& m : U) m = Mat<N>::eye(); }
  // scalar_field.hpp:40-51
  void gaussian_field(std::vector<cd>& f, std::uint64_t stream, double sigma) const {
#pragma omp parallel for schedule(static)
    for (std::int64_t s = 0; s < g.vol; ++s)
      for (int k = 0; k < d; ++k) {
        const double re = sigma * draw_gauss(seed, stream_key(stream, s, 2 * k));
        const double im = rep->real ? 0.0 : sigma * draw_gauss(seed, stream_key(stream, s, 2 * k + 1));
        f[s * d + k] = cd(re, im);
      }
  }
  void phi_cold(double c) {  // scalar_field.hpp:33-37
    for (auto& z : phi) z = cd(0, 0);
    for (std::int64_t s = 0; s < g.vol; ++s) phi[s * d] = cd(c, 0);
  }
  // fields.hpp:59-66
  void refresh_P(std::uint64_t stream) {
#pragma omp parallel for schedule(static)
    for (std::int64_t s = 0; s < g.vol; ++s)
      for (int mu = 0; mu < D; ++mu)
        for (int a = 0; a < ngen<N>(); ++a) P[s * D + mu][a] = draw_gauss(seed, stream_key(stream, s, mu, a));
  }
  void refresh_momenta() {  // gauge_higgs_hmc.hpp:167-173 / hmc.hpp:74-75
    if (!rep) { refresh_P(stream_key(0xA1, traj_count)); return; }
    refresh_P(stream_key(0xC3, traj_count));
    gaussian_field(pi, stream_key(0xD4, traj_count), 1.0);
    if (frozen) project_pi_tangent();
  }
  double kinetic_P() const {  // fields.hpp:69-74
    double e = 0.0;
    for (const auto& v : P) for (int a = 0; a < ngen<N>(); ++a) e += v[a] * v[a];
    return 0.5 * e;
  }
  double kinetic_pi() const {  // scalar_field.hpp:53-57
    double e = 0.0;
    for (const auto& z : pi) e += std::norm(z);
    return 0.5 * e;
  }

  // gauge_wilson.hpp:12-27
  Mat<N> staple(std::int64_t s, int mu) const {
    Mat<N> sig;
    const std::int64_t spm = g.up(s, mu);
    for (int nu = 0; nu < D; ++nu) {
      if (nu == mu) continue;
      const std::int64_t spn = g.up(s, nu), smn = g.dn(s, nu), spmmn = g.dn(spm, nu);
      sig += U[spm * D + nu] * U[spn * D + mu].dag() * U[s * D + nu].dag();
      sig += U[spmmn * D + nu].dag() * U[smn * D + mu].dag() * U[smn * D + nu];
    }
    return sig;
  }
  double sum_retr_plaq() const {  // gauge_wilson.hpp:32-47 (serial order)
    double sum = 0.0;
    for (std::int64_t s = 0; s < g.vol; ++s)
      for (int mu = 0; mu < D; ++mu) {
        const std::int64_t spm = g.up(s, mu);
        for (int nu = mu + 1; nu < D; ++nu) {
          const std::int64_t spn = g.up(s, nu);
          Mat<N> Pl = U[s * D + mu] * U[spm * D + nu] * U[spn * D + mu].dag() * U[s * D + nu].dag();
          sum += Pl.tr().real();
        }
      }
    return sum;
  }
  double gauge_action() const { const double sr = sum_retr_plaq(); return (beta / N) * (double(N) * g.n_plaq() - sr); }
  double plaquette() const { const double sr = sum_retr_plaq(); return sr / (double(N) * g.n_plaq()); }
  void add_gauge_force() {  // gauge_wilson.hpp:71-83
    const double c = beta / N;
#pragma omp parallel for schedule(static)
    for (std::int64_t s = 0; s < g.vol; ++s)
      for (int mu = 0; mu < D; ++mu) {
        const Mat<N> Om = U[s * D + mu] * staple(s, mu);
        Alg<N>& Fa = F[s * D + mu];
        for (int a = 0; a < ngen<N>(); ++a) Fa[a] += c * tr_prod(gens<N>()[a], Om).imag();
      }
  }
  // scalar_higgs.hpp:25-39 (QuarticPotential)
  // on-site potential: the quartic default (scalar_higgs.hpp:25-39) or the plugged-in
  // multi-invariant one (GaugeHiggsHMC::potential, gauge_higgs_hmc.hpp:27)
  const Channels<N>* multi = nullptr;
  std::vector<double> multi_f;
  double multi_mu2 = 0.0;
  double pot_value(const DV& x) const {
    if (multi) return multi->value(x, multi_f, multi_mu2);
    const double n2 = vnorm2(x);
    return n2 + lambda * (n2 - 1.0) * (n2 - 1.0);
  }
  DV pot_grad(const DV& x) const {
    if (multi) return multi->grad(x, multi_f, multi_mu2);
    const double n2 = vnorm2(x);
    const double c = 1.0 + 2.0 * lambda * (n2 - 1.0);
    DV gr(x.size());
    for (std::size_t k = 0; k < x.size(); ++k) gr[k] = cd(c, 0) * x[k];
    return gr;
  }
  double scalar_action() const {  // scalar_higgs.hpp:42-58
    double S = 0.0;
    for (std::int64_t s = 0; s < g.vol; ++s) {
      const DV px = site_phi(s);
      S += pot_value(px);
      for (int mu = 0; mu < D; ++mu) {
        const DV Dpy = rep->rotate(U[s * D + mu], site_phi(g.up(s, mu)));
        S -= 2.0 * kappa * vdot(px, Dpy).real();
      }
    }
    return S;
  }
  void scalar_force() {  // scalar_higgs.hpp:67-99
#pragma omp parallel for schedule(static)
    for (std::int64_t s = 0; s < g.vol; ++s) {
      const DV px = site_phi(s);
      const DV gr = pot_grad(px);
      DV Fs(d);
      for (int k = 0; k < d; ++k) Fs[k] = -2.0 * gr[k];
      for (int mu = 0; mu < D; ++mu) {
        const std::int64_t yf = g.up(s, mu), yb = g.dn(s, mu);
        const DV Dpf = rep->rotate(U[s * D + mu], site_phi(yf));
        const DV Dpb = rep->rotate_dag(U[yb * D + mu], site_phi(yb));
        for (int k = 0; k < d; ++k) Fs[k] += 2.0 * kappa * (Dpf[k] + Dpb[k]);
      }
      for (int k = 0; k < d; ++k) Fphi[s * d + k] = Fs[k];
    }
  }
  void add_matter_link_force() {  // scalar_higgs.hpp:110-128
#pragma omp parallel for schedule(static)
    for (std::int64_t s = 0; s < g.vol; ++s)
      for (int mu = 0; mu < D; ++mu) {
        const Alg<N> gg = rep->hop_g(U[s * D + mu], site_phi(s), site_phi(g.up(s, mu)));
        Alg<N>& Fa = F[s * D + mu];
        for (int a = 0; a < ngen<N>(); ++a) Fa[a] += -kappa * gg[a];
      }
  }
  double link_energy() const {  // scalar_higgs.hpp:130-144
    double sm = 0.0;
    for (std::int64_t x = 0; x < g.vol; ++x)
      for (int mu = 0; mu < D; ++mu) {
        const DV Dpy = rep->rotate(U[x * D + mu], site_phi(g.up(x, mu)));
        sm += vdot(site_phi(x), Dpy).real();
      }
    return sm / (double(g.vol) * D);
  }
  double higgs_length() const {  // scalar_higgs.hpp:146-154
    double sm = 0.0;
    for (std::int64_t x = 0; x < g.vol; ++x) sm += vnorm2(site_phi(x));
    return sm / double(g.vol);
  }
  double polyakov(int tdir) const {  // observables.hpp:18-34
    const int Lt = g.L[tdir];
    double sum = 0.0;
    std::int64_t count = 0;
    for (std::int64_t s = 0; s < g.vol; ++s) {
      if (g.coord(s, tdir) != 0) continue;
      Mat<N> Pl = U[s * D + tdir];
      std::int64_t cur = g.up(s, tdir);
      for (int t = 1; t < Lt; ++t) { Pl = Pl * U[cur * D + tdir]; cur = g.up(cur, tdir); }
      sum += Pl.tr().real();
      ++count;
    }
    return sum / (double(count) * N);
  }
  // observables.hpp:36-48: path-ordered product around the Ra x Rb rectangle
  // based at x in the (mu, nu) plane (Ra steps along mu, Rb along nu, back).
  Mat<N> wilson_rect(std::int64_t x, int mu, int nu, int Ra, int Rb) const {
    Mat<N> W = Mat<N>::eye();
    std::int64_t s = x;
    for (int i = 0; i < Ra; ++i) { W = W * U[s * D + mu]; s = g.up(s, mu); }
    for (int j = 0; j < Rb; ++j) { W = W * U[s * D + nu]; s = g.up(s, nu); }
    for (int i = 0; i < Ra; ++i) { s = g.dn(s, mu); W = W * U[s * D + mu].dag(); }
    for (int j = 0; j < Rb; ++j) { s = g.dn(s, nu); W = W * U[s * D + nu].dag(); }
    return W;
  }
  // Re Tr D^(R)(M) / d_R; probe null = fundamental (observables.hpp:50-81).
  static double probe_trace(const Rep<N>* probe, const Mat<N>& M) {
    if (!probe) return M.tr().real();
    return probe->rep_matrix(M).tr().real();
  }
  // observables.hpp:50-81: <(1/d_R) Re Tr D^(R)(W(Ra,Rb))> over sites and planes mu < nu.
  double wilson_loop(const Rep<N>* probe, int Ra, int Rb) const {
    double sum = 0.0;
    std::int64_t count = 0;
    for (std::int64_t x = 0; x < g.vol; ++x)
      for (int mu = 0; mu < D; ++mu)
        for (int nu = mu + 1; nu < D; ++nu) {
          sum += probe_trace(probe, wilson_rect(x, mu, nu, Ra, Rb));
          ++count;
        }
    return sum / (double(count) * (probe ? probe->d : N));
  }
  // observables.hpp:83-101: Polyakov loop in direction tdir traced in the probe rep.
  double polyakov_rep(const Rep<N>* probe, int tdir) const {
    const int Lt = g.L[tdir];
    double sum = 0.0;
    std::int64_t count = 0;
    for (std::int64_t s = 0; s < g.vol; ++s) {
      if (g.coord(s, tdir) != 0) continue;
      Mat<N> Pl = U[s * D + tdir];
      std::int64_t cur = g.up(s, tdir);
      for (int t = 1; t < Lt; ++t) { Pl = Pl * U[cur * D + tdir]; cur = g.up(cur, tdir); }
      sum += probe_trace(probe, Pl);
      ++count;
    }
    return sum / (double(count) * (probe ? probe->d : N));
  }
  void reunitarize_all() { for (auto& m : U) reunitarize<N>(m); }
  double max_unitarity_violation() const { double w = 0.0; for (const auto& m : U) w = std::max(w, unitarity_violation<N>(m)); return w; }

  // Link force F and scalar force Fphi for the current fields (kick without the update).
  void forces() {
    for (auto& v : F) v.fill(0.0);
    add_gauge_force();
    if (rep) { add_matter_link_force(); scalar_force(); }
  }
  double hamiltonian() const {  // hmc.hpp:32, gauge_higgs_hmc.hpp:105-110
    if (!rep) return kinetic_P() + gauge_action();
    return kinetic_P() + kinetic_pi() + gauge_action() + scalar_action();
  }
  // Frozen-length scalar |phi_x| = 1 — gauge_higgs_hmc.hpp:61-103.
  // Tangent projection: pi -= Re(phi^dag pi) phi per site.
  void project_pi_tangent() {
#pragma omp parallel for schedule(static)
    for (std::int64_t s = 0; s < g.vol; ++s) {
      double rp = 0.0;
      for (int k = 0; k < d; ++k) rp += (std::conj(phi[s * d + k]) * pi[s * d + k]).real();
      for (int k = 0; k < d; ++k) pi[s * d + k] -= cd(rp, 0.0) * phi[s * d + k];
    }
  }
  void normalize_phi() {
#pragma omp parallel for schedule(static)
    for (std::int64_t s = 0; s < g.vol; ++s) {
      double n2 = 0.0;
      for (int k = 0; k < d; ++k) n2 += std::norm(phi[s * d + k]);
      if (n2 > 0.0) {
        const double inv = 1.0 / std::sqrt(n2);
        for (int k = 0; k < d; ++k) phi[s * d + k] *= cd(inv, 0.0);
      }
    }
  }
  // Great-circle flow by eps: phi' = cos(r eps) phi + sin(r eps)/r pi, pi' = -r sin(r eps) phi + cos(r eps) pi.
  void geodesic_drift(double eps) {
#pragma omp parallel for schedule(static)
    for (std::int64_t s = 0; s < g.vol; ++s) {
      double r2 = 0.0;
      for (int k = 0; k < d; ++k) r2 += std::norm(pi[s * d + k]);
      const double r = std::sqrt(r2);
      if (r < 1e-300) continue;
      const double c = std::cos(r * eps), sn = std::sin(r * eps);
      for (int k = 0; k < d; ++k) {
        const cd f = phi[s * d + k], p = pi[s * d + k];
        phi[s * d + k] = cd(c, 0.0) * f + cd(sn / r, 0.0) * p;
        pi[s * d + k] = cd(-r * sn, 0.0) * f + cd(c, 0.0) * p;
      }
    }
  }
  void kick(double eps) {  // gauge_higgs_hmc.hpp:112-134, hmc.hpp:35-42
    for (auto& v : F) v.fill(0.0);
    add_gauge_force();
    if (rep) add_matter_link_force();
#pragma omp parallel for schedule(static)
    for (std::size_t i = 0; i < P.size(); ++i)
      for (int a = 0; a < ngen<N>(); ++a) P[i][a] -= eps * F[i][a];
    if (!rep) return;
    scalar_force();
#pragma omp parallel for schedule(static)
    for (std::size_t i = 0; i < pi.size(); ++i) pi[i] += eps * Fphi[i];
    if (frozen) project_pi_tangent();
  }
  void drift(double eps) {  // gauge_higgs_hmc.hpp:136-147, hmc.hpp:44-52
#pragma omp parallel for schedule(static)
    for (std::int64_t s = 0; s < g.vol; ++s)
      for (int mu = 0; mu < D; ++mu) {
        Mat<N> Pm = alg_to_mat<N>(P[s * D + mu]);
        Pm *= cd(eps, 0.0);
        U[s * D + mu] = expi<N>(Pm) * U[s * D + mu];
      }
    if (!rep) return;
    if (frozen) { geodesic_drift(eps); return; }
#pragma omp parallel for schedule(static)
    for (std::size_t i = 0; i < phi.size(); ++i) phi[i] += eps * pi[i];
  }
  void md_evolve() {  // gauge_higgs_hmc.hpp:149-165, hmc.hpp:54-70
    const double eps = tau / nmd;
    if (!omelyan) {
      kick(0.5 * eps);
      for (int i = 0; i < nmd; ++i) { drift(eps); if (i != nmd - 1) kick(eps); }
      kick(0.5 * eps);
    } else {
      const double lam = kOmelyanLambda;
      kick(lam * eps);
      for (int i = 0; i < nmd; ++i) {
        drift(0.5 * eps);
        kick((1.0 - 2.0 * lam) * eps);
        drift(0.5 * eps);
        kick((i != nmd - 1 ? 2.0 * lam : lam) * eps);
      }
    }
  }
  // gauge_higgs_hmc.hpp:175-192, hmc.hpp:73-91. Returns accept; H parts of both ends out.
  bool trajectory(double* Hi_out = nullptr, double* Hf_out = nullptr) {
    refresh_momenta();
    const std::vector<Mat<N>> Usave = U;
    const std::vector<cd> phisave = phi;
    const double Hi = hamiltonian();
    md_evolve();
    if (reunit_each_traj) reunitarize_all();
    const double Hf = hamiltonian();
    last_dH = Hf - Hi;
    ++traj_count;
    const double r = draw_uniform(seed, stream_key(rep ? 0xE5 : 0xB2, traj_count));
    const bool acc = (last_dH <= 0.0) || (r < std::exp(-last_dH));
    if (acc) ++accept_count;
    else { U = Usave; phi = phisave; }
    if (Hi_out) *Hi_out = Hi;
    if (Hf_out) *Hf_out = Hf;
    return acc;
  }
};

}  // namespace orc
