#include "rep_build.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <map>
#include <numeric>

namespace ghx {

using cplx = std::complex<double>;

std::vector<cplx> fundamental_generators(int N) {
  const int ng = N * N - 1;
  std::vector<cplx> T(std::size_t(ng) * N * N, cplx(0, 0));
  auto at = [&](int a, int i, int j) -> cplx& { return T[(std::size_t(a) * N + i) * N + j]; };
  int a = 0;
  for (int j = 0; j < N; ++j)
    for (int k = j + 1; k < N; ++k, ++a) { at(a, j, k) = 0.5; at(a, k, j) = 0.5; }
  for (int j = 0; j < N; ++j)
    for (int k = j + 1; k < N; ++k, ++a) { at(a, j, k) = cplx(0, -0.5); at(a, k, j) = cplx(0, 0.5); }
  for (int l = 1; l < N; ++l, ++a) {
    const double c = std::sqrt(2.0 / (double(l) * (l + 1)));
    for (int j = 0; j < l; ++j) at(a, j, j) = 0.5 * c;
    at(a, l, l) = -0.5 * c * l;
  }
  return T;
}

namespace {

// Tensor index helpers: slot 0 is the most significant base-N digit.
struct TensorSpace {
  int N, n;
  long dim;
  std::vector<long> pw;
  TensorSpace(int N_, int n_) : N(N_), n(n_), pw(n_ + 1, 1) {
    for (int k = 1; k <= n; ++k) pw[k] = pw[k - 1] * N;
    dim = pw[n];
  }
  int digit(long idx, int slot) const { return int((idx / pw[n - 1 - slot]) % N); }
  long index(const std::vector<int>& t) const {
    long v = 0;
    for (int s = 0; s < n; ++s) v = v * N + t[s];
    return v;
  }
  std::vector<int> tuple(long idx) const {
    std::vector<int> t(n);
    for (int s = n - 1; s >= 0; --s) { t[s] = int(idx % N); idx /= N; }
    return t;
  }
};

// All permutations of positions that permute within each block (product of
// symmetric groups), with signs. Enumeration order does not matter here: the
// symmetrizers only sum integer coefficients.
void block_perms(int n, const std::vector<std::vector<int>>& blocks, std::vector<std::vector<int>>& perms,
                 std::vector<int>& signs) {
  std::vector<int> id(n);
  std::iota(id.begin(), id.end(), 0);
  perms.assign(1, id);
  for (const auto& b : blocks) {
    std::vector<int> order = b;
    std::sort(order.begin(), order.end());
    std::vector<std::vector<int>> next;
    do {
      for (const auto& p : perms) {
        std::vector<int> q = p;
        for (std::size_t k = 0; k < b.size(); ++k) q[std::size_t(b[k])] = order[k];
        next.push_back(q);
      }
    } while (std::next_permutation(order.begin(), order.end()));
    perms.swap(next);
  }
  signs.clear();
  for (const auto& p : perms) {
    std::vector<char> seen(p.size(), 0);
    int sg = 1;
    for (std::size_t i = 0; i < p.size(); ++i) {
      if (seen[i]) continue;
      std::size_t j = i, len = 0;
      while (!seen[j]) { seen[j] = 1; j = std::size_t(p[j]); ++len; }
      if (len % 2 == 0) sg = -sg;
    }
    signs.push_back(sg);
  }
}

// Place tuple entry j at position sigma[j].
std::vector<int> act(const std::vector<int>& sigma, const std::vector<int>& t) {
  std::vector<int> u(t.size());
  for (std::size_t j = 0; j < t.size(); ++j) u[std::size_t(sigma[j])] = t[j];
  return u;
}

void symmetric_basis(const TensorSpace& ts, std::vector<std::vector<double>>& W) {
  std::vector<int> t(ts.n, 0);
  for (;;) {
    std::vector<double> v(ts.dim, 0.0);
    std::vector<int> p = t;
    long cnt = 0;
    do { v[std::size_t(ts.index(p))] = 1.0; ++cnt; } while (std::next_permutation(p.begin(), p.end()));
    const double inv = 1.0 / std::sqrt(double(cnt));
    for (auto& x : v) x *= inv;
    W.push_back(std::move(v));
    int j = ts.n - 1;
    while (j >= 0 && t[j] == ts.N - 1) --j;
    if (j < 0) break;
    ++t[j];
    for (int k = j + 1; k < ts.n; ++k) t[k] = t[j];
  }
}

void young_basis(const TensorSpace& ts, const std::vector<int>& rows, std::vector<std::vector<double>>& W) {
  std::vector<std::vector<int>> rb(rows.size()), cb(rows[0]);
  int pos = 0;
  for (std::size_t i = 0; i < rows.size(); ++i)
    for (int j = 0; j < rows[i]; ++j, ++pos) { rb[i].push_back(pos); cb[std::size_t(j)].push_back(pos); }
  std::vector<std::vector<int>> Rp, Cp;
  std::vector<int> Rs, Cs;
  block_perms(ts.n, rb, Rp, Rs);
  block_perms(ts.n, cb, Cp, Cs);
  for (long tc = 0; tc < ts.dim; ++tc) {
    const std::vector<int> t = ts.tuple(tc);
    std::map<long, double> a;  // row-symmetrized (sparse, exact integers)
    for (const auto& p : Rp) a[ts.index(act(p, t))] += 1.0;
    std::vector<double> v(ts.dim, 0.0);
    for (const auto& kv : a)
      for (std::size_t q = 0; q < Cp.size(); ++q) v[std::size_t(ts.index(act(Cp[q], ts.tuple(kv.first))))] += Cs[q] * kv.second;
    for (const auto& w : W) {
      double ov = 0.0;
      for (long i = 0; i < ts.dim; ++i) ov += w[std::size_t(i)] * v[std::size_t(i)];
      if (ov != 0.0)
        for (long i = 0; i < ts.dim; ++i) v[std::size_t(i)] -= ov * w[std::size_t(i)];
    }
    double nr = 0.0;
    for (double x : v) nr += x * x;
    nr = std::sqrt(std::max(0.0, nr));
    if (nr > 1e-9) {
      for (auto& x : v) x /= nr;
      W.push_back(std::move(v));
    }
  }
}

}  // namespace

std::vector<int> parse_young_rows(int N, std::string spec, bool* real) {
  bool re = false;
  const auto colon = spec.find(':');
  if (colon != std::string::npos) { re = spec.substr(colon + 1) == "real"; spec = spec.substr(0, colon); }
  if (real) *real = re;
  std::vector<int> rows;
  std::size_t p = 0;
  while (p <= spec.size()) {
    const std::size_t q = spec.find(',', p);
    const std::string tok = spec.substr(p, q == std::string::npos ? std::string::npos : q - p);
    if (!tok.empty()) rows.push_back(std::atoi(tok.c_str()));
    if (q == std::string::npos) break;
    p = q + 1;
  }
  if (rows.empty()) throw std::runtime_error("bad rep spec '" + spec + "'");
  while (!rows.empty() && rows.back() == 0) rows.pop_back();
  if (rows.empty()) throw std::runtime_error("GeneralRep: empty diagram (singlet not supported)");
  for (std::size_t i = 0; i + 1 < rows.size(); ++i)
    if (rows[i] < rows[i + 1]) throw std::runtime_error("GeneralRep: rows must be non-increasing");
  if (int(rows.size()) > N - 1) throw std::runtime_error("GeneralRep: > N-1 rows (strip full columns of N)");
  const int n = std::accumulate(rows.begin(), rows.end(), 0);
  double dimT = 1.0;
  for (int k = 0; k < n; ++k) dimT *= N;
  if (dimT > 200000.0) throw std::runtime_error("GeneralRep: N^n exceeds cap; irrep too large for tensor build");
  return rows;
}

long young_dimension(int N, const std::vector<int>& rows) {
  // hook-content formula: prod over boxes (N + column - row) / hook length
  double num = 1.0, den = 1.0;
  for (std::size_t i = 0; i < rows.size(); ++i)
    for (int j = 0; j < rows[i]; ++j) {
      int below = 0;
      for (std::size_t k = i + 1; k < rows.size(); ++k)
        if (rows[k] > j) ++below;
      num *= double(N + j - int(i));
      den *= double(rows[i] - j - 1 + below + 1);
    }
  return long(std::llround(num / den));
}

RepInfo build_rep(int N, const std::string& spec_in) {
  RepInfo r;
  r.N = N;
  const int ng = N * N - 1;
  std::string spec = spec_in;
  if (spec.empty() || spec == "none") { r.kind = 0; r.name = "none"; return r; }
  if (spec == "fund") {
    r.kind = 1; r.d = N; r.nality = 1; r.name = "fundamental";
    r.T = fundamental_generators(N);
    return r;
  }
  if (spec == "adj") {
    r.kind = 2; r.d = ng; r.nality = 0; r.real = true; r.name = "adjoint";
    const auto F = fundamental_generators(N);
    auto f = [&](int a, int b, int c) {  // f^{abc} = 2 Im Tr([Ta,Tb]Tc)
      cplx t(0, 0);
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
          for (int k = 0; k < N; ++k) {
            const cplx ab = F[(std::size_t(a) * N + i) * N + j] * F[(std::size_t(b) * N + j) * N + k] -
                            F[(std::size_t(b) * N + i) * N + j] * F[(std::size_t(a) * N + j) * N + k];
            t += ab * F[(std::size_t(c) * N + k) * N + i];
          }
      return 2.0 * t.imag();
    };
    r.T.assign(std::size_t(ng) * ng * ng, cplx(0, 0));
    for (int a = 0; a < ng; ++a)
      for (int b = 0; b < ng; ++b)
        for (int c = 0; c < ng; ++c) r.T[(std::size_t(a) * ng + b) * ng + c] = cplx(0, -f(a, b, c));
    return r;
  }
  bool real = false;
  const std::vector<int> rows = parse_young_rows(N, spec, &real);
  const int n = std::accumulate(rows.begin(), rows.end(), 0);
  TensorSpace ts(N, n);
  std::vector<std::vector<double>> W;
  if (rows.size() == 1) symmetric_basis(ts, W); else young_basis(ts, rows, W);
  r.kind = 3; r.rows = rows; r.nbox = n; r.dimT = ts.dim; r.d = int(W.size()); r.nality = n % N; r.real = real;
  r.name = "general[";
  for (std::size_t i = 0; i < rows.size(); ++i) r.name += (i ? "," : "") + std::to_string(rows[i]);
  r.name += "]";
  r.W.resize(std::size_t(r.d) * ts.dim);
  for (int l = 0; l < r.d; ++l) std::copy(W[std::size_t(l)].begin(), W[std::size_t(l)].end(), r.W.begin() + std::size_t(l) * ts.dim);
  // T^a_R[k][l] = <W_k| sum_s T^a(slot s) |W_l>; W is sparse, so walk its nonzeros.
  const auto F = fundamental_generators(N);
  r.T.assign(std::size_t(ng) * r.d * r.d, cplx(0, 0));
  std::vector<cplx> v(ts.dim);
  for (int a = 0; a < ng; ++a)
    for (int l = 0; l < r.d; ++l) {
      std::fill(v.begin(), v.end(), cplx(0, 0));
      const double* wl = &r.W[std::size_t(l) * ts.dim];
      for (long idx = 0; idx < ts.dim; ++idx) {
        if (wl[idx] == 0.0) continue;
        for (int s = 0; s < n; ++s) {
          const long blk = ts.pw[n - 1 - s];
          const int dg = ts.digit(idx, s);
          const long base = idx - dg * blk;
          for (int j = 0; j < N; ++j) v[std::size_t(base + j * blk)] += F[(std::size_t(a) * N + j) * N + dg] * wl[idx];
        }
      }
      for (int k = 0; k < r.d; ++k) {
        const double* wk = &r.W[std::size_t(k) * ts.dim];
        cplx acc(0, 0);
        for (long i = 0; i < ts.dim; ++i) if (wk[i] != 0.0) acc += wk[i] * v[std::size_t(i)];
        r.T[(std::size_t(a) * r.d + k) * r.d + l] = acc;
      }
    }
  return r;
}

}  // namespace ghx
