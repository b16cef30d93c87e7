// Host construction of the symmetric-power engine (see symrep.hpp).
#include "symrep.hpp"

#include <cmath>
#include <complex>
#include <map>

namespace ghx {

namespace {

using cplx = std::complex<double>;

// Multisets of size k over N colours (type vectors), in a fixed order.
std::vector<std::vector<int>> multisets(int N, int k) {
  std::vector<std::vector<int>> out;
  std::vector<int> cur(N, 0);
  auto rec = [&](auto&& self, int i, int rem) -> void {
    if (i == N - 1) { cur[i] = rem; out.push_back(cur); return; }
    for (int a = rem; a >= 0; --a) { cur[i] = a; self(self, i + 1, rem - a); }
  };
  rec(rec, 0, k);
  return out;
}

double factorial(int n) { double f = 1.0; for (int i = 2; i <= n; ++i) f *= i; return f; }
double mfact(const std::vector<int>& m) { double f = 1.0; for (int x : m) f *= factorial(x); return f; }

// Sym^k(U) in the orthonormal monomial basis, from the recursion of the engine.
std::vector<cplx> sym_matrix(const SymBuild& s, const std::vector<cplx>& U, int k) {
  const int N = s.N;
  std::vector<cplx> prev(1, cplx(1, 0)), cur;
  int sp = 1;
  for (int l = 1; l <= k; ++l) {
    const int off = s.lvl_off[l], sk = s.lvl_off[l + 1] - off;
    cur.assign(size_t(sk) * sk, cplx(0, 0));
    for (int m = 0; m < sk; ++m)
      for (int n = 0; n < sk; ++n) {
        const int j = s.colj[off + n], np = s.colsrc[off + n];
        cplx acc(0, 0);
        for (int i = 0; i < N; ++i) {
          const int ri = s.rowsrc[size_t(off + m) * N + i];
          if (ri >= 0) acc += U[size_t(i) * N + j] * prev[size_t(ri) * sp + np];
        }
        cur[size_t(m) * sk + n] = acc;
      }
    prev.swap(cur);
    sp = sk;
  }
  const int off = s.lvl_off[k];
  for (int m = 0; m < sp; ++m)
    for (int n = 0; n < sp; ++n) prev[size_t(m) * sp + n] *= s.fnorm[off + m] / s.fnorm[off + n];
  return prev;
}

// U^(x n) applied to v (length N^n, slot 0 most significant).
std::vector<cplx> tensor_apply(const std::vector<cplx>& U, std::vector<cplx> v, int N, int n) {
  long dimT = 1;
  for (int i = 0; i < n; ++i) dimT *= N;
  std::vector<cplx> w(v.size());
  long blk = dimT;
  for (int slot = 0; slot < n; ++slot) {
    blk /= N;
    std::fill(w.begin(), w.end(), cplx(0, 0));
    for (long t = 0; t < dimT; ++t) {
      const int digit = int((t / blk) % N);
      const long base = t - digit * blk;
      for (int i = 0; i < N; ++i) w[size_t(base + i * blk)] += U[size_t(i) * N + digit] * v[size_t(t)];
    }
    v.swap(w);
  }
  return v;
}

}  // namespace

SymBuild build_sym(const RepInfo& r) {
  SymBuild s;
  if (r.kind != 3 || r.rows.empty() || r.rows.size() > 2) return s;
  const int N = r.N, d = r.d;
  s.N = N;
  s.d = d;
  s.k1 = r.rows[0];
  s.k2 = r.rows.size() > 1 ? r.rows[1] : 0;
  // per-level tables, levels 0..k1
  std::vector<std::vector<std::vector<int>>> lv;
  std::vector<std::map<std::vector<int>, int>> idx;
  for (int k = 0; k <= s.k1; ++k) {
    lv.push_back(multisets(N, k));
    std::map<std::vector<int>, int> mp;
    for (size_t i = 0; i < lv[k].size(); ++i) mp[lv[k][i]] = int(i);
    idx.push_back(mp);
  }
  if (int(lv[s.k1].size()) > kSymMaxS) return s;
  s.s1 = int(lv[s.k1].size());
  s.s2 = s.k2 > 0 ? int(lv[s.k2].size()) : 1;
  s.dimS = s.s1 * s.s2;
  if (s.dimS > kSymMaxDimS) return s;
  s.lvl_off.assign(s.k1 + 2, 0);
  for (int k = 0; k <= s.k1; ++k) s.lvl_off[k + 1] = s.lvl_off[k] + int(lv[k].size());
  const int tot = s.lvl_off[s.k1 + 1];
  s.colj.assign(tot, 0);
  s.colsrc.assign(tot, 0);
  s.rowsrc.assign(size_t(tot) * N, -1);
  s.fnorm.assign(tot, 1.0);
  for (int k = 0; k <= s.k1; ++k)
    for (size_t e = 0; e < lv[k].size(); ++e) {
      const int g = s.lvl_off[k] + int(e);
      const std::vector<int>& m = lv[k][e];
      s.fnorm[g] = std::sqrt(mfact(m));
      if (k == 0) continue;
      int j = 0;
      while (m[j] == 0) ++j;
      std::vector<int> mj = m;
      --mj[j];
      s.colj[g] = j;
      s.colsrc[g] = idx[k - 1].at(mj);
      for (int i = 0; i < N; ++i) {
        if (m[i] == 0) continue;
        std::vector<int> mi = m;
        --mi[i];
        s.rowsrc[size_t(g) * N + i] = idx[k - 1].at(mi);
      }
    }
  // device recipes and normalisations
  s.rec_off.assign(s.k1 + 2, 0);
  for (int k = 0; k <= s.k1; ++k) s.rec_off[k + 1] = s.rec_off[k] + int(lv[k].size() * lv[k].size());
  s.recipe.assign(size_t(s.rec_off[s.k1 + 1]) * 4, -1);
  for (int k = 1; k <= s.k1; ++k) {
    const int sk = int(lv[k].size()), sp = int(lv[k - 1].size()), off = s.lvl_off[k];
    for (int m = 0; m < sk; ++m)
      for (int n = 0; n < sk; ++n) {
        int* rc = &s.recipe[size_t(s.rec_off[k] + m * sk + n) * 4];
        const int np = s.colsrc[off + n];
        for (int i = 0; i < N; ++i) {
          const int ri = s.rowsrc[size_t(off + m) * N + i];
          rc[i] = ri >= 0 ? ri * sp + np : -1;
        }
        rc[3] = s.colj[off + n];
      }
  }
  auto scales = [&](int k) {
    std::vector<double> sc;
    const int sk = int(lv[k].size()), off = s.lvl_off[k];
    for (int m = 0; m < sk; ++m)
      for (int n = 0; n < sk; ++n) sc.push_back(s.fnorm[off + m] / s.fnorm[off + n]);
    return sc;
  };
  s.scale1 = scales(s.k1);
  s.scale2 = s.k2 > 0 ? scales(s.k2) : std::vector<double>(1, 1.0);
  // Yc[(a,b)][l] = sum over tuples t (row-1 slots then row-2 slots) of the
  // normalised monomial vectors times W_l[t]
  const int n = s.k1 + s.k2;
  const long dimT = r.dimT;
  const double n1 = factorial(s.k1), n2 = factorial(s.k2);
  s.Yc.assign(size_t(s.dimS) * d, 0.0);
  std::vector<int> ty1(N), ty2(N);
  for (long t = 0; t < dimT; ++t) {
    std::fill(ty1.begin(), ty1.end(), 0);
    std::fill(ty2.begin(), ty2.end(), 0);
    long rem = t;
    for (int slot = n - 1; slot >= 0; --slot) {
      const int digit = int(rem % N);
      rem /= N;
      if (slot < s.k1) ++ty1[digit];
      else ++ty2[digit];
    }
    const int a = idx[s.k1].at(ty1), b = s.k2 > 0 ? idx[s.k2].at(ty2) : 0;
    const double w = std::sqrt(mfact(ty1) / n1) * (s.k2 > 0 ? std::sqrt(mfact(ty2) / n2) : 1.0);
    for (int l = 0; l < d; ++l) {
      const double wl = r.W[size_t(l) * dimT + t];
      if (wl != 0.0) s.Yc[size_t(a * s.s2 + b) * d + l] += w * wl;
    }
  }
  // Yc^T Yc = c I
  double c = 0.0, off = 0.0;
  for (int l = 0; l < d; ++l)
    for (int m = 0; m < d; ++m) {
      double g = 0.0;
      for (int q = 0; q < s.dimS; ++q) g += s.Yc[size_t(q) * d + l] * s.Yc[size_t(q) * d + m];
      if (l == m) c += g / d;
      else off = std::max(off, std::fabs(g));
    }
  if (!(c > 1e-12) || off > 1e-10 * c) return s;
  s.invc = 1.0 / c;
  // check against the tensor apply on generic links
  double worst = 0.0;
  for (int trial = 0; trial < 3; ++trial) {
    // a generic SU(N) element: Gram-Schmidt of a fixed complex matrix
    std::vector<cplx> U(size_t(N) * N);
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j) U[size_t(i) * N + j] = cplx(std::sin(1.3 * i + 0.7 * j + trial), std::cos(0.9 * i - 1.1 * j + 2.0 * trial));
    for (int i = 0; i < N; ++i) {
      for (int p = 0; p < i; ++p) {
        cplx dot(0, 0);
        for (int j = 0; j < N; ++j) dot += std::conj(U[size_t(p) * N + j]) * U[size_t(i) * N + j];
        for (int j = 0; j < N; ++j) U[size_t(i) * N + j] -= dot * U[size_t(p) * N + j];
      }
      double nr = 0.0;
      for (int j = 0; j < N; ++j) nr += std::norm(U[size_t(i) * N + j]);
      for (int j = 0; j < N; ++j) U[size_t(i) * N + j] /= std::sqrt(nr);
    }
    const std::vector<cplx> S1 = sym_matrix(s, U, s.k1);
    const std::vector<cplx> S2 = s.k2 > 0 ? sym_matrix(s, U, s.k2) : std::vector<cplx>(1, cplx(1, 0));
    for (int l = 0; l < d; ++l) {
      // tensor: W U^(x n) W_l
      std::vector<cplx> v(static_cast<size_t>(dimT));
      for (long t = 0; t < dimT; ++t) v[size_t(t)] = r.W[size_t(l) * dimT + t];
      const std::vector<cplx> tv = tensor_apply(U, v, N, n);
      // engine: (1/c) Yc^T (S1 (x) S2) Yc e_l
      std::vector<cplx> yl(s.dimS), w(s.dimS, cplx(0, 0));
      for (int q = 0; q < s.dimS; ++q) yl[size_t(q)] = s.Yc[size_t(q) * d + l];
      for (int a = 0; a < s.s1; ++a)
        for (int b = 0; b < s.s2; ++b) {
          cplx acc(0, 0);
          for (int a2 = 0; a2 < s.s1; ++a2)
            for (int b2 = 0; b2 < s.s2; ++b2)
              acc += S1[size_t(a) * s.s1 + a2] * S2[size_t(b) * s.s2 + b2] * yl[size_t(a2 * s.s2 + b2)];
          w[size_t(a * s.s2 + b)] = acc;
        }
      for (int k = 0; k < d; ++k) {
        cplx ref(0, 0), eng(0, 0);
        for (long t = 0; t < dimT; ++t) ref += r.W[size_t(k) * dimT + t] * tv[size_t(t)];
        for (int q = 0; q < s.dimS; ++q) eng += s.Yc[size_t(q) * d + k] * w[size_t(q)];
        worst = std::max(worst, std::abs(eng * s.invc - ref));
      }
    }
  }
  s.max_err = worst;
  s.ok = worst < 1e-10;
  return s;
}

}  // namespace ghx
