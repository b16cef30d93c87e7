// Host construction of the symmetric-power engine (see symrep.hpp).
#include "symrep.hpp"

#include <cmath>
#include <complex>
#include <algorithm>
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

// A generic unitary N x N matrix: Gram-Schmidt of a fixed complex matrix.
std::vector<cplx> generic_sun(int N, int trial) {
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
  return U;
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
  if (s.dimS > kSymMaxDimS || d > kSymMaxD) return s;
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
    const std::vector<cplx> U = generic_sun(N, trial);
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

SymProgram sym_program(const SymBuild& s) {
  SymProgram p;
  const int N = s.N, k1 = s.k1, k2 = s.k2;
  if (N > 3) return p;  // an entry has three sources and the link's columns 0..2 (3 is the copy column)
  auto sk = [&](int k) { return s.lvl_off[k + 1] - s.lvl_off[k]; };
  // level k lives in Ca when k1 - k is even, so that C_k1 ends in Ca
  p.s1p = s.s1 | 1;  // odd row stride of C_k1
  auto stride = [&](int k) { return k == k1 ? p.s1p : sk(k); };
  int ca = 0, cb = 0;
  for (int k = 0; k <= k1; ++k) ((k1 - k) % 2 == 0 ? ca : cb) = std::max((k1 - k) % 2 == 0 ? ca : cb, sk(k) * stride(k));
  const int Ca = kSymLdsBase, Cb = Ca + ca;
  p.lds_Cb = Cb;
  auto base = [&](int k) { return (k1 - k) % 2 == 0 ? Ca : Cb; };
  const bool copy2 = k2 > 0 && k2 < k1;
  p.lds_S1 = Ca;
  p.lds_S2 = copy2 ? Cb + cb : (k2 == k1 ? Ca : 0);
  p.s2p = (k2 == k1) ? p.s1p : s.s2;
  const int after = Cb + cb + (copy2 ? s.s2 * s.s2 : 0);
  // V, T1 and W hold both hops of a link, P = 4 ng entries per direction; W
  // takes V's place once C1 V is formed (or is T1 itself for one-row irreps)
  p.ng = (s.dimS + 3) / 4;
  const int P = 4 * p.ng;
  p.lds_V = cb >= 4 * P ? Cb : after;  // Cb is free once the build is done
  p.lds_T1 = p.lds_V + 2 * P;
  p.lds_W = s.s2 == 1 ? p.lds_T1 : p.lds_V;
  p.lds_total = std::max(after, p.lds_T1 + 2 * P);
  p.lds_one = base(0);
  p.pair_C = 0;
  p.pair_S2 = copy2 ? s.s1 * p.s1p : 0;
  p.pair_V = s.s1 * p.s1p + (copy2 ? s.s2 * s.s2 : 0);
  p.pair_T1 = s.s2 <= kSymRowS2 ? p.pair_V : p.pair_V + 2 * P;
  p.pair_total = p.pair_T1 + 2 * P;
  auto flush = [&](std::vector<int>& pass) {  // pad a pass to 64 entries (idle lanes write slot 1)
    while (pass.size() < size_t(4 * 64)) { pass.push_back(0); pass.push_back(0); pass.push_back(0); pass.push_back((1 << 2) | 3); }
    p.prog.insert(p.prog.end(), pass.begin(), pass.end());
    pass.clear();
    ++p.npass;
  };
  std::vector<int> pass;
  for (int k = 1; k <= k1; ++k) {
    const int n_k = sk(k), sp = sk(k - 1), off = s.lvl_off[k];
    for (int m = 0; m < n_k; ++m)
      for (int n = 0; n < n_k; ++n) {
        const int np = s.colsrc[off + n];
        int src[3] = {0, 0, 0};
        for (int i = 0; i < N; ++i) {
          const int ri = s.rowsrc[size_t(off + m) * N + i];
          if (ri >= 0) src[i] = base(k - 1) + ri * sp + np;
        }
        pass.insert(pass.end(), {src[0], src[1], src[2], ((base(k) + m * stride(k) + n) << 2) | s.colj[off + n]});
        if (pass.size() == size_t(4 * 64)) flush(pass);
      }
    if (!pass.empty()) flush(pass);
    if (k == k2 && copy2) {  // save C_k2 before level k2 + 2 overwrites it
      for (int e = 0; e < s.s2 * s.s2; ++e) {
        pass.insert(pass.end(), {base(k2) + e, 0, 0, ((p.lds_S2 + e) << 2) | 3});
        if (pass.size() == size_t(4 * 64)) flush(pass);
      }
      if (!pass.empty()) flush(pass);
    }
  }
  if (p.npass % 2) flush(pass);  // an even number of passes (an idle one at the end)
  // F over q = (a, b) and the folded lift / projection tables
  std::vector<double> F(s.dimS);
  const int o1 = s.lvl_off[k1], o2 = s.lvl_off[k2];
  for (int a = 0; a < s.s1; ++a)
    for (int b = 0; b < s.s2; ++b) F[size_t(a * s.s2 + b)] = s.fnorm[o1 + a] * (k2 > 0 ? s.fnorm[o2 + b] : 1.0);
  p.pre.resize(size_t(2) * s.dimS);
  p.Yp.assign(size_t(2) * p.ng * s.d * 4, 0.0);
  auto yp = [&](int dir, int q, int l) -> double& { return p.Yp[((size_t(dir) * p.ng + q / 4) * s.d + l) * 4 + q % 4]; };
  for (int q = 0; q < s.dimS; ++q) {
    p.pre[q] = 1.0 / F[q];
    p.pre[s.dimS + q] = F[q];
    for (int l = 0; l < s.d; ++l) {
      const double y = s.Yc[size_t(q) * s.d + l] * s.invc;
      yp(0, q, l) = y * F[q];
      yp(1, q, l) = y / F[q];
    }
  }
  // the same tables as sparse rows (exact zeros dropped: entries of another
  // colour content are products with zero)
  for (int dir = 0; dir < 2; ++dir)
    for (int l = 0; l < s.d; ++l) {
      int nz = 0;
      for (int q = 0; q < s.dimS; ++q) nz += yp(dir, q, l) != 0.0;
      p.ynz = std::max(p.ynz, nz);
    }
  p.ysq.assign(size_t(2) * p.ynz * s.d, 0);
  p.ysv.assign(p.ysq.size(), 0.0);
  for (int dir = 0; dir < 2; ++dir)
    for (int l = 0; l < s.d; ++l) {
      int k = 0;
      for (int q = 0; q < s.dimS; ++q)
        if (yp(dir, q, l) != 0.0) {
          const size_t e = (size_t(dir) * p.ynz + k++) * s.d + l;
          p.ysq[e] = q;
          p.ysv[e] = yp(dir, q, l);
        }
    }
  // run the program on the host, as the device does, against the engine's
  // matrices (already checked against the tensor apply): every column of
  // D(U) and of D(U)^dag on two generic links
  double worst = 0.0;
  for (int trial = 0; trial < 2; ++trial) {
    const std::vector<cplx> U = generic_sun(N, trial + 7);
    std::vector<cplx> L(size_t(p.lds_total), cplx(0, 0));
    L[size_t(p.lds_one)] = cplx(1, 0);
    for (int i = 0; i < N; ++i) {
      for (int j = 0; j < N; ++j) L[size_t(kSymLdsUs + 4 * i + j)] = U[size_t(i) * N + j];
      L[size_t(kSymLdsUs + 4 * i + 3)] = cplx(i == 0 ? 1 : 0, 0);
    }
    for (int ps = 0; ps < p.npass; ++ps) {
      std::vector<cplx> out(64);
      for (int t = 0; t < 64; ++t) {
        const int* e = &p.prog[size_t(ps * 64 + t) * 4];
        const int j = e[3] & 3;
        cplx acc(0, 0);
        for (int i = 0; i < N; ++i) acc += L[size_t(kSymLdsUs + 4 * i + j)] * L[size_t(e[i])];
        out[size_t(t)] = acc;
      }
      for (int t = 0; t < 64; ++t) L[size_t(p.prog[size_t(ps * 64 + t) * 4 + 3] >> 2)] = out[size_t(t)];
      L[0] = cplx(0, 0);
    }
    const std::vector<cplx> S1 = sym_matrix(s, U, k1);
    const std::vector<cplx> S2 = k2 > 0 ? sym_matrix(s, U, k2) : std::vector<cplx>(1, cplx(1, 0));
    for (int dir = 0; dir < 2; ++dir)
      for (int l = 0; l < s.d; ++l) {
        // device path: V = pre (Yc e_l), W = (C1 x C2) V (or its adjoint), out = the sparse rows of Yp^T applied to W
        std::vector<cplx> V(s.dimS), W(s.dimS, cplx(0, 0)), ref(s.dimS, cplx(0, 0));
        for (int q = 0; q < s.dimS; ++q) V[size_t(q)] = p.pre[size_t(dir) * s.dimS + q] * s.Yc[size_t(q) * s.d + l];
        for (int a = 0; a < s.s1; ++a)
          for (int b = 0; b < s.s2; ++b)
            for (int a2 = 0; a2 < s.s1; ++a2)
              for (int b2 = 0; b2 < s.s2; ++b2) {
                const cplx c1 = dir ? std::conj(L[size_t(p.lds_S1 + a2 * p.s1p + a)]) : L[size_t(p.lds_S1 + a * p.s1p + a2)];
                const cplx c2 = k2 == 0 ? cplx(1, 0)
                                        : (dir ? std::conj(L[size_t(p.lds_S2 + b2 * p.s2p + b)]) : L[size_t(p.lds_S2 + b * p.s2p + b2)]);
                W[size_t(a * s.s2 + b)] += c1 * c2 * V[size_t(a2 * s.s2 + b2)];
                const cplx r1 = dir ? std::conj(S1[size_t(a2) * s.s1 + a]) : S1[size_t(a) * s.s1 + a2];
                const cplx r2 = dir ? std::conj(S2[size_t(b2) * s.s2 + b]) : S2[size_t(b) * s.s2 + b2];
                ref[size_t(a * s.s2 + b)] += r1 * r2 * s.Yc[size_t(a2 * s.s2 + b2) * s.d + l];
              }
        for (int k = 0; k < s.d; ++k) {
          cplx dev(0, 0), eng(0, 0);
          for (int z = 0; z < p.ynz; ++z) {
            const size_t e = (size_t(dir) * p.ynz + z) * s.d + k;
            dev += p.ysv[e] * W[size_t(p.ysq[e])];
          }
          for (int q = 0; q < s.dimS; ++q) eng += s.Yc[size_t(q) * s.d + k] * s.invc * ref[size_t(q)];
          worst = std::max(worst, std::abs(dev - eng));
        }
      }
  }
  p.max_err = worst;
  p.ok = worst < 1e-10;
  return p;
}

}  // namespace ghx
