// Device-side building blocks for the gfx950 kernels: complex scalars, N x N
// link algebra, su(N) generator maps (Bertlmann-Krammer order, as in the
// reference algebra.hpp:24-81), exp(iH), the counter-based RNG (rng.hpp:11-46)
// and the representation applies (fundamental, adjoint, small tensor irreps).
//
// Storage type T (float or double) is what sits in HBM; R is the compute type.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "gauss_cr.hpp"

namespace ghx {

#define GHX_DEV __device__ __forceinline__

template <typename R>
struct cx {
  R r, i;
};
template <typename R> GHX_DEV cx<R> cadd(cx<R> a, cx<R> b) { return {a.r + b.r, a.i + b.i}; }
template <typename R> GHX_DEV cx<R> csub(cx<R> a, cx<R> b) { return {a.r - b.r, a.i - b.i}; }
template <typename R> GHX_DEV cx<R> cmul(cx<R> a, cx<R> b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
// conj(a) * b
template <typename R> GHX_DEV cx<R> cmulc(cx<R> a, cx<R> b) { return {a.r * b.r + a.i * b.i, a.r * b.i - a.i * b.r}; }
template <typename R> GHX_DEV cx<R> cscale(cx<R> a, R s) { return {a.r * s, a.i * s}; }
template <typename R> GHX_DEV cx<R> cconj(cx<R> a) { return {a.r, -a.i}; }
// acc += a*b
template <typename R> GHX_DEV void cfma(cx<R>& acc, cx<R> a, cx<R> b) {
  acc.r = fma(a.r, b.r, acc.r); acc.r = fma(-a.i, b.i, acc.r);
  acc.i = fma(a.r, b.i, acc.i); acc.i = fma(a.i, b.r, acc.i);
}
// acc += conj(a)*b
template <typename R> GHX_DEV void cfmac(cx<R>& acc, cx<R> a, cx<R> b) {
  acc.r = fma(a.r, b.r, acc.r); acc.r = fma(a.i, b.i, acc.r);
  acc.i = fma(a.r, b.i, acc.i); acc.i = fma(-a.i, b.r, acc.i);
}
// acc += a*conj(b)
template <typename R> GHX_DEV void cfmcc(cx<R>& acc, cx<R> a, cx<R> b) {
  acc.r = fma(a.r, b.r, acc.r); acc.r = fma(a.i, b.i, acc.r);
  acc.i = fma(a.i, b.r, acc.i); acc.i = fma(-a.r, b.i, acc.i);
}
// acc += conj(a)*conj(b)
template <typename R> GHX_DEV void cfmcc2(cx<R>& acc, cx<R> a, cx<R> b) {
  acc.r = fma(a.r, b.r, acc.r); acc.r = fma(-a.i, b.i, acc.r);
  acc.i = fma(-a.r, b.i, acc.i); acc.i = fma(-a.i, b.r, acc.i);
}

template <typename R, int N>
struct Mat {
  cx<R> e[N][N];
};

template <typename R, int N> GHX_DEV Mat<R, N> mzero() {
  Mat<R, N> m;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) m.e[i][j] = {R(0), R(0)};
  return m;
}
template <typename R, int N> GHX_DEV Mat<R, N> meye() {
  Mat<R, N> m = mzero<R, N>();
#pragma unroll
  for (int i = 0; i < N; ++i) m.e[i][i].r = R(1);
  return m;
}
// C = A B
template <typename R, int N> GHX_DEV Mat<R, N> mul_nn(const Mat<R, N>& A, const Mat<R, N>& B) {
  Mat<R, N> C;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      cx<R> s = cmul(A.e[i][0], B.e[0][j]);
#pragma unroll
      for (int k = 1; k < N; ++k) cfma(s, A.e[i][k], B.e[k][j]);
      C.e[i][j] = s;
    }
  return C;
}
// C = A B^dag
template <typename R, int N> GHX_DEV Mat<R, N> mul_nd(const Mat<R, N>& A, const Mat<R, N>& B) {
  Mat<R, N> C;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      cx<R> s = {R(0), R(0)};
#pragma unroll
      for (int k = 0; k < N; ++k) cfmcc(s, A.e[i][k], B.e[j][k]);
      C.e[i][j] = s;
    }
  return C;
}
// C = A^dag B
template <typename R, int N> GHX_DEV Mat<R, N> mul_dn(const Mat<R, N>& A, const Mat<R, N>& B) {
  Mat<R, N> C;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      cx<R> s = {R(0), R(0)};
#pragma unroll
      for (int k = 0; k < N; ++k) cfmac(s, A.e[k][i], B.e[k][j]);
      C.e[i][j] = s;
    }
  return C;
}
// C = A^dag B^dag
template <typename R, int N> GHX_DEV Mat<R, N> mul_dd(const Mat<R, N>& A, const Mat<R, N>& B) {
  Mat<R, N> C;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      cx<R> s = {R(0), R(0)};
#pragma unroll
      for (int k = 0; k < N; ++k) cfmcc2(s, A.e[k][i], B.e[j][k]);
      C.e[i][j] = s;
    }
  return C;
}
// Acc += A B
template <typename R, int N> GHX_DEV void mac_nn(Mat<R, N>& Acc, const Mat<R, N>& A, const Mat<R, N>& B) {
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j)
#pragma unroll
      for (int k = 0; k < N; ++k) cfma(Acc.e[i][j], A.e[i][k], B.e[k][j]);
}
// Acc += A B^dag
template <typename R, int N> GHX_DEV void mac_nd(Mat<R, N>& Acc, const Mat<R, N>& A, const Mat<R, N>& B) {
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j)
#pragma unroll
      for (int k = 0; k < N; ++k) cfmcc(Acc.e[i][j], A.e[i][k], B.e[j][k]);
}
// Acc += A^dag B
template <typename R, int N> GHX_DEV void mac_dn(Mat<R, N>& Acc, const Mat<R, N>& A, const Mat<R, N>& B) {
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j)
#pragma unroll
      for (int k = 0; k < N; ++k) cfmac(Acc.e[i][j], A.e[k][i], B.e[k][j]);
}
// Rows 0..N-2 of A B and of A B^dag (the last row of A is not read; the last
// row of the result is left unset).
template <typename R, int N> GHX_DEV Mat<R, N> mul_nn_top(const Mat<R, N>& A, const Mat<R, N>& B) {
  Mat<R, N> C;
#pragma unroll
  for (int i = 0; i < N - 1; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      cx<R> s = cmul(A.e[i][0], B.e[0][j]);
#pragma unroll
      for (int k = 1; k < N; ++k) cfma(s, A.e[i][k], B.e[k][j]);
      C.e[i][j] = s;
    }
  return C;
}
template <typename R, int N> GHX_DEV Mat<R, N> mul_nd_top(const Mat<R, N>& A, const Mat<R, N>& B) {
  Mat<R, N> C;
#pragma unroll
  for (int i = 0; i < N - 1; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      cx<R> s = {R(0), R(0)};
#pragma unroll
      for (int k = 0; k < N; ++k) cfmcc(s, A.e[i][k], B.e[j][k]);
      C.e[i][j] = s;
    }
  return C;
}
// Rows 0..N-2 of A^dag B (every row of A and B is read).
template <typename R, int N> GHX_DEV Mat<R, N> mul_dn_top(const Mat<R, N>& A, const Mat<R, N>& B) {
  Mat<R, N> C;
#pragma unroll
  for (int i = 0; i < N - 1; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      cx<R> s = {R(0), R(0)};
#pragma unroll
      for (int k = 0; k < N; ++k) cfmac(s, A.e[k][i], B.e[k][j]);
      C.e[i][j] = s;
    }
  return C;
}
// Rows 0..N-2 of Acc += A B^dag and of Acc += A^dag B.
template <typename R, int N> GHX_DEV void mac_nd_top(Mat<R, N>& Acc, const Mat<R, N>& A, const Mat<R, N>& B) {
#pragma unroll
  for (int i = 0; i < N - 1; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j)
#pragma unroll
      for (int k = 0; k < N; ++k) cfmcc(Acc.e[i][j], A.e[i][k], B.e[j][k]);
}
template <typename R, int N> GHX_DEV void mac_dn_top(Mat<R, N>& Acc, const Mat<R, N>& A, const Mat<R, N>& B) {
#pragma unroll
  for (int i = 0; i < N - 1; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j)
#pragma unroll
      for (int k = 0; k < N; ++k) cfmac(Acc.e[i][j], A.e[k][i], B.e[k][j]);
}
template <typename R, int N> GHX_DEV R retrace(const Mat<R, N>& A) {
  R t = A.e[0][0].r;
#pragma unroll
  for (int i = 1; i < N; ++i) t += A.e[i][i].r;
  return t;
}
// Re Tr(A B)
template <typename R, int N> GHX_DEV R retr_nn(const Mat<R, N>& A, const Mat<R, N>& B) {
  R t = R(0);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int k = 0; k < N; ++k) t = fma(A.e[i][k].r, B.e[k][i].r, fma(-A.e[i][k].i, B.e[k][i].i, t));
  return t;
}
// Re Tr(A B^dag)
template <typename R, int N> GHX_DEV R retr_nd(const Mat<R, N>& A, const Mat<R, N>& B) {
  R t = R(0);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int k = 0; k < N; ++k) t = fma(A.e[i][k].r, B.e[i][k].r, fma(A.e[i][k].i, B.e[i][k].i, t));
  return t;
}
template <typename R, int N> GHX_DEV R fnorm2(const Mat<R, N>& A) {
  R t = R(0);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) t = fma(A.e[i][j].r, A.e[i][j].r, fma(A.e[i][j].i, A.e[i][j].i, t));
  return t;
}

// ---- su(N) algebra in the reference generator order (algebra.hpp:24-50) ----
template <int N> constexpr int NG = N * N - 1;
template <int N> constexpr int NPAIR = N * (N - 1) / 2;

// H = sum_a p[a] T^a, scaled by eps (Hermitian traceless)
template <typename R, int N> GHX_DEV Mat<R, N> alg_to_mat(const R* p, R eps) {
  Mat<R, N> H = mzero<R, N>();
  int a = 0;
#pragma unroll
  for (int j = 0; j < N; ++j)
#pragma unroll
    for (int k = j + 1; k < N; ++k) {
      const R s = R(0.5) * eps * p[a];
      const R t = R(0.5) * eps * p[a + NPAIR<N>];
      H.e[j][k] = {s, -t};
      H.e[k][j] = {s, t};
      ++a;
    }
#pragma unroll
  for (int l = 1; l < N; ++l) {
    const R c = R(0.5) * eps * p[2 * NPAIR<N> + l - 1] * R(sqrt(2.0 / (double(l) * (l + 1))));
#pragma unroll
    for (int j = 0; j < l; ++j) H.e[j][j].r += c;
    H.e[l][l].r -= c * R(l);
  }
  return H;
}
// f[a] = 2 Im Tr(T^a X) for all generators: the factor 1/2 of T^a left to the
// caller's coefficient (halving is exact, so c * im_tr_gen(X)[a] and
// (c / 2) * im_tr_gen2(X)[a] are the same number)
template <typename R, int N> GHX_DEV void im_tr_gen2(const Mat<R, N>& X, R* f) {
  int a = 0;
#pragma unroll
  for (int j = 0; j < N; ++j)
#pragma unroll
    for (int k = j + 1; k < N; ++k) {
      f[a] = X.e[k][j].i + X.e[j][k].i;
      f[a + NPAIR<N>] = X.e[j][k].r - X.e[k][j].r;
      ++a;
    }
#pragma unroll
  for (int l = 1; l < N; ++l) {
    R s = R(0);
#pragma unroll
    for (int j = 0; j < l; ++j) s += X.e[j][j].i;
    s -= R(l) * X.e[l][l].i;
    f[2 * NPAIR<N> + l - 1] = R(sqrt(2.0 / (double(l) * (l + 1)))) * s;
  }
}
// f[a] = Im Tr(T^a X) for all generators
template <typename R, int N> GHX_DEV void im_tr_gen(const Mat<R, N>& X, R* f) {
  int a = 0;
#pragma unroll
  for (int j = 0; j < N; ++j)
#pragma unroll
    for (int k = j + 1; k < N; ++k) {
      f[a] = R(0.5) * (X.e[k][j].i + X.e[j][k].i);
      f[a + NPAIR<N>] = R(0.5) * (X.e[j][k].r - X.e[k][j].r);
      ++a;
    }
#pragma unroll
  for (int l = 1; l < N; ++l) {
    R s = R(0);
#pragma unroll
    for (int j = 0; j < l; ++j) s += X.e[j][j].i;
    s -= R(l) * X.e[l][l].i;
    f[2 * NPAIR<N> + l - 1] = R(0.5) * R(sqrt(2.0 / (double(l) * (l + 1)))) * s;
  }
}
// complex components c[a] = 2 Tr(T^a X) (inverse of alg_to_mat for complex coefficients)
template <typename R, int N> GHX_DEV void tr_gen2(const Mat<R, N>& X, cx<R>* c) {
  int a = 0;
#pragma unroll
  for (int j = 0; j < N; ++j)
#pragma unroll
    for (int k = j + 1; k < N; ++k) {
      // 2 Tr(T_sym X) = X_kj + X_jk ; 2 Tr(T_asym X) = -i X_kj + i X_jk
      c[a] = cadd(X.e[k][j], X.e[j][k]);
      const cx<R> d = csub(X.e[j][k], X.e[k][j]);
      c[a + NPAIR<N>] = {-d.i, d.r};
      ++a;
    }
#pragma unroll
  for (int l = 1; l < N; ++l) {
    cx<R> s = {R(0), R(0)};
#pragma unroll
    for (int j = 0; j < l; ++j) s = cadd(s, X.e[j][j]);
    s = csub(s, cscale(X.e[l][l], R(l)));
    c[2 * NPAIR<N> + l - 1] = cscale(s, R(sqrt(2.0 / (double(l) * (l + 1)))));
  }
}
// X = sum_a c[a] T^a for complex c
template <typename R, int N> GHX_DEV Mat<R, N> calg_to_mat(const cx<R>* c) {
  Mat<R, N> H = mzero<R, N>();
  int a = 0;
#pragma unroll
  for (int j = 0; j < N; ++j)
#pragma unroll
    for (int k = j + 1; k < N; ++k) {
      const cx<R> s = cscale(c[a], R(0.5));
      const cx<R> t = cscale(c[a + NPAIR<N>], R(0.5));
      H.e[j][k] = {s.r + t.i, s.i - t.r};   // s - i t
      H.e[k][j] = {s.r - t.i, s.i + t.r};   // s + i t
      ++a;
    }
#pragma unroll
  for (int l = 1; l < N; ++l) {
    const cx<R> c2 = cscale(c[2 * NPAIR<N> + l - 1], R(0.5) * R(sqrt(2.0 / (double(l) * (l + 1)))));
#pragma unroll
    for (int j = 0; j < l; ++j) H.e[j][j] = cadd(H.e[j][j], c2);
    H.e[l][l] = csub(H.e[l][l], cscale(c2, R(l)));
  }
  return H;
}

// ---- exp(iH) -------------------------------------------------------------
// SU(2) closed form (sun.hpp:34-46): cos r I + i (sin r / r) H, r^2 = Tr(H^2)/2
template <typename R> GHX_DEV Mat<R, 2> expi_su2(const Mat<R, 2>& H) {
  const R r2 = R(0.5) * (H.e[0][0].r * H.e[0][0].r + H.e[1][1].r * H.e[1][1].r +
                         R(2) * (H.e[0][1].r * H.e[0][1].r + H.e[0][1].i * H.e[0][1].i));
  const R r = sqrt(r2 > R(0) ? r2 : R(0));
  R c, sc;
  if (r < R(1e-8)) { c = R(1) - R(0.5) * r2; sc = R(1) - r2 / R(6); }
  else { c = cos(r); sc = sin(r) / r; }
  Mat<R, 2> U;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) U.e[i][j] = {-sc * H.e[i][j].i, sc * H.e[i][j].r};
  U.e[0][0].r += c;
  U.e[1][1].r += c;
  return U;
}
// General N: scale until ||H||_F <= 1/2, Taylor to ORDER terms, square back
// (sun.hpp:12-30 uses ORDER = 18; the FP32 path uses a shorter series whose
// truncation error is below float rounding at ||H|| <= 1/2).
template <typename R, int N, int ORDER> GHX_DEV Mat<R, N> expi_taylor(const Mat<R, N>& H) {
  const R nrm = sqrt(fnorm2(H));
  int s = 0;
  R scaled = nrm, inv2s = R(1);
  while (scaled > R(0.5) && s < 64) { scaled *= R(0.5); inv2s *= R(0.5); ++s; }  // bounded: non-finite input cannot spin
  Mat<R, N> A;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) A.e[i][j] = {-inv2s * H.e[i][j].i, inv2s * H.e[i][j].r};
  Mat<R, N> term = meye<R, N>(), res = meye<R, N>();
#pragma unroll 1
  for (int k = 1; k <= ORDER; ++k) {
    term = mul_nn(term, A);
    const R ik = R(1) / R(k);
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) {
        term.e[i][j] = cscale(term.e[i][j], ik);
        res.e[i][j] = cadd(res.e[i][j], term.e[i][j]);
      }
  }
#pragma unroll 1
  for (int i = 0; i < s; ++i) res = mul_nn(res, res);
  return res;
}
// SU(3), FP32 drift: Cayley-Hamilton reduction of the exponential series.
// For traceless Hermitian Q, Q^3 = c1 Q + c0 1 with c1 = Tr(Q^2)/2, c0 = det Q,
// so Q^n = a_n 1 + b_n Q + c_n Q^2 with (a,b,c)_{n+1} = (c0 c_n, a_n + c1 c_n, b_n)
// and exp(iQ) = f0 1 + f1 Q + f2 Q^2, f_j = sum_n i^n/n! (a,b,c)_n. After scaling
// to ||Q||_F <= 1/2 the series is cut at n = 8 (remainder <= 2^-9 / 9! = 5e-9,
// below float rounding); the result is squared back. An MD step's eps P has
// ||.||_F near 0.1, so the squaring is rare. Costs one Hermitian Q^2 instead
// of a Taylor chain.
template <typename R> GHX_DEV Mat<R, 3> expi_su3_ch(const Mat<R, 3>& H) {
  R n2 = fnorm2(H), sc = R(1);
  int s = 0;
  while (n2 * sc * sc > R(0.25) && s < 64) { sc *= R(0.5); ++s; }  // bounded: non-finite input cannot spin
  Mat<R, 3> Q;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Q.e[i][j] = cscale(H.e[i][j], sc);
  Mat<R, 3> Q2;  // Hermitian: fill the upper triangle, mirror the rest
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = i; j < 3; ++j) {
      cx<R> z = cmul(Q.e[i][0], Q.e[0][j]);
      cfma(z, Q.e[i][1], Q.e[1][j]);
      cfma(z, Q.e[i][2], Q.e[2][j]);
      Q2.e[i][j] = z;
      if (j > i) Q2.e[j][i] = cconj(z);
    }
  Q2.e[0][0].i = Q2.e[1][1].i = Q2.e[2][2].i = R(0);
  const R c1 = R(0.5) * (Q2.e[0][0].r + Q2.e[1][1].r + Q2.e[2][2].r);
  R c0 = R(0);  // det Q = Tr(Q^3)/3 = Re Tr(Q Q^2)/3
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) c0 = fma(Q.e[i][k].r, Q2.e[k][i].r, fma(-Q.e[i][k].i, Q2.e[k][i].i, c0));
  c0 *= R(1.0 / 3.0);
  R a = R(1), b = R(0), c = R(0);
  R f0r = R(1), f0i = R(0), f1r = R(0), f1i = R(0), f2r = R(0), f2i = R(0);
  R coef = R(1);
#pragma unroll
  for (int n = 1; n <= 8; ++n) {
    const R an = c0 * c, bn = fma(c1, c, a), cn = b;
    a = an; b = bn; c = cn;
    coef *= R(1.0 / n);
    // i^n: n%4 == 0 -> 1, 1 -> i, 2 -> -1, 3 -> -i
    if (n % 4 == 0) { f0r += coef * a; f1r += coef * b; f2r += coef * c; }
    else if (n % 4 == 1) { f0i += coef * a; f1i += coef * b; f2i += coef * c; }
    else if (n % 4 == 2) { f0r -= coef * a; f1r -= coef * b; f2r -= coef * c; }
    else { f0i -= coef * a; f1i -= coef * b; f2i -= coef * c; }
  }
  const cx<R> f1 = {f1r, f1i}, f2 = {f2r, f2i};
  Mat<R, 3> E;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cx<R> z = cmul(f2, Q2.e[i][j]);
      cfma(z, f1, Q.e[i][j]);
      E.e[i][j] = z;
    }
#pragma unroll
  for (int i = 0; i < 3; ++i) { E.e[i][i].r += f0r; E.e[i][i].i += f0i; }
#pragma unroll 1
  for (int i = 0; i < s; ++i) E = mul_nn(E, E);
  return E;
}
template <typename R, int N> GHX_DEV Mat<R, N> expi(const Mat<R, N>& H) {
  if constexpr (N == 2) return expi_su2<R>(H);
  else if constexpr (sizeof(R) == 8) return expi_taylor<R, N, 18>(H);
  else return expi_su3_ch<R>(H);
}

// Modified Gram-Schmidt on rows + det phase on row 0 (sun.hpp:49-72), in R.
template <typename R, int N> GHX_DEV void reunitarize(Mat<R, N>& U) {
#pragma unroll
  for (int i = 0; i < N; ++i) {
#pragma unroll
    for (int k = 0; k < i; ++k) {
      cx<R> pr = {R(0), R(0)};
#pragma unroll
      for (int j = 0; j < N; ++j) cfmac(pr, U.e[k][j], U.e[i][j]);
#pragma unroll
      for (int j = 0; j < N; ++j) U.e[i][j] = csub(U.e[i][j], cmul(pr, U.e[k][j]));
    }
    R nr = R(0);
#pragma unroll
    for (int j = 0; j < N; ++j) nr += U.e[i][j].r * U.e[i][j].r + U.e[i][j].i * U.e[i][j].i;
    const R inv = R(1) / sqrt(nr);
#pragma unroll
    for (int j = 0; j < N; ++j) U.e[i][j] = cscale(U.e[i][j], inv);
  }
  cx<R> d;
  if constexpr (N == 2) {
    d = csub(cmul(U.e[0][0], U.e[1][1]), cmul(U.e[0][1], U.e[1][0]));
  } else {
    const cx<R> m0 = csub(cmul(U.e[1][1], U.e[2][2]), cmul(U.e[1][2], U.e[2][1]));
    const cx<R> m1 = csub(cmul(U.e[1][0], U.e[2][2]), cmul(U.e[1][2], U.e[2][0]));
    const cx<R> m2 = csub(cmul(U.e[1][0], U.e[2][1]), cmul(U.e[1][1], U.e[2][0]));
    d = cadd(csub(cmul(U.e[0][0], m0), cmul(U.e[0][1], m1)), cmul(U.e[0][2], m2));
  }
  const R ad = sqrt(d.r * d.r + d.i * d.i);
  if (ad > R(0)) {
    const cx<R> ph = {d.r / ad, -d.i / ad};  // conj(det/|det|)
#pragma unroll
    for (int j = 0; j < N; ++j) U.e[0][j] = cmul(U.e[0][j], ph);
  }
}
template <typename R, int N> GHX_DEV R unitarity_violation(const Mat<R, N>& U) {
  Mat<R, N> d = mul_dn(U, U);
#pragma unroll
  for (int i = 0; i < N; ++i) d.e[i][i].r -= R(1);
  return sqrt(fnorm2(d));
}

// ---- counter-based RNG (rng.hpp:11-46) -----------------------------------
GHX_DEV uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}
__host__ __device__ inline uint64_t mix64_hd(uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}
__host__ __device__ inline uint64_t stream_key(uint64_t a, uint64_t b = 0, uint64_t c = 0, uint64_t d = 0, uint64_t e = 0) {
  const uint64_t g = 0x9E3779B97F4A7C15ULL;
  uint64_t h = mix64_hd(a);
  h = mix64_hd(h ^ (b + g));
  h = mix64_hd(h ^ (c + g));
  h = mix64_hd(h ^ (d + g));
  h = mix64_hd(h ^ (e + g));
  return h;
}
__host__ __device__ inline double draw_uniform(uint64_t seed, uint64_t k) {
  const uint64_t h = mix64_hd(seed ^ mix64_hd(k));
  return (double(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
// Box-Muller with correctly rounded log, cos and sqrt (gauss_cr.hpp) and IEEE
// products, as the reference's glibc/x86 arithmetic, so a draw equals the
// reference's wherever the host libm's log and cos round correctly.
GHX_DEV double draw_gauss(uint64_t seed, uint64_t k) {
  const double u1 = draw_uniform(seed, mix64(k * 2 + 1));
  const double u2 = draw_uniform(seed, mix64(k * 2 + 2));
  return gcr::sqrt_cr(-2.0 * gcr::log_cr(u1)) * gcr::cos_cr(6.283185307179586476925286766559 * u2);
}
// The same draw rounded to float, for FP32 storage. A float only needs the
// draw to the nearest float: the device's own double log, cos and sqrt (a few
// double ulp) give it unless the draw lies within 2^-46 (relative) of a float
// rounding boundary, about 2 draws in 10^7; there the correctly rounded draw
// decides. The result is float(draw_gauss(seed, k)) either way.
GHX_DEV float draw_gauss_f32(uint64_t seed, uint64_t k) {
  const double u1 = draw_uniform(seed, mix64(k * 2 + 1));
  const double u2 = draw_uniform(seed, mix64(k * 2 + 2));
  const double g = sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925286766559 * u2);
  const float lo = float(g * (1.0 - 0x1p-46)), hi = float(g * (1.0 + 0x1p-46));
  if (lo == hi) return lo;
  return float(draw_gauss(seed, k));
}

}  // namespace ghx
