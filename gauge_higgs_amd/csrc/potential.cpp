// Host construction of the multi-invariant quartic potential
// (scalar_invariants.hpp:20-208): the adjoint Casimir superoperator on
// M = phi phi^dag, its distinct eigenvalues (the channels of R x Rbar) and the
// phi-independent superoperator A = sum_c 2 f_c P_c, with P_c the channel
// projector written as the Lagrange polynomial in Chat. The device evaluates
// dV/dphi^* = -mu2 phi + A(M) phi and V = -mu2 |phi|^2 + Re <M, A(M)>/2 per site.
#include "potential.hpp"

#include <algorithm>
#include <cmath>

namespace ghx {

using cplx = std::complex<double>;

// Chat[(p,q),(i,j)] = sum_a (T_a^2)_{pi} d_{jq} - 2 (T_a)_{pi} (T_a)_{jq} + d_{pi} (T_a^2)_{jq},
// the matrix of M -> sum_a [T_a, [T_a, M]] on vec(M) (index i*d + j).
std::vector<cplx> casimir_superop(const RepInfo& r) {
  const int d = r.d, ngn = r.N * r.N - 1, dd = d * d;
  std::vector<cplx> C(size_t(dd) * dd, cplx(0, 0));
  std::vector<cplx> T2(size_t(d) * d);
  for (int a = 0; a < ngn; ++a) {
    const cplx* T = r.T.data() + size_t(a) * d * d;
    for (int i = 0; i < d; ++i)
      for (int j = 0; j < d; ++j) {
        cplx s(0, 0);
        for (int k = 0; k < d; ++k) s += T[i * d + k] * T[k * d + j];
        T2[size_t(i) * d + j] = s;
      }
    for (int p = 0; p < d; ++p)
      for (int q = 0; q < d; ++q)
        for (int i = 0; i < d; ++i)
          for (int j = 0; j < d; ++j) {
            cplx v = -2.0 * T[p * d + i] * T[j * d + q];
            if (j == q) v += T2[size_t(p) * d + i];
            if (p == i) v += T2[size_t(j) * d + q];
            C[size_t(p * d + q) * dd + (i * d + j)] += v;
          }
  }
  return C;
}

namespace {
// Eigenvalues of a real symmetric n x n matrix (cyclic Jacobi rotations).
std::vector<double> symmetric_eigenvalues(std::vector<double> A, int n) {
  for (int sweep = 0; sweep < 200; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) off += A[size_t(p) * n + q] * A[size_t(p) * n + q];
    if (off < 1e-30) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = A[size_t(p) * n + q];
        if (std::fabs(apq) < 1e-300) continue;
        const double th = (A[size_t(q) * n + q] - A[size_t(p) * n + p]) / (2.0 * apq);
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          double& x = A[size_t(k) * n + p];
          double& y = A[size_t(k) * n + q];
          const double xv = x, yv = y;
          x = c * xv - s * yv;
          y = s * xv + c * yv;
        }
        for (int k = 0; k < n; ++k) {
          double& x = A[size_t(p) * n + k];
          double& y = A[size_t(q) * n + k];
          const double xv = x, yv = y;
          x = c * xv - s * yv;
          y = s * xv + c * yv;
        }
      }
  }
  std::vector<double> ev(n);
  for (int i = 0; i < n; ++i) ev[i] = A[size_t(i) * n + i];
  return ev;
}
}  // namespace

// Distinct eigenvalues of the Hermitian Chat, through its real 2dd x 2dd
// embedding [[Re, -Im], [Im, Re]] (every eigenvalue appears twice there),
// merged within 1e-4 as scalar_invariants.hpp:186 does.
std::vector<double> casimir_channels(const RepInfo& r) {
  const int dd = r.d * r.d, n = 2 * dd;
  const std::vector<cplx> C = casimir_superop(r);
  std::vector<double> E(size_t(n) * n);
  for (int i = 0; i < dd; ++i)
    for (int j = 0; j < dd; ++j) {
      const cplx z = C[size_t(i) * dd + j];
      E[size_t(i) * n + j] = z.real();
      E[size_t(i) * n + dd + j] = -z.imag();
      E[size_t(dd + i) * n + j] = z.imag();
      E[size_t(dd + i) * n + dd + j] = z.real();
    }
  std::vector<double> ev = symmetric_eigenvalues(E, n);
  std::sort(ev.begin(), ev.end());
  std::vector<double> out;
  for (double e : ev)
    if (out.empty() || std::fabs(e - out.back()) > 1e-4) out.push_back(e);
  return out;
}

// A = sum_c 2 f_c P_c, P_c = prod_{c' != c} (Chat - lambda_c') / (lambda_c - lambda_c').
std::vector<cplx> multi_superop(const RepInfo& r, const std::vector<double>& lambda, const std::vector<double>& f) {
  const int dd = r.d * r.d;
  const std::vector<cplx> C = casimir_superop(r);
  std::vector<cplx> A(size_t(dd) * dd, cplx(0, 0)), P(size_t(dd) * dd), Q(size_t(dd) * dd);
  for (size_t c = 0; c < lambda.size(); ++c) {
    if (f[c] == 0.0) continue;
    std::fill(P.begin(), P.end(), cplx(0, 0));
    for (int i = 0; i < dd; ++i) P[size_t(i) * dd + i] = 1.0;
    for (size_t cp = 0; cp < lambda.size(); ++cp) {
      if (cp == c) continue;
      const double inv = 1.0 / (lambda[c] - lambda[cp]);
      // Q = (C - lambda_cp) P * inv
      for (int i = 0; i < dd; ++i)
        for (int j = 0; j < dd; ++j) {
          cplx s = -lambda[cp] * P[size_t(i) * dd + j];
          for (int k = 0; k < dd; ++k) s += C[size_t(i) * dd + k] * P[size_t(k) * dd + j];
          Q[size_t(i) * dd + j] = s * inv;
        }
      std::swap(P, Q);
    }
    for (size_t k = 0; k < A.size(); ++k) A[k] += 2.0 * f[c] * P[k];
  }
  return A;
}

namespace {
// E_k of the orthonormal Hermitian basis as a d x d matrix.
std::vector<cplx> hermitian_unit(int d, int k) {
  std::vector<cplx> E(size_t(d) * d, cplx(0, 0));
  if (k < d) { E[size_t(k) * d + k] = 1.0; return E; }
  const double h = std::sqrt(0.5);
  int pr = (k - d) / 2;
  const bool imag = (k - d) % 2 == 1;
  int i = 0;
  while (pr >= d - 1 - i) { pr -= d - 1 - i; ++i; }
  const int j = i + 1 + pr;
  E[size_t(i) * d + j] = imag ? cplx(0, h) : cplx(h, 0);
  E[size_t(j) * d + i] = imag ? cplx(0, -h) : cplx(h, 0);
  return E;
}
}  // namespace

std::vector<double> hermitian_superop(int d, const std::vector<cplx>& A) {
  const int dd = d * d;
  std::vector<std::vector<cplx>> E(dd);
  for (int k = 0; k < dd; ++k) E[size_t(k)] = hermitian_unit(d, k);
  std::vector<double> AR(size_t(dd) * dd);
  std::vector<cplx> T(dd);
  for (int l = 0; l < dd; ++l) {
    for (int p = 0; p < dd; ++p) {  // T = A vec(E_l)
      cplx s(0, 0);
      for (int q = 0; q < dd; ++q) s += A[size_t(p) * dd + q] * E[size_t(l)][size_t(q)];
      T[size_t(p)] = s;
    }
    for (int k = 0; k < dd; ++k) {  // Tr(E_k T) = sum_ij (E_k)_ij T_ji
      cplx tr(0, 0);
      for (int i = 0; i < d; ++i)
        for (int j = 0; j < d; ++j) tr += E[size_t(k)][size_t(i) * d + j] * T[size_t(j) * d + i];
      AR[size_t(k) * dd + l] = tr.real();
    }
  }
  return AR;
}

}  // namespace ghx
