// Device representation policies. Each policy supplies, for a link U:
//   apply(U, phi, out)       out = D^(R)(U) phi
//   apply_dag(U, phi, out)   out = D^(R)(U)^dag phi
//   fwd_M(U, phy, phx, psi, M)
//       psi = D^(R)(U) phy and the N x N matrix M with
//       phx^dag T^a_R psi = M_SCALE Tr(T^a M) for every fundamental generator
//       T^a (M_SCALE an optional integer member, 1 when absent: a policy whose
//       M carries an exact factor leaves it to the caller's coefficient),
//   so the matter back-reaction g^a = -2 Im phx^dag T^a_R D(U) phy
//   (representation.hpp:36-46) is -2 Im Tr(T^a M), and the total link force is
//   one TA projection of (beta/N) U Sigma + 2 kappa M (theory_notes 3.3).
//
// Tensor irreps: phi lives in the orthonormal real basis W (d x N^n) of the
// Young-symmetrized subspace (rep_general.hpp:54-84); D(U) acts as U on each of
// the n tensor slots. Because the subspace is invariant, the pre-projection
// tensor Psi = U^{xn} W^T phy is exactly the lift of psi, so M is the sum over
// slots of the partial trace of Psi (x) conj(lift(phx)).
#pragma once
#include "device_math.hpp"

namespace ghx {

enum RepKind : int { REP_NONE = 0, REP_FUND = 1, REP_ADJ = 2, REP_TENSOR = 3, REP_EXT = 4 };

struct NoParams {};

// ---------------------------------------------------------------- none (pure gauge)
template <typename R, int N>
struct RepNone {
  static constexpr int DR = 0;
  static constexpr int KIND = REP_NONE;
  template <typename P> using Params = NoParams;
};

// ---------------------------------------------------------------- external matter
// Large irreps: the scalar force, pi/phi update and the matter link force are
// computed by big_matter_kernel; the step kernel only adds that link force
// (read from the G planes) to the gauge force.
template <typename R, int N>
struct RepExt {
  static constexpr int DR = 0;
  static constexpr int KIND = REP_EXT;
  template <typename P> using Params = NoParams;
};

// ---------------------------------------------------------------- fundamental
template <typename R, int N>
struct RepFund {
  static constexpr int DR = N;
  static constexpr int KIND = REP_FUND;
  template <typename P> using Params = NoParams;
  template <class PR>
  static GHX_DEV void apply(const Mat<R, N>& U, const cx<R>* x, cx<R>* y, const PR&) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
      cx<R> s = cmul(U.e[i][0], x[0]);
#pragma unroll
      for (int j = 1; j < N; ++j) cfma(s, U.e[i][j], x[j]);
      y[i] = s;
    }
  }
  template <class PR>
  static GHX_DEV void apply_dag(const Mat<R, N>& U, const cx<R>* x, cx<R>* y, const PR&) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
      cx<R> s = {R(0), R(0)};
#pragma unroll
      for (int j = 0; j < N; ++j) cfmac(s, U.e[j][i], x[j]);
      y[i] = s;
    }
  }
  // M = psi phx^dag
  template <class PR>
  static GHX_DEV void fwd_M(const Mat<R, N>& U, const cx<R>* py, const cx<R>* px, cx<R>* psi, Mat<R, N>& M, const PR& p) {
    apply(U, py, psi, p);
#pragma unroll
    for (int j = 0; j < N; ++j)
#pragma unroll
      for (int i = 0; i < N; ++i) { cx<R> z = {R(0), R(0)}; cfmcc(z, psi[j], px[i]); M.e[j][i] = z; }
  }
};

// ---------------------------------------------------------------- adjoint
// D_ab(U) = 2 Re Tr(T^a U T^b U^dag) (rep_adjoint.hpp:27-40) applied as the
// adjoint action on su(N): psi = components of U Phi U^dag, Phi = phi_b T^b.
// T^a_adj = -i f^{abc} gives phx^dag T^a psi = -i f^{abc} conj(phx_b) psi_c, so
// M = -2 [conj(Phx), Psi] (see the header comment for the normalisation).
template <typename R, int N>
struct RepAdj {
  static constexpr int DR = N * N - 1;
  static constexpr int KIND = REP_ADJ;
  template <typename P> using Params = NoParams;
  template <class PR>
  static GHX_DEV void apply(const Mat<R, N>& U, const cx<R>* x, cx<R>* y, const PR&) {
    const Mat<R, N> X = calg_to_mat<R, N>(x);
    const Mat<R, N> Y = mul_nd(mul_nn(U, X), U);
    tr_gen2<R, N>(Y, y);
  }
  template <class PR>
  static GHX_DEV void apply_dag(const Mat<R, N>& U, const cx<R>* x, cx<R>* y, const PR&) {
    const Mat<R, N> X = calg_to_mat<R, N>(x);
    const Mat<R, N> Y = mul_nn(mul_dn(U, X), U);
    tr_gen2<R, N>(Y, y);
  }
  template <class PR>
  static GHX_DEV void fwd_M(const Mat<R, N>& U, const cx<R>* py, const cx<R>* px, cx<R>* psi, Mat<R, N>& M, const PR&) {
    const Mat<R, N> Y = calg_to_mat<R, N>(py);
    const Mat<R, N> Psi = mul_nd(mul_nn(U, Y), U);
    tr_gen2<R, N>(Psi, psi);
    cx<R> pxc[DR];
#pragma unroll
    for (int b = 0; b < DR; ++b) pxc[b] = cconj(px[b]);
    const Mat<R, N> Xc = calg_to_mat<R, N>(pxc);
    const Mat<R, N> A = mul_nn(Xc, Psi), B = mul_nn(Psi, Xc);
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) M.e[i][j] = csub(A.e[i][j], B.e[i][j]);
  }
  static constexpr int M_SCALE = -2;
};

// ---------------------------------------------------------------- small tensor irreps
constexpr int ipow(int b, int e) { return e == 0 ? 1 : b * ipow(b, e - 1); }

template <typename P, int DR_, int DIMT_>
struct TensorParams {
  P W[DR_][DIMT_];  // orthonormal real basis rows (rep_general.hpp:40)
};

template <typename R, int N, int NB, int DR_>
struct RepTensor {
  static constexpr int DR = DR_;
  static constexpr int DIMT = ipow(N, NB);
  static constexpr int KIND = REP_TENSOR;
  template <typename P> using Params = TensorParams<P, DR_, DIMT>;

  template <class PR>
  static GHX_DEV void lift(const cx<R>* x, cx<R>* X, const PR& p) {
#pragma unroll
    for (int t = 0; t < DIMT; ++t) X[t] = {R(0), R(0)};
#pragma unroll
    for (int l = 0; l < DR; ++l)
#pragma unroll
      for (int t = 0; t < DIMT; ++t) {
        const R w = R(p.W[l][t]);
        X[t].r = fma(w, x[l].r, X[t].r);
        X[t].i = fma(w, x[l].i, X[t].i);
      }
  }
  template <class PR>
  static GHX_DEV void project(const cx<R>* X, cx<R>* y, const PR& p) {
#pragma unroll
    for (int k = 0; k < DR; ++k) {
      cx<R> s = {R(0), R(0)};
#pragma unroll
      for (int t = 0; t < DIMT; ++t) {
        const R w = R(p.W[k][t]);
        s.r = fma(w, X[t].r, s.r);
        s.i = fma(w, X[t].i, s.i);
      }
      y[k] = s;
    }
  }
  // dst = (M on slot s) src ; slot 0 is the most significant digit (rep_general.hpp:184-194)
  template <int S, bool DAG>
  static GHX_DEV void slot(const Mat<R, N>& M, const cx<R>* src, cx<R>* dst) {
    constexpr int blk = ipow(N, NB - 1 - S);
#pragma unroll
    for (int t = 0; t < DIMT; ++t) {
      if ((t / blk) % N != 0) continue;
#pragma unroll
      for (int j = 0; j < N; ++j) {
        cx<R> s = {R(0), R(0)};
#pragma unroll
        for (int k = 0; k < N; ++k) {
          if (DAG) cfmac(s, M.e[k][j], src[t + k * blk]);
          else cfma(s, M.e[j][k], src[t + k * blk]);
        }
        dst[t + j * blk] = s;
      }
    }
  }
  template <int S, bool DAG>
  static GHX_DEV void all_slots(const Mat<R, N>& M, cx<R>* a, cx<R>* b) {
    if constexpr (S < NB) {
      slot<S, DAG>(M, a, b);
      all_slots<S + 1, DAG>(M, b, a);
    }
  }
  // After all_slots the result sits in `a` if NB is even, else in `b`.
  template <bool DAG, class PR>
  static GHX_DEV void apply_impl(const Mat<R, N>& U, const cx<R>* x, cx<R>* y, const PR& p) {
    cx<R> a[DIMT], b[DIMT];
    lift(x, a, p);
    all_slots<0, DAG>(U, a, b);
    if constexpr (NB % 2 == 0) project(a, y, p); else project(b, y, p);
  }
  template <class PR> static GHX_DEV void apply(const Mat<R, N>& U, const cx<R>* x, cx<R>* y, const PR& p) { apply_impl<false>(U, x, y, p); }
  template <class PR> static GHX_DEV void apply_dag(const Mat<R, N>& U, const cx<R>* x, cx<R>* y, const PR& p) { apply_impl<true>(U, x, y, p); }

  // M_{ji} = sum_s sum_{t: t_s = i} Psi[t with slot s set to j] conj(Phx[t])
  template <int S>
  static GHX_DEV void partial_trace(const cx<R>* Psi, const cx<R>* Phx, Mat<R, N>& M) {
    if constexpr (S < NB) {
      constexpr int blk = ipow(N, NB - 1 - S);
#pragma unroll
      for (int t = 0; t < DIMT; ++t) {
        const int i = (t / blk) % N;
        const int base = t - i * blk;
#pragma unroll
        for (int j = 0; j < N; ++j) cfmcc(M.e[j][i], Psi[base + j * blk], Phx[t]);
      }
      partial_trace<S + 1>(Psi, Phx, M);
    }
  }
  template <class PR>
  static GHX_DEV void fwd_M(const Mat<R, N>& U, const cx<R>* py, const cx<R>* px, cx<R>* psi, Mat<R, N>& M, const PR& p) {
    cx<R> a[DIMT], b[DIMT];
    lift(py, a, p);
    all_slots<0, false>(U, a, b);
    cx<R> Phx[DIMT];
    lift(px, Phx, p);
    M = mzero<R, N>();
    if constexpr (NB % 2 == 0) { project(a, psi, p); partial_trace<0>(a, Phx, M); }
    else { project(b, psi, p); partial_trace<0>(b, Phx, M); }
  }
};

// ---------------------------------------------------------------- two-index symmetric {2}
// The single-row n=2 basis of rep_general.hpp:338-362 in odometer order over
// i <= j: W = e_i(x)e_i, or (e_i(x)e_j + e_j(x)e_i)/sqrt 2. With Phi the
// symmetric matrix Phi_ii = phi_k, Phi_ij = Phi_ji = phi_k/sqrt 2:
//   D(U) phi  <->  U Phi U^T,   D(U)^dag phi  <->  U^dag Phi conj(U),
// and the back-reaction matrix of both slots is M = 2 Psi Phx^dag.
// Selected by the host only after checking W has exactly this form.
template <typename R, int N>
struct RepSym2 {
  static constexpr int DR = N * (N + 1) / 2;
  static constexpr int KIND = REP_TENSOR;
  template <typename P> using Params = NoParams;
  static GHX_DEV Mat<R, N> lift(const cx<R>* x) {
    Mat<R, N> F;
    int k = 0;
    const R h = R(0.70710678118654752440);
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = i; j < N; ++j, ++k) {
        if (i == j) F.e[i][i] = x[k];
        else { F.e[i][j] = cscale(x[k], h); F.e[j][i] = F.e[i][j]; }
      }
    return F;
  }
  static GHX_DEV void project(const Mat<R, N>& S, cx<R>* y) {
    int k = 0;
    const R r2 = R(1.41421356237309504880);
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = i; j < N; ++j, ++k) y[k] = (i == j) ? S.e[i][i] : cscale(S.e[i][j], r2);
  }
  // S = A F B^T with F, S symmetric (upper triangle of S computed, then mirrored)
  template <bool DAG>
  static GHX_DEV Mat<R, N> sandwich(const Mat<R, N>& U, const Mat<R, N>& F) {
    Mat<R, N> G;  // G = U F  (or U^dag F)
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) {
        cx<R> z = {R(0), R(0)};
#pragma unroll
        for (int k = 0; k < N; ++k) {
          if (DAG) cfmac(z, U.e[k][i], F.e[k][j]);
          else cfma(z, U.e[i][k], F.e[k][j]);
        }
        G.e[i][j] = z;
      }
    Mat<R, N> S;  // S = G U^T (or G conj(U))
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = i; j < N; ++j) {
        cx<R> z = {R(0), R(0)};
#pragma unroll
        for (int k = 0; k < N; ++k) {
          if (DAG) cfmcc(z, G.e[i][k], U.e[k][j]);
          else cfma(z, G.e[i][k], U.e[j][k]);
        }
        S.e[i][j] = z;
        S.e[j][i] = z;
      }
    return S;
  }
  template <class PR>
  static GHX_DEV void apply(const Mat<R, N>& U, const cx<R>* x, cx<R>* y, const PR&) { project(sandwich<false>(U, lift(x)), y); }
  template <class PR>
  static GHX_DEV void apply_dag(const Mat<R, N>& U, const cx<R>* x, cx<R>* y, const PR&) { project(sandwich<true>(U, lift(x)), y); }
  template <class PR>
  static GHX_DEV void fwd_M(const Mat<R, N>& U, const cx<R>* py, const cx<R>* px, cx<R>* psi, Mat<R, N>& M, const PR&) {
    const Mat<R, N> S = sandwich<false>(U, lift(py));
    project(S, psi);
    const Mat<R, N> X = lift(px);
    M = mul_nd(S, X);  // Psi Phx^dag (Phx symmetric)
  }
  static constexpr int M_SCALE = 2;
};

}  // namespace ghx
