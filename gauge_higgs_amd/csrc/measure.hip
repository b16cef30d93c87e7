// Measurement kernels: Wilson loops W_R'(Ra, Rb) and Polyakov loops traced in
// a probe irrep R' (observables.hpp:36-101).
//
// The reference materialises the d x d matrix D^(R')(U_loop) and takes its
// trace. Here only the character is formed: chi_R'(W) = Tr D^(R')(W) depends
// on W through its eigenvalues alone, so
//   fundamental: Re Tr W;
//   adjoint:     |Tr W|^2 - |W|_F^2 / N  (= sum_a 2 Re Tr(T^a W T^a W^dag),
//                the trace of rep_adjoint.hpp's matrix, by completeness of T^a);
//   Young rows:  the Schur polynomial s_lambda of the eigenvalues, by the
//                Jacobi-Trudi determinant det[h_{lambda_i - i + j}] over the
//                complete symmetric polynomials h_k, which obey the
//                characteristic-polynomial recurrence
//                h_k = e_1 h_{k-1} - e_2 h_{k-2} + ... (+/-) e_N h_{k-N}
//                with e_i the elementary symmetric polynomials of W (e_1 = Tr W,
//                e_N = det W). This is the trace of W^{(x)n} on the
//                Young-symmetrised subspace for any matrix W.
// One thread per site, FP64 arithmetic whatever the link storage type, block
// partial sums reduced per replica in a fixed order.
#include "kernels.hpp"
#include "measure.hpp"

namespace ghx {

template <int N>
GHX_DEV cx<double> mdet(const Mat<double, N>& W) {
  if constexpr (N == 2) {
    cx<double> d = cmul(W.e[0][0], W.e[1][1]);
    const cx<double> o = cmul(W.e[0][1], W.e[1][0]);
    return csub(d, o);
  } else {
    static_assert(N == 3, "N = 2 or 3");
    cx<double> d = {0.0, 0.0};
    const cx<double> m0 = csub(cmul(W.e[1][1], W.e[2][2]), cmul(W.e[1][2], W.e[2][1]));
    const cx<double> m1 = csub(cmul(W.e[1][0], W.e[2][2]), cmul(W.e[1][2], W.e[2][0]));
    const cx<double> m2 = csub(cmul(W.e[1][0], W.e[2][1]), cmul(W.e[1][1], W.e[2][0]));
    cfma(d, W.e[0][0], m0);
    d = csub(d, cmul(W.e[0][1], m1));
    cfma(d, W.e[0][2], m2);
    return d;
  }
}

// Re chi_R'(W) for the probe p (see the file comment).
template <int N>
GHX_DEV double probe_character(const Mat<double, N>& W, const Probe& p) {
  cx<double> tr = W.e[0][0];
#pragma unroll
  for (int i = 1; i < N; ++i) tr = cadd(tr, W.e[i][i]);
  if (p.kind == 1) return tr.r;
  if (p.kind == 2) {
    double f2 = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) f2 += W.e[i][j].r * W.e[i][j].r + W.e[i][j].i * W.e[i][j].i;
    return tr.r * tr.r + tr.i * tr.i - f2 / N;
  }
  // elementary symmetric polynomials e1..eN of the eigenvalues
  const cx<double> e1 = tr, eN = mdet<N>(W);
  cx<double> e2 = eN;  // N = 2: e2 = det
  if constexpr (N == 3) {
    e2 = {0.0, 0.0};
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = i + 1; j < N; ++j) e2 = cadd(e2, csub(cmul(W.e[i][i], W.e[j][j]), cmul(W.e[i][j], W.e[j][i])));
  }
  // h_k by the recurrence, capturing the entries of the Jacobi-Trudi determinant
  const int r0 = p.rows[0], r1 = p.nrows > 1 ? p.rows[1] : 0;
  const int kmax = p.nrows > 1 ? r0 + 1 : r0;
  const cx<double> one = {1.0, 0.0}, zero = {0.0, 0.0};
  cx<double> hA = r0 == 0 ? one : zero;            // h_{r0}
  cx<double> hB = zero;                            // h_{r0+1}
  cx<double> hC = r1 == 0 ? one : zero;            // h_{r1}
  cx<double> hD = r1 - 1 == 0 ? one : zero;        // h_{r1-1} (0 when r1 = 0)
  cx<double> h1 = one, h2 = zero, h3 = zero;       // h_{k-1}, h_{k-2}, h_{k-3}
  for (int k = 1; k <= kmax; ++k) {
    cx<double> h = cmul(e1, h1);
    h = csub(h, cmul(e2, h2));
    if constexpr (N == 3) cfma(h, eN, h3);
    h3 = h2;
    h2 = h1;
    h1 = h;
    if (k == r0) hA = h;
    if (k == r0 + 1) hB = h;
    if (k == r1) hC = h;
    if (k == r1 - 1) hD = h;
  }
  if (p.nrows == 1) return hA.r;
  return cmul(hA, hC).r - cmul(hB, hD).r;
}

// One thread per site of its replica: sum over planes mu < nu of the probe
// character of the Ra x Rb loop based at the site (observables.hpp:38-48 path:
// Ra links along mu, Rb along nu, then back along mu and nu, daggered).
template <typename T, int N, int D>
__global__ __launch_bounds__(kRedBlock) void wilson_kernel(ActArgs a, const T* __restrict__ U, Probe p, int Ra, int Rb) {
  const int r = a.red_rep[blockIdx.x];
  const RepDesc& rd = a.reps[r];
  const unsigned s = a.red_first[blockIdx.x] - rd.offset + threadIdx.x;
  double q[1] = {0.0};
  if (s < rd.vol) {
#pragma unroll 1
    for (int mu = 0; mu < D; ++mu) {
#pragma unroll 1
      for (int nu = mu + 1; nu < D; ++nu) {
        const unsigned sm = rd.stride[mu], sn = rd.stride[nu];
        const int Lm = rd.L[mu], Ln = rd.L[nu];
        int xm = int((s / sm) % unsigned(Lm)), xn = int((s / sn) % unsigned(Ln));
        const unsigned base = rd.offset + s - unsigned(xm) * sm - unsigned(xn) * sn;
        auto site = [&](int cm, int cn) { return base + unsigned(cm) * sm + unsigned(cn) * sn; };
        Mat<double, N> W = meye<double, N>();
#pragma unroll 1
        for (int i = 0; i < Ra; ++i) {
          W = mul_nn(W, load_link<T, double, N>(U, a.Vp, mu, site(xm, xn)));
          xm = xm + 1 == Lm ? 0 : xm + 1;
        }
#pragma unroll 1
        for (int j = 0; j < Rb; ++j) {
          W = mul_nn(W, load_link<T, double, N>(U, a.Vp, nu, site(xm, xn)));
          xn = xn + 1 == Ln ? 0 : xn + 1;
        }
#pragma unroll 1
        for (int i = 0; i < Ra; ++i) {
          xm = xm == 0 ? Lm - 1 : xm - 1;
          W = mul_nd(W, load_link<T, double, N>(U, a.Vp, mu, site(xm, xn)));
        }
#pragma unroll 1
        for (int j = 0; j < Rb; ++j) {
          xn = xn == 0 ? Ln - 1 : xn - 1;
          W = mul_nd(W, load_link<T, double, N>(U, a.Vp, nu, site(xm, xn)));
        }
        q[0] += probe_character<N>(W, p);
      }
    }
  }
  block_reduce_store<1>(q, a.partials + blockIdx.x);
}

// One thread per transverse site (x_tdir = 0, the other coordinates from the
// thread's index in direction order): probe character of the Polyakov line.
template <typename T, int N, int D>
__global__ __launch_bounds__(kRedBlock) void polyakov_rep_kernel(ActArgs a, const T* __restrict__ U, Probe p, int tdir) {
  const int r = a.red_rep[blockIdx.x];
  const RepDesc& rd = a.reps[r];
  const unsigned t = a.red_first[blockIdx.x] - rd.offset + threadIdx.x;
  const unsigned Lt = unsigned(rd.L[tdir]);
  double q[1] = {0.0};
  if (t < rd.vol / Lt) {
    unsigned rem = t, s = 0;
#pragma unroll
    for (int m = 0; m < D; ++m) {
      if (m == tdir) continue;
      const unsigned Lm = unsigned(rd.L[m]);
      s += (rem % Lm) * rd.stride[m];
      rem /= Lm;
    }
    const unsigned st = rd.stride[tdir], g = rd.offset + s;
    Mat<double, N> Pl = load_link<T, double, N>(U, a.Vp, tdir, g);
#pragma unroll 1
    for (unsigned k = 1; k < Lt; ++k) Pl = mul_nn(Pl, load_link<T, double, N>(U, a.Vp, tdir, g + k * st));
    q[0] = probe_character<N>(Pl, p);
  }
  block_reduce_store<1>(q, a.partials + blockIdx.x);
}

template <typename T, int N, int D>
struct MeasLaunch {
  static void wilson(const LaunchCtx& c, const void* U, Probe p, int Ra, int Rb) {
    ActArgs a{c.reps, c.red_rep, c.red_first, c.Vp, c.partials};
    hipLaunchKernelGGL((wilson_kernel<T, N, D>), dim3(c.nred), dim3(kRedBlock), 0, c.stream, a, (const T*)U, p, Ra, Rb);
  }
  static void polyakov(const LaunchCtx& c, const void* U, Probe p, int tdir) {
    ActArgs a{c.reps, c.red_rep, c.red_first, c.Vp, c.partials};
    hipLaunchKernelGGL((polyakov_rep_kernel<T, N, D>), dim3(c.nred), dim3(kRedBlock), 0, c.stream, a, (const T*)U, p, tdir);
  }
  static MeasOps ops() { return MeasOps{&wilson, &polyakov}; }
};

template <int N>
static bool meas_pick(int prec, int D, MeasOps* o) {
  if (D == 4) { *o = prec ? MeasLaunch<float, N, 4>::ops() : MeasLaunch<double, N, 4>::ops(); return true; }
  if (D == 3) { *o = prec ? MeasLaunch<float, N, 3>::ops() : MeasLaunch<double, N, 3>::ops(); return true; }
  return false;
}

bool meas_ops(int prec, int N, int D, MeasOps* o) {
  if (N == 2) return meas_pick<2>(prec, D, o);
  if (N == 3) return meas_pick<3>(prec, D, o);
  return false;
}

}  // namespace ghx
