// Device side of the symmetric-power engine (symrep.hpp) for large tensor
// irreps: D(U) phi = invc Yc^T (Sym^k1(U) (x) Sym^k2(U)) Yc phi.
//
// sym_lift_kernel stores Yc phi for every site (dim S complex planes). In
// sym_link_kernel one wave takes one link: it builds Sym^k(U) (levels k =
// 1..k1 in LDS, the row-2 power saved on the way), applies it and its adjoint
// to the lifted scalars at both ends as S1 V S2^T (V the lifted vector viewed
// as s1 x s2) and projects back with Yc^T. sym_finish_kernel (one wave per
// site, a lane per component) turns a site's 2D hops into the matter link
// force, the scalar force and the scalar's update.
#pragma once
#include "bigrep.hpp"

namespace ghx {

constexpr int kSymWave = 64;
constexpr int kLiftChunk = 16;  // lifted components per block of sym_lift_kernel

// One complex value as a single 8- or 16-byte access (p at an even index).
template <typename T>
GHX_DEV cx<T> ld_cx(const T* p) {
  using V = std::conditional_t<sizeof(T) == 4, float2, double2>;
  const V v = *reinterpret_cast<const V*>(p);
  return {v.x, v.y};
}
template <typename T>
GHX_DEV void st_cx(T* p, cx<T> v) {
  using V = std::conditional_t<sizeof(T) == 4, float2, double2>;
  *reinterpret_cast<V*>(p) = V{v.r, v.i};
}

// Sym^k1(U) (normalised) into Ca; Sym^k2(U) into S2 when the irrep has a second
// row. Ca holds the levels of the parity of k1 (s1^2 entries), Cb the others;
// each level entry gathers its N sources of the previous level by recipe. The
// link goes through LDS (Us, N x N) so that its runtime column index j reads
// LDS rather than a private array.
template <typename R, int N>
GHX_DEV cx<R>* sym_build(const BigRepTables::Sym& st, const Mat<R, N>& U, cx<R>* Us, cx<R>* Ca, cx<R>* Cb, cx<R>* S2) {
  const int t = threadIdx.x;
  if (t == 0)
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) Us[i * N + j] = U.e[i][j];
  // level k lives in Ca when k1 - k is even (so the result lands in Ca)
  cx<R>* prev = (st.k1 % 2 == 0) ? Ca : Cb;
  cx<R>* cur = (st.k1 % 2 == 0) ? Cb : Ca;
  if (t == 0) prev[0] = {R(1), R(0)};
  __syncthreads();
  for (int k = 1; k <= st.k1; ++k) {
    const int r0 = st.rec_off[k], ne = st.rec_off[k + 1] - r0;
    for (int e = t; e < ne; e += kSymWave) {
      const int4 rc = st.recipe[r0 + e];
      cx<R> acc = {R(0), R(0)};
      if (rc.x >= 0) cfma(acc, Us[rc.w], prev[rc.x]);
      if (rc.y >= 0) cfma(acc, Us[N + rc.w], prev[rc.y]);
      if constexpr (N > 2)
        if (rc.z >= 0) cfma(acc, Us[2 * N + rc.w], prev[rc.z]);
      cur[e] = acc;
    }
    __syncthreads();
    if (k == st.k2)
      for (int e = t; e < ne; e += kSymWave)
        S2[e] = cscale(cur[e], sizeof(R) == 4 ? R(st.scale2f[e]) : R(st.scale2[e]));
    cx<R>* tmp = prev;
    prev = cur;
    cur = tmp;
  }
  const int ne = st.s1 * st.s1;
  for (int e = t; e < ne; e += kSymWave) prev[e] = cscale(prev[e], sizeof(R) == 4 ? R(st.scale1f[e]) : R(st.scale1[e]));
  __syncthreads();
  return prev;
}

// W = (S1 (x) S2) V = S1 V S2^T, or its adjoint S1^dag V conj(S2) (V, W: s1 x s2).
template <typename R>
GHX_DEV void sym_apply(const BigRepTables::Sym& st, const cx<R>* S1, const cx<R>* S2, const cx<R>* V, cx<R>* T1, cx<R>* W,
                       bool dag) {
  const int t = threadIdx.x, s1 = st.s1, s2 = st.s2;
  const int da = kSymWave / s2, db = kSymWave % s2;  // (a, b) of e advance by (da, db) per pass
  cx<R>* out1 = s2 == 1 ? W : T1;
  for (int e = t, a = t / s2, b = t % s2; e < s1 * s2; e += kSymWave) {
    cx<R> acc = {R(0), R(0)};
    if (!dag) {
      for (int a2 = 0; a2 < s1; ++a2) cfma(acc, S1[a * s1 + a2], V[a2 * s2 + b]);
    } else {
      for (int a2 = 0; a2 < s1; ++a2) cfmac(acc, S1[a2 * s1 + a], V[a2 * s2 + b]);
    }
    out1[e] = acc;
    a += da;
    b += db;
    if (b >= s2) { b -= s2; ++a; }
  }
  __syncthreads();
  if (s2 == 1) return;
  for (int e = t, a = t / s2, b = t % s2; e < s1 * s2; e += kSymWave, a += da, b += db) {
    if (b >= s2) { b -= s2; ++a; }
    cx<R> acc = {R(0), R(0)};
    if (!dag) {
      for (int b2 = 0; b2 < s2; ++b2) cfma(acc, S2[b * s2 + b2], T1[a * s2 + b2]);
    } else {
      for (int b2 = 0; b2 < s2; ++b2) cfmac(acc, S2[b2 * s2 + b], T1[a * s2 + b2]);
    }
    W[e] = acc;
  }
  __syncthreads();
}

// out[l] = invc sum_q Yc[q][l] W[q], l < d (threads over l).
template <typename R>
GHX_DEV void sym_project(const BigRepTables::Sym& st, int d, const cx<R>* W, cx<R>* out) {
  const R invc = R(st.invc);
  for (int l = threadIdx.x; l < d; l += kSymWave) {
    cx<R> acc = {R(0), R(0)};
#pragma unroll 4
    for (int q = 0; q < st.dimS; ++q) {
      const R y = sizeof(R) == 4 ? R(st.Ycf[q * d + l]) : R(st.Yc[q * d + l]);
      acc.r = fma(y, W[q].r, acc.r);
      acc.i = fma(y, W[q].i, acc.i);
    }
    out[l] = cscale(acc, invc);
  }
  __syncthreads();
}

// Yc phi of every site, site-major: lift[2 (g dimS + q) + (re, im)], so that
// the wave of a link reads one site's lifted vector in one contiguous run.
// Block (x, y) lifts components [y kLiftChunk, (y+1) kLiftChunk) of 64 sites:
// phi staged in LDS, the results through an LDS tile and out in runs of
// kLiftChunk complex values per site.
template <typename T>
__global__ __launch_bounds__(kSymWave) void sym_lift_kernel(unsigned Vp, const T* phi, int d, BigRepTables::Sym st) {
  using R = T;
  constexpr int ts = 2 * kLiftChunk + 1;  // tile row stride (odd: no bank conflicts)
  extern __shared__ unsigned char smem_raw[];
  cx<R>* ph = reinterpret_cast<cx<R>*>(smem_raw);  // [d][64]
  R* tile = reinterpret_cast<R*>(ph + d * kSymWave);  // [64][ts]
  const unsigned g0 = blockIdx.x * kSymWave;
  const int t = threadIdx.x, dimS = st.dimS, q0 = blockIdx.y * kLiftChunk;
  const int nq = min(kLiftChunk, dimS - q0);
  // padding sites are lifted too (never read): the block stays whole
  for (int l = 0; l < d; ++l) ph[l * kSymWave + t] = {R(phi[size_t(2 * l) * Vp + g0 + t]), R(phi[size_t(2 * l + 1) * Vp + g0 + t])};
  __syncthreads();
  for (int q = 0; q < nq; ++q) {
    cx<R> acc = {R(0), R(0)};
    for (int l = 0; l < d; ++l) {
      const R y = sizeof(R) == 4 ? R(st.Ycf[(q0 + q) * d + l]) : R(st.Yc[(q0 + q) * d + l]);
      acc.r = fma(y, ph[l * kSymWave + t].r, acc.r);
      acc.i = fma(y, ph[l * kSymWave + t].i, acc.i);
    }
    tile[t * ts + 2 * q] = acc.r;
    tile[t * ts + 2 * q + 1] = acc.i;
  }
  __syncthreads();
  T* out = static_cast<T*>(st.lift) + size_t(g0) * 2 * dimS + 2 * q0;
  const int w = 2 * nq;  // values per site in this chunk
  for (int e = t; e < kSymWave * w; e += kSymWave) {
    const int site = e / w, j = e - site * w;
    out[size_t(site) * 2 * dimS + j] = T(tile[site * ts + j]);
  }
}

// One link per wave (blockIdx.y = mu): Sym(U_mu(x)) is built once and serves
// both hops through the link, D(U) phi_{x+mu} (the forward hop of x, slot 2 mu)
// and D(U)^dag phi_x (the backward hop of x+mu, slot 2 mu + 1; skipped unless
// both). Results (d complex) go to the receiving site's hop record, site-major:
// hops[2 ((g 2D + slot) d + l) + (re, im)].
template <typename T, int N, int D>
__global__ __launch_bounds__(kSymWave) void sym_link_kernel(BigArgs<T> a, BigRepTables tb, T* out, int both) {
  using R = T;
  const BigRepTables::Sym& st = tb.sym;
  extern __shared__ unsigned char smem_raw[];
  cx<R>* L = reinterpret_cast<cx<R>*>(smem_raw);
  cx<R>* Ca = L + st.lds_Ca;
  cx<R>* Cb = L + st.lds_Cb;
  cx<R>* S2 = L + st.lds_S2;
  cx<R>* V = L + st.lds_V;
  cx<R>* T1 = L + st.lds_T1;
  cx<R>* W = L + st.lds_W;
  __shared__ cx<R> res[kBigMaxD];
  __shared__ cx<R> Us[N * N];
  const unsigned g = blockIdx.x, Vp = a.Vp;
  const int r = a.blk_rep[g / kBlock];
  const RepDesc& rd = a.reps[r];
  const unsigned s = g - rd.offset;
  if (s >= rd.vol) return;  // whole block exits together
  const int mu = int(blockIdx.y), t = threadIdx.x, d = tb.d;
  const Nbr<D> nb(rd, s);
  int up = nb.up[0];  // nb.up[mu] without a private array
#pragma unroll
  for (int m = 1; m < D; ++m) up = (mu == m) ? nb.up[m] : up;
  const unsigned gp = g + up;
  const Mat<R, N> Uk = load_link<T, R, N>(a.U, Vp, mu, g);
  const cx<R>* S1 = sym_build<R, N>(st, Uk, Us, Ca, Cb, S2);
  const T* lift = static_cast<const T*>(st.lift);
  for (int dir = 0; dir <= both; ++dir) {
    const unsigned gv = dir ? g : gp;   // the scalar's site
    const unsigned go = dir ? gp : g;   // the site that receives the hop
    const T* lv = lift + size_t(gv) * 2 * st.dimS;
    for (int q = t; q < st.dimS; q += kSymWave) V[q] = ld_cx(lv + 2 * q);
    __syncthreads();
    sym_apply<R>(st, S1, S2, V, T1, W, dir != 0);
    sym_project<R>(st, d, W, res);
    T* ho = out + 2 * ((size_t(go) * 2 * D + 2 * mu + dir) * d);
    for (int l = t; l < d; l += kSymWave) st_cx(ho + 2 * l, res[l]);
    __syncthreads();
  }
}

template <typename R>
GHX_DEV R wave_sum(R v) {
#pragma unroll
  for (int o = kSymWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// A site's matter terms from its 2D hops (sym_link_kernel), one wave per site
// with lane l holding component l (d <= 64): the matter link force
// -kappa g^a, g^a = -2 Im phi^dag T^a_R psi_mu (scalar_higgs.hpp:198-213), and
// the scalar force 2 kappa sum(hops) - 2 dV/dphi^* with the pi kick, the
// tangent projection and great-circle drift of a frozen scalar and the phi
// drift (big_matter_kernel's contract), or q1 = sum_mu Re phi^dag psi_mu for
// the action. Sums over components are wave reductions.
template <typename T, int N, int D>
__global__ __launch_bounds__(kSymWave) void sym_finish_kernel(BigArgs<T> a, BigRepTables tb, const T* __restrict__ hops) {
  using R = T;
  constexpr int NGN = N * N - 1;
  __shared__ cx<R> psi[kBigMaxD];
  const unsigned g = blockIdx.x, Vp = a.Vp;
  const int r = a.blk_rep[g / kBlock];
  const RepDesc& rd = a.reps[r];
  const unsigned s = g - rd.offset;
  if (s >= rd.vol) return;  // whole block exits together
  const int d = tb.d, l = threadIdx.x;
  const bool on = l < d;
  const T* hg = hops + 2 * (size_t(g) * 2 * D * d);
  auto hop = [&](int slot) { return on ? ld_cx(hg + 2 * (slot * d + l)) : cx<R>{R(0), R(0)}; };
  const cx<R> ph = on ? cx<R>{R(a.phi_in[size_t(2 * l) * Vp + g]), R(a.phi_in[size_t(2 * l + 1) * Vp + g])} : cx<R>{R(0), R(0)};
  const int nmu = (a.flags & F_NO_HOP) ? 0 : D;
  if (a.mode == BIG_ACTION) {
    double q = 0.0;
    for (int mu = 0; mu < nmu; ++mu) {
      const cx<R> ps = hop(2 * mu);
      q += double(ph.r) * double(ps.r) + double(ph.i) * double(ps.i);
    }
    q = wave_sum(q);
    if (l == 0) a.q1[g] = q;
    return;
  }
  const R k2 = R(2.0 * rd.kappa);
  cx<R> f = {R(0), R(0)};
  for (int mu = 0; mu < nmu; ++mu) {
    const cx<R> ps = hop(2 * mu), ch = hop(2 * mu + 1);
    psi[l] = ps;
    __syncthreads();
    // (T^a psi)_l from the CSR rows (a, l), then -kappa g^a summed over l
    R ga[NGN];
#pragma unroll
    for (int c = 0; c < NGN; ++c) {
      cx<R> tt = {R(0), R(0)};
      if (on)
        for (int e = tb.T_ptr[c * d + l]; e < tb.T_ptr[c * d + l + 1]; ++e)
          cfma(tt, cx<R>{R(tb.T_val[2 * e]), R(tb.T_val[2 * e + 1])}, psi[tb.T_col[e]]);
      cx<R> z = {R(0), R(0)};
      cfmac(z, ph, tt);
      ga[c] = wave_sum(z.i);
    }
    if (l < NGN) {
      R v = ga[0];
#pragma unroll
      for (int c = 1; c < NGN; ++c) v = (l == c) ? ga[c] : v;
      a.G[size_t(mu * NGN + l) * Vp + g] = T(R(-rd.kappa) * (R(-2) * v));
    }
    f.r += k2 * ps.r;
    f.i += k2 * ps.i;
    f.r += k2 * ch.r;
    f.i += k2 * ch.i;
    __syncthreads();
  }
  // on-site term -2 dV/dphi^* (scalar_higgs.hpp:111-125) or the plugged-in potential's planes
  const R n2 = wave_sum(ph.r * ph.r + ph.i * ph.i);
  if (!a.Fon) {
    const R c = (a.flags & F_NO_ONSITE) ? R(0) : R(-2) * (R(1) + R(2.0 * rd.lambda) * (n2 - R(1)));
    f.r = fma(c, ph.r, f.r);
    f.i = fma(c, ph.i, f.i);
  } else if (!(a.flags & F_NO_ONSITE) && on) {
    f.r += R(a.Fon[size_t(2 * l) * Vp + g]);
    f.i += R(a.Fon[size_t(2 * l + 1) * Vp + g]);
  }
  if (a.mode == BIG_FORCE) {
    if (on) {
      a.Fphi[size_t(2 * l) * Vp + g] = T(f.r);
      a.Fphi[size_t(2 * l + 1) * Vp + g] = T(f.i);
    }
    return;
  }
  const bool abs_eps = a.flags & F_ABS_EPS;
  const R ek = R(abs_eps ? a.ck : a.ck * rd.eps), ed = R(abs_eps ? a.cd : a.cd * rd.eps);
  const bool frozen = a.flags & F_FROZEN;
  cx<R> pv = on ? cx<R>{R(a.pi[size_t(2 * l) * Vp + g]), R(a.pi[size_t(2 * l + 1) * Vp + g])} : cx<R>{R(0), R(0)};
  if (a.flags & F_KICK) {
    pv.r = fma(ek, f.r, pv.r);
    pv.i = fma(ek, f.i, pv.i);
    if (frozen) {  // |phi| = 1 (gauge_higgs_hmc.hpp:61-104): tangent projection
      const R rp = wave_sum(ph.r * pv.r + ph.i * pv.i);
      pv.r -= rp * ph.r;
      pv.i -= rp * ph.i;
    }
  }
  cx<R> pn = ph;
  if (a.flags & F_DRIFT) {
    if (frozen) {  // great-circle drift of phi and pi
      const R rr = sqrt(wave_sum(pv.r * pv.r + pv.i * pv.i));
      const R rmin = sizeof(R) == 8 ? R(1e-300) : R(1e-30);
      if (rr >= rmin) {
        const R cs = cos(rr * ed), sn = sin(rr * ed), q = sn / rr, w = -rr * sn;
        pn = {cs * ph.r + q * pv.r, cs * ph.i + q * pv.i};
        pv = {w * ph.r + cs * pv.r, w * ph.i + cs * pv.i};
      }
    } else {
      pn = {fma(ed, pv.r, ph.r), fma(ed, pv.i, ph.i)};
    }
  }
  if (!on) return;
  if ((a.flags & F_KICK) || (frozen && (a.flags & F_DRIFT))) {
    a.pi[size_t(2 * l) * Vp + g] = T(pv.r);
    a.pi[size_t(2 * l + 1) * Vp + g] = T(pv.i);
  }
  if (a.flags & F_DRIFT) {
    a.phi_out[size_t(2 * l) * Vp + g] = T(pn.r);
    a.phi_out[size_t(2 * l + 1) * Vp + g] = T(pn.i);
  }
}

}  // namespace ghx
