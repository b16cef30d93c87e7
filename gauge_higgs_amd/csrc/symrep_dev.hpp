// Device side of the symmetric-power engine (symrep.hpp) for large tensor
// irreps: D(U) phi = Yc^T (Sym^k1(U) (x) Sym^k2(U)) Yc phi / c.
//
// sym_lift_kernel stores Yc phi for every site. In sym_link_kernel one wave
// takes one link: it runs the host-built LDS program that forms the
// unnormalised C_k1 (and C_k2) of the link, applies C1 (x) C2 and its adjoint
// to the lifted scalars at both ends as C1 V C2^T (V viewed as s1 x s2, scaled
// by F^-1 or F on load), and projects both hops back with the sparse rows of
// Yc F / c and Yc F^-1 / c. sym_finish_kernel (one wave per site, a lane per
// component) turns a site's hops into the matter link force, the scalar force
// and the scalar's update.
#pragma once
#include "bigrep.hpp"
#include "symrep.hpp"

namespace ghx {

constexpr int kSymWave = 64;
constexpr int kFinishSites = 8;  // sites (waves) per workgroup of sym_finish_kernel (16: +3.5 %, 4: +3.7 % per {6,1} step)
constexpr int kLiftChunk = 16;  // lifted components per block of sym_lift_kernel

// The link kernel runs one wave per workgroup. LDS instructions of a wave
// complete in issue order, so ordering one step's LDS writes before the next
// step's reads needs only a wave-scope fence (no s_barrier, and no wait for the
// global load prefetched for the next step).
GHX_DEV void wave_sync() { wave_fence(); }

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

// C_k1 (and C_k2) of the link U by the LDS program: one entry per lane per
// pass, L[dst] = sum_i Us[i][j] L[s_i]. Slot 0 is zero, slot 1 takes idle
// lanes' writes, Us holds U with a fourth column e_0 (j = 3 copies).
template <typename R, int N>
GHX_DEV void sym_build(const BigRepTables::Sym& st, const Mat<R, N>& U, cx<R>* L) {
  const int t = threadIdx.x;
  if (t == 0) {
    L[0] = {R(0), R(0)};
    L[st.lds_one] = {R(1), R(0)};
#pragma unroll
    for (int i = 0; i < N; ++i) {
#pragma unroll
      for (int j = 0; j < N; ++j) L[kSymLdsUs + 4 * i + j] = U.e[i][j];
      L[kSymLdsUs + 4 * i + 3] = {R(i == 0 ? 1 : 0), R(0)};
    }
  }
  wave_sync();
  const int4* pg = st.prog + t;
  const int last = st.npass - 1;
  unsigned char* Lb = reinterpret_cast<unsigned char*>(L);  // the program's offsets are in bytes
  auto at = [&](int off) -> cx<R>& { return *reinterpret_cast<cx<R>*>(Lb + off); };
  auto run = [&](const int4 rc) {
    const cx<R>* u = L + kSymLdsUs + (rc.w & 3);
    cx<R> acc = cmul(u[0], at(rc.x));
    cfma(acc, u[4], at(rc.y));
    if constexpr (N > 2) cfma(acc, u[8], at(rc.z));
    at(rc.w & ~3) = acc;
    wave_sync();
  };
  // two register sets, each refilled right after use: a pass's entries load
  // while the other set's pass runs
  int4 X = pg[0], Y = pg[min(1, last) * kSymWave];
  for (int p = 0; p < st.npass; p += 2) {  // npass is even (the host pads it)
    run(X);
    X = pg[min(p + 2, last) * kSymWave];
    run(Y);
    Y = pg[min(p + 3, last) * kSymWave];
  }
}

// Both hops of a link in one sweep. Direction 0 (the forward hop) needs
// (C1 (x) C2) V0 = C1 V0 C2^T, direction 1 (the adjoint) C1^dag V1 conj(C2).
// Direction 1 is carried conjugated (V1, T1 and W hold the conjugates of the
// true values), so that both are plain products that differ only in how a
// lane walks C1 and C2:
//   conj(C1^dag V1) = C1^T conj(V1),   conj(T conj(C2)) = conj(T) C2,
// and the projection conjugates its result back (Yp is real). Regions V, T1
// and W hold the two directions P = 4 ng entries apart (V, W: s1 x s2).
// With s2 <= 4 a lane takes one row a of one direction: the s2 entries of
// C1 V (or C1^T V) for that row stay in registers, so C1 is read once per lane
// and V's row a2 is the same for all lanes of a direction (an LDS broadcast),
// and C2 is applied in registers before W is written. Larger s2 use the
// two-stage form below (a lane per entry, C1 V through LDS).
// W is separate from V: lanes of a later pass may still read V.
// t and nl: this lane's index among the nl lanes that serve the link.
template <typename R>
GHX_DEV void sym_apply2_rows(const BigRepTables::Sym& st, const cx<R>* S1, const cx<R>* S2, const cx<R>* V, cx<R>* W, int t, int nl) {
  const int s1 = st.s1, s2 = st.s2, P = 4 * st.ng, s2p = st.s2p;
  for (int e = t; e < 2 * s1; e += nl) {
    const int dir = e >= s1, a = e - dir * s1;
    const cx<R>* c = S1 + (dir ? a : a * st.s1p);  // row a of C1, or column a
    const int cs = dir ? st.s1p : 1;
    const cx<R>* v = V + dir * P;
    cx<R> acc[kSymRowS2];
#pragma unroll
    for (int b = 0; b < kSymRowS2; ++b) acc[b] = {R(0), R(0)};
#pragma unroll 4
    for (int a2 = 0; a2 < s1; ++a2, c += cs, v += s2) {
      const cx<R> cv = *c;
#pragma unroll
      for (int b = 0; b < kSymRowS2; ++b)
        if (b < s2) cfma(acc[b], cv, v[b]);
    }
    cx<R>* w = W + dir * P + a * s2;
    if (s2 == 1) { w[0] = acc[0]; continue; }
#pragma unroll
    for (int b = 0; b < kSymRowS2; ++b) {
      if (b >= s2) break;
      cx<R> z = {R(0), R(0)};
#pragma unroll
      for (int b2 = 0; b2 < kSymRowS2; ++b2)
        if (b2 < s2) cfma(z, dir ? S2[b2 * s2p + b] : S2[b * s2p + b2], acc[b2]);
      w[b] = z;
    }
  }
  wave_sync();
}

// Returns where the result lies: W, or T1 on the row path.
template <typename R>
GHX_DEV const cx<R>* sym_apply2(const BigRepTables::Sym& st, const cx<R>* S1, const cx<R>* S2, const cx<R>* V, cx<R>* T1, cx<R>* W,
                                int t, int nl) {
  const int s1 = st.s1, s2 = st.s2, dimS = st.dimS, P = 4 * st.ng, s1p = st.s1p, s2p = st.s2p;
  if (s2 <= kSymRowS2) {
    sym_apply2_rows<R>(st, S1, S2, V, T1, t, nl);
    return T1;
  }
  cx<R>* out1 = s2 == 1 ? W : T1;
  for (int e = t; e < 2 * dimS; e += nl) {
    const int dir = e >= dimS, q = e - dir * dimS, a = q / s2, b = q - a * s2;
    const cx<R>* c = S1 + (dir ? a : a * s1p);  // row a of C1, or column a
    const int cs = dir ? s1p : 1;
    const cx<R>* v = V + dir * P + b;
    cx<R> acc = {R(0), R(0)};
#pragma unroll 4
    for (int a2 = 0; a2 < s1; ++a2, c += cs, v += s2) cfma(acc, *c, *v);
    out1[dir * P + q] = acc;
  }
  wave_sync();
  if (s2 == 1) return W;
  for (int e = t; e < 2 * dimS; e += nl) {
    const int dir = e >= dimS, q = e - dir * dimS, a = q / s2, b = q - a * s2;
    const cx<R>* c = S2 + (dir ? b : b * s2p);  // row b of C2, or column b
    const int cs = dir ? s2p : 1;
    const cx<R>* u = T1 + dir * P + a * s2;
    cx<R> acc = {R(0), R(0)};
    for (int b2 = 0; b2 < s2; ++b2, c += cs) cfma(acc, *c, u[b2]);
    W[dir * P + q] = acc;
  }
  wave_sync();
  return W;
}

// A hop projected back to the irrep, out[l] = sum_k ysv[dir][k][l] W[q_k] with
// q_k = ysq[dir][k][l] (lanes over l; 1/c and F or F^-1 folded into the
// values), stored straight to the hop record. The rows hold a few nonzeros
// each (the q that split the colour content of component l).
template <typename T, typename R>
GHX_DEV void sym_project(const BigRepTables::Sym& st, int dir, int d, const cx<R>* W, T* out, int t, int nl) {
  const int nz = st.ynz;
  const int* __restrict__ qs = st.ysq + size_t(dir) * nz * d;
  const R* __restrict__ vs =
      (sizeof(R) == 4 ? reinterpret_cast<const R*>(st.ysvf) : reinterpret_cast<const R*>(st.ysv)) + size_t(dir) * nz * d;
  for (int l = t; l < d; l += nl) {
    cx<R> acc = {R(0), R(0)};
    for (int k = 0; k < nz; ++k) {
      const R v = vs[k * d + l];
      const cx<R> w = W[qs[k * d + l]];
      acc.r = fma(v, w.r, acc.r);
      acc.i = fma(v, w.i, acc.i);
    }
    st_cx(out + 2 * l, cx<T>{T(acc.r), T(acc.i)});
  }
}

// Yc phi of every site, site-major: lift[2 (g dimS + q) + (re, im)], so that
// the wave of a link reads one site's lifted vector in one contiguous run.
// Block (x, y) lifts components [y kLiftChunk, (y+1) kLiftChunk) of 64 sites,
// a lane per site with the chunk's accumulators in registers: phi_l is read
// once per lane (coalesced over sites), the table row Yl[l][chunk] is uniform
// (scalar loads). The results go out through an LDS tile in runs of
// kLiftChunk complex values per site.
template <typename T>
__global__ __launch_bounds__(kSymWave) void sym_lift_kernel(unsigned Vp, const T* __restrict__ phi, int d, BigRepTables::Sym st) {
  using R = T;
  constexpr int ts = 2 * kLiftChunk + 1;  // tile row stride (odd: no bank conflicts)
  __shared__ R tile[kSymWave * ts];
  const unsigned g0 = blockIdx.x * kSymWave;
  const int t = threadIdx.x, dimS = st.dimS, q0 = blockIdx.y * kLiftChunk;
  const int nq = min(kLiftChunk, dimS - q0);
  const R* __restrict__ yl = (sizeof(R) == 4 ? reinterpret_cast<const R*>(st.Ylf) : reinterpret_cast<const R*>(st.Yl)) + q0;
  cx<R> acc[kLiftChunk];
#pragma unroll
  for (int q = 0; q < kLiftChunk; ++q) acc[q] = {R(0), R(0)};
  // padding sites are lifted too (never read): the block stays whole
  for (int l = 0; l < d; ++l, yl += st.dimSp) {
    const cx<R> p = {R(phi[size_t(2 * l) * Vp + g0 + t]), R(phi[size_t(2 * l + 1) * Vp + g0 + t])};
#pragma unroll
    for (int q = 0; q < kLiftChunk; ++q) {  // the table's zero padding covers q >= nq
      acc[q].r = fma(yl[q], p.r, acc[q].r);
      acc[q].i = fma(yl[q], p.i, acc[q].i);
    }
  }
#pragma unroll
  for (int q = 0; q < kLiftChunk; ++q) {
    tile[t * ts + 2 * q] = acc[q].r;
    tile[t * ts + 2 * q + 1] = acc[q].i;
  }
  __syncthreads();
  T* out = static_cast<T*>(st.lift) + size_t(g0) * 2 * dimS + 2 * q0;
  const int w = 2 * nq;  // values per site in this chunk
  for (int e = t; e < kSymWave * w; e += kSymWave) {
    const int site = e / w, j = e - site * w;
    out[size_t(site) * 2 * dimS + j] = T(tile[site * ts + j]);
  }
}

// FP32: the same lift on the matrix cores. Over all sites it is one GEMM,
// lift[site][q] = sum_l phi_l(site) Yc[q][l]: A = the scalars (M = two rows per
// site, re and im; K = d in steps of 4, zero beyond d), B = Yc^T (N = dim S in
// 16-column tiles, the table's zero padding), v_mfma_f32_16x16x4f32 with lane
// t holding A[t % 16][t / 16] and B[t / 16][t % 16] and receiving
// D[4 (t / 16) + r][t % 16] in accumulator r. Accumulators r = 0, 1 (and 2, 3)
// are the re and im of one site at one q, so a lane stores them as one pair
// and 16 lanes store 16 consecutive q of a site. A wave takes 8 sites (one
// 16-row tile), four waves a workgroup of 32 sites.
constexpr int kLiftMfmaWaves = 4;
constexpr int kLiftMfmaSites = 8;  // per wave
typedef float mfma_f4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(kSymWave* kLiftMfmaWaves) void sym_lift_mfma_kernel(unsigned Vp, const float* __restrict__ phi, int d,
                                                                                  BigRepTables::Sym st) {
  constexpr int kMaxK = 16;                     // d <= 64
  constexpr int kTiles = kLiftMfmaSites / 8;  // 16-row tiles per wave
  const int lane = int(threadIdx.x) % kSymWave, wave = int(threadIdx.x) / kSymWave;
  const int col = lane % 16, kk = lane / 16;
  const unsigned site0 = (blockIdx.x * kLiftMfmaWaves + unsigned(wave)) * kLiftMfmaSites;
  const int dimS = st.dimS, dimSp = st.dimSp, ks = (d + 3) / 4, nt = dimSp / 16;
  const float* __restrict__ yl = st.Ylf;
  float* lift = static_cast<float*>(st.lift);
  float a[kTiles][kMaxK];  // A rows: site site0 + 8 m + col / 2, re (col even) or im
#pragma unroll
  for (int m = 0; m < kTiles; ++m)
#pragma unroll
    for (int s2 = 0; s2 < kMaxK; ++s2) {
      const int l = 4 * s2 + kk;
      a[m][s2] = (s2 < ks && l < d) ? phi[size_t(2 * l + (col & 1)) * Vp + site0 + unsigned(8 * m + col / 2)] : 0.f;
    }
#pragma unroll 2
  for (int t = 0; t < nt; ++t) {
    float b[kMaxK];
#pragma unroll
    for (int s2 = 0; s2 < kMaxK; ++s2) {
      const int l = 4 * s2 + kk;
      b[s2] = (s2 < ks && l < d) ? yl[size_t(l) * dimSp + 16 * t + col] : 0.f;
    }
    const int q = 16 * t + col;
#pragma unroll
    for (int m = 0; m < kTiles; ++m) {
      mfma_f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s2 = 0; s2 < kMaxK; ++s2) {
        if (s2 >= ks) break;
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[m][s2], b[s2], acc, 0, 0, 0);
      }
      if (q < dimS) {  // rows 4 kk + r: sites 8 m + 2 kk (r = 0, 1) and 8 m + 2 kk + 1 (r = 2, 3)
        const unsigned sa = site0 + unsigned(8 * m + 2 * kk);
        *reinterpret_cast<float2*>(lift + (size_t(sa) * dimS + q) * 2) = float2{acc[0], acc[1]};
        *reinterpret_cast<float2*>(lift + (size_t(sa + 1) * dimS + q) * 2) = float2{acc[2], acc[3]};
      }
    }
  }
}

// One link per wave (blockIdx.y = mu): Sym(U_mu(x)) is built once and serves
// both hops through the link: D(U) phi_{x+mu}, the forward hop of x, into
// hops[2 ((g D + mu) d + l) + (re, im)]; and D(U)^dag phi_x, the backward hop
// of x+mu (skipped unless both), conjugated, into
// hopsL[2 ((g' D + mu) d + l) + (re, im)] with g' = x+mu.
template <typename T, int N, int D>
__global__ __launch_bounds__(kSymWave) void sym_link_kernel(BigArgs<T> a, BigRepTables tb, T* out, int both) {
  using R = T;
  const BigRepTables::Sym& st = tb.sym;
  extern __shared__ unsigned char smem_raw[];
  cx<R>* L = reinterpret_cast<cx<R>*>(smem_raw);
  cx<R>* V = L + st.lds_V;
  cx<R>* T1 = L + st.lds_T1;
  cx<R>* W = L + st.lds_W;
  const unsigned g = blockIdx.x, Vp = a.Vp;
  const int r = a.blk_rep[g / kBlock];
  const RepDesc& rd = a.reps[r];
  const unsigned s = g - rd.offset;
  if (s >= rd.vol) return;  // whole block exits together
  const int mu = int(blockIdx.y), t = threadIdx.x, d = tb.d;
  const Nbr<D> nb(rd, s);
  const unsigned gp = g + nb.upv(rd, mu);
  sym_build<R, N>(st, load_link<T, R, N>(a.U, Vp, mu, g), L);
  const cx<R>* S1 = L + st.lds_S1;
  const cx<R>* S2 = L + st.lds_S2;
  const int P = 4 * st.ng, dimS = st.dimS, nd = both + 1;
  // V0 = F^-1 Yc phi(x+mu) for the forward hop of x, V1 = conj(F Yc phi(x)) for
  // the adjoint hop of x+mu
  const T* lift = static_cast<const T*>(st.lift);
  const R* pre = sizeof(R) == 4 ? reinterpret_cast<const R*>(st.pref) : reinterpret_cast<const R*>(st.pre);
  for (int e = t; e < nd * dimS; e += kSymWave) {
    const int dir = e >= dimS, q = e - dir * dimS;
    const cx<R> v = cscale(ld_cx(lift + size_t(dir ? g : gp) * 2 * dimS + 2 * q), pre[e]);
    V[dir * P + q] = dir ? cconj(v) : v;
  }
  wave_sync();
  const cx<R>* Wr = sym_apply2<R>(st, S1, S2, V, T1, W, t, kSymWave);
  sym_project<T, R>(st, 0, d, Wr, out + 2 * ((size_t(g) * D + mu) * d), t, kSymWave);
  if (both) sym_project<T, R>(st, 1, d, Wr + P, static_cast<T*>(st.hopsL) + 2 * ((size_t(gp) * D + mu) * d), t, kSymWave);
}

// ---------------------------------------------------------------------------
// SU(3) with dim Sym^k1 <= 32 (k1 <= 6): two links per wave, C_k built a column
// per lane. Lanes 0-31 take link mu = 2 blockIdx.y of site blockIdx.x, lanes
// 32-63 link mu + 1, each half in its own LDS region laid out as for the
// per-entry program. Column n of level k follows from column n - e_j of level
// k - 1 (j = the first colour of n):
//   C_k[m][n] = sum_i U[i][j] C_(k-1)[m - e_i][n - e_j],
// so a lane reads that one column of the previous level into registers,
// forms every entry of its column from them with the lane's fixed U[i][j],
// and writes it: the build moves one read of level k-1 and one write of level
// k through LDS, against three reads of C_(k-1), three of U and one write per
// entry in the program. Level 1 is U itself and stays in registers.
// Monomials of degree k over the three colours are ordered as on the host:
// first exponent descending, then the second.
namespace symc {
__host__ __device__ constexpr int sdim(int k) { return (k + 1) * (k + 2) / 2; }
__host__ __device__ constexpr int tri(int r) { return r * (r + 1) / 2; }
// the index of (a0, a1, k - a0 - a1) among the monomials of degree k
__host__ __device__ constexpr int midx(int k, int a0, int a1) { return tri(k - a0) + (k - a0 - a1); }
struct Mono {
  int a[3];
};
// the exponents of monomial m of degree k
__host__ __device__ constexpr Mono mono(int k, int m) {
  int r = 0;
  while (tri(r + 1) <= m) ++r;
  const int a0 = k - r, a1 = r - (m - tri(r));
  return Mono{{a0, a1, k - a0 - a1}};
}
// the row of level k - 1 that row m of level k draws on through colour i, or -1
__host__ __device__ constexpr int rowsrc(int k, int m, int i) {
  Mono x = mono(k, m);
  if (x.a[i] == 0) return -1;
  x.a[i] -= 1;
  return midx(k - 1, x.a[0], x.a[1]);
}
}  // namespace symc

constexpr int kSymPairLanes = kSymWave / 2;  // dim Sym^6 = 28 columns fit a half-wave; Sym^7 (36) does not

template <int B, int E, class F>
GHX_DEV void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// row[j] for a per-lane j in {0, 1, 2} as a one-hot weighted sum (exact for
// finite entries). A chain of selects would be folded into an indexed load,
// which puts the whole matrix in scratch memory.
template <typename R>
GHX_DEV cx<R> sel3(const cx<R>* row, int j) {
  const R w0 = R(j == 0), w1 = R(j == 1), w2 = R(j == 2);
  return {fma(w0, row[0].r, fma(w1, row[1].r, w2 * row[2].r)), fma(w0, row[0].i, fma(w1, row[1].i, w2 * row[2].i))};
}

// Level K (>= 2) of the build for column n = t of this lane, written with row
// stride SC into Lc, and copied into S2c (row stride s2p) when that is set.
template <typename R, int K, int SC>
GHX_DEV void sym_level(const Mat<R, 3>& U, const cx<R>* Lp, cx<R>* Lc, cx<R>* S2c, int s2p, int t) {
  constexpr int SK = symc::sdim(K), SP = symc::sdim(K - 1);
  if (t < SK) {
    const symc::Mono x = symc::mono(K, t);
    const int j = x.a[0] > 0 ? 0 : (x.a[1] > 0 ? 1 : 2);
    const int np = j == 0 ? symc::midx(K - 1, x.a[0] - 1, x.a[1]) : (j == 1 ? symc::midx(K - 1, 0, x.a[1] - 1) : symc::midx(K - 1, 0, 0));
    cx<R> col[SP];
    if constexpr (K == 2) {
#pragma unroll
      for (int i = 0; i < 3; ++i) col[i] = sel3(U.e[i], np);
    } else {
      static_for<0, SP>([&](auto mp) { col[int(mp)] = Lp[int(mp) * SP + np]; });
    }
    cx<R> u[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) u[i] = sel3(U.e[i], j);
    static_for<0, SK>([&](auto mc) {
      constexpr int m = mc;
      constexpr int r0 = symc::rowsrc(K, m, 0), r1 = symc::rowsrc(K, m, 1), r2 = symc::rowsrc(K, m, 2);
      cx<R> acc = {R(0), R(0)};
      if constexpr (r0 >= 0) cfma(acc, u[0], col[r0]);
      if constexpr (r1 >= 0) cfma(acc, u[1], col[r1]);
      if constexpr (r2 >= 0) cfma(acc, u[2], col[r2]);
      Lc[m * SC + t] = acc;
      if (S2c) S2c[m * s2p + t] = acc;
    });
  }
  wave_sync();
}

// The row-blocked apply for a half-wave (s2 <= kSymRowS2, s1 <= 32): lane a
// takes row a of C1 for the forward hop and column a for the adjoint at once,
// two independent accumulations that share each step's loop and hide each
// other's LDS latency. Same arithmetic as sym_apply2_rows, with the row length
// S1 = dim Sym^k1 known at compile time; the result lands in W (both
// directions, P apart). W may be V itself: every lane's reads
// of V precede its writes in the wave's instruction stream.
template <typename R, int S1N>
GHX_DEV void sym_apply_pair_rows(const BigRepTables::Sym& st, const cx<R>* S1, const cx<R>* S2, const cx<R>* V, cx<R>* W, int t) {
  constexpr int s1 = S1N, s1p = S1N | 1;
  const int s2 = st.s2, P = 4 * st.ng, s2p = st.s2p;
  if (t < s1) {
    const int a = t;
    const cx<R>* c0 = S1 + a * s1p;  // row a (forward)
    const cx<R>* c1 = S1 + a;        // column a (adjoint)
    const cx<R>* v0 = V;
    const cx<R>* v1 = V + P;
    cx<R> f[kSymRowS2], b[kSymRowS2];
#pragma unroll
    for (int k = 0; k < kSymRowS2; ++k) f[k] = b[k] = {R(0), R(0)};
#pragma unroll 7
    for (int a2 = 0; a2 < s1; ++a2, v0 += s2, v1 += s2) {
      const cx<R> x0 = c0[a2], x1 = c1[a2 * s1p];
#pragma unroll
      for (int k = 0; k < kSymRowS2; ++k)
        if (k < s2) {
          cfma(f[k], x0, v0[k]);
          cfma(b[k], x1, v1[k]);
        }
    }
    cx<R>* w0 = W + a * s2;
    cx<R>* w1 = W + P + a * s2;
    if (s2 == 1) {
      w0[0] = f[0];
      w1[0] = b[0];
    } else {
#pragma unroll
      for (int k = 0; k < kSymRowS2; ++k) {
        if (k >= s2) break;
        cx<R> z0 = {R(0), R(0)}, z1 = {R(0), R(0)};
#pragma unroll
        for (int k2 = 0; k2 < kSymRowS2; ++k2)
          if (k2 < s2) {
            cfma(z0, S2[k * s2p + k2], f[k2]);
            cfma(z1, S2[k2 * s2p + k], b[k2]);
          }
        w0[k] = z0;
        w1[k] = z1;
      }
    }
  }
  wave_sync();
}

template <typename T, int D, int K1>
__global__ __launch_bounds__(kSymWave) void sym_link_pair_kernel(BigArgs<T> a, BigRepTables tb, T* out, int both) {
  using R = T;
  constexpr int N = 3, NL = kSymPairLanes;
  constexpr int kPre = 2 * kSymMaxDimS / NL;  // lifted entries a lane loads (both directions)
  const BigRepTables::Sym& st = tb.sym;
  extern __shared__ unsigned char smem_raw[];
  const int lane = threadIdx.x, h = lane / NL, t = lane % NL;
  cx<R>* L = reinterpret_cast<cx<R>*>(smem_raw) + h * st.pair_total;
  const unsigned g = blockIdx.x, Vp = a.Vp;
  const int r = a.blk_rep[g / kBlock];
  const RepDesc& rd = a.reps[r];
  const unsigned s = g - rd.offset;
  if (s >= rd.vol) return;  // whole block exits together
  const int mu = 2 * int(blockIdx.y) + h;
  const bool act = mu < D;
  const int ml = act ? mu : 0;  // an idle half (odd D) repeats link 0 and stores nothing
  const Nbr<D> nb(rd, s);
  const unsigned gp = g + nb.upv(rd, ml);
  const Mat<R, N> U = load_link<T, R, N>(a.U, Vp, ml, g);
  const int P = 4 * st.ng, dimS = st.dimS, nd = both + 1, k2 = st.k2;
  // the lifted scalars of both ends, loaded now so that the build hides their latency
  const T* lift = static_cast<const T*>(st.lift);
  cx<R> pv[kPre];
#pragma unroll
  for (int c = 0; c < kPre; ++c) {
    const int e = t + c * NL;
    if (e < nd * dimS) {
      const int dir = e >= dimS, q = e - dir * dimS;
      pv[c] = ld_cx(lift + size_t(dir ? g : gp) * 2 * dimS + 2 * q);
    }
  }
  // every level in place in C: a level's reads all precede its writes in the
  // wave's instruction stream, and a wave's LDS operations execute in order
  cx<R>* C = L;
  cx<R>* S2 = L + st.pair_S2;
  const bool copy2 = k2 > 0 && k2 < K1;
  if (copy2 && k2 == 1 && t < N) {
#pragma unroll
    for (int i = 0; i < N; ++i) S2[i * st.s2p + t] = sel3(U.e[i], t);
  }
  static_for<2, K1 + 1>([&](auto kc) {
    constexpr int k = kc;
    sym_level<R, k, (k == K1 ? (symc::sdim(k) | 1) : symc::sdim(k))>(U, C, C, (copy2 && k == k2) ? S2 : nullptr, st.s2p, t);
  });
  cx<R>* V = L + st.pair_V;
  cx<R>* T1 = L + st.pair_T1;
  cx<R>* W = V;
  // V0 = F^-1 Yc phi(x+mu) for the forward hop of x, V1 = conj(F Yc phi(x)) for
  // the adjoint hop of x+mu
  const R* pre = sizeof(R) == 4 ? reinterpret_cast<const R*>(st.pref) : reinterpret_cast<const R*>(st.pre);
#pragma unroll
  for (int c = 0; c < kPre; ++c) {
    const int e = t + c * NL;
    if (e < nd * dimS) {
      const int dir = e >= dimS, q = e - dir * dimS;
      const cx<R> v = cscale(pv[c], pre[e]);
      V[dir * P + q] = dir ? cconj(v) : v;
    }
  }
  wave_sync();
  const cx<R>* Wr = T1;
  if (st.s2 <= kSymRowS2) sym_apply_pair_rows<R, symc::sdim(K1)>(st, C, S2, V, T1, t);  // T1 = V here
  else Wr = sym_apply2<R>(st, C, S2, V, T1, W, t, NL);
  if (!act) return;
  sym_project<T, R>(st, 0, tb.d, Wr, out + 2 * ((size_t(g) * D + mu) * tb.d), t, NL);
  if (both) sym_project<T, R>(st, 1, tb.d, Wr + P, static_cast<T*>(st.hopsL) + 2 * ((size_t(gp) * D + mu) * tb.d), t, NL);
}

// ---------------------------------------------------------------------------
// FP32, SU(3), first Young row 2..6 and second row <= 1: two links per wave,
// Sym^k(U) never stored. Lane t of a half-wave builds column t of C_k in
// registers, level by level: column n of level k extends column n - e_j of
// level k - 1 (as in sym_level), which the lane takes from the lane holding it
// (ds_bpermute). The adjoint hop is then a column dot, (C1^T V1)[t]. The
// forward hop needs rows of C1; since
//   C_k(U)^T = F_k^2 C_k(U^T) F_k^-2,  F_k = diag sqrt(m!),
// row t of C1 is column t of C_k(U^T) up to those factors, so a second build
// from U^T gives the forward hop as the same column dot, with V0 lifted by F
// instead of F^-1 and projected with the adjoint's rows (Yc F^-1 / c). Only
// the two lifted vectors pass through LDS, and the results over them.

// Column t of C_K(A) (t >= dim Sym^K repeats the last column); lane0: the
// first lane of this half-wave. KEEP: prev receives this lane's column of
// level K - 1 as well.
template <int K, bool KEEP = false>
GHX_DEV void sym_column(const Mat<float, 3>& A, int t, int lane0, cx<float> (&c)[symc::sdim(K)], cx<float>* prev = nullptr) {
  constexpr int SK = symc::sdim(K), SP = symc::sdim(K - 1);
  const int n = t < SK ? t : SK - 1;
  const symc::Mono x = symc::mono(K, n);
  const int j = x.a[0] > 0 ? 0 : (x.a[1] > 0 ? 1 : 2);
  const int np = j == 0 ? symc::midx(K - 1, x.a[0] - 1, x.a[1]) : (j == 1 ? symc::midx(K - 1, 0, x.a[1] - 1) : symc::midx(K - 1, 0, 0));
  cx<float> col[SP];
  if constexpr (K == 2) {
#pragma unroll
    for (int i = 0; i < 3; ++i) col[i] = sel3(A.e[i], np);
    if constexpr (KEEP) {
#pragma unroll
      for (int i = 0; i < 3; ++i) prev[i] = sel3(A.e[i], t < 3 ? t : 2);
    }
  } else {
    cx<float> p[SP];
    sym_column<K - 1>(A, t, lane0, p);
    const int src = (lane0 + np) * 4;
    static_for<0, SP>([&](auto m) { col[int(m)] = {lane_take(p[int(m)].r, src), lane_take(p[int(m)].i, src)}; });
    if constexpr (KEEP) static_for<0, SP>([&](auto m) { prev[int(m)] = p[int(m)]; });
  }
  cx<float> u[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) u[i] = sel3(A.e[i], j);
  static_for<0, SK>([&](auto mc) {
    constexpr int m = mc;
    constexpr int r0 = symc::rowsrc(K, m, 0), r1 = symc::rowsrc(K, m, 1), r2 = symc::rowsrc(K, m, 2);
    cx<float> acc = {0.f, 0.f};
    if constexpr (r0 >= 0) cfma(acc, u[0], col[r0]);
    if constexpr (r1 >= 0) cfma(acc, u[1], col[r1]);
    if constexpr (r2 >= 0) cfma(acc, u[2], col[r2]);
    c[m] = acc;
  });
}

// Column t of C_K(A) for K >= 4 as a polynomial product: column n of C_K
// holds the coefficients of prod_j l_j^{n_j} (l_j = sum_i A_ij x_i), the
// product of columns a (level KA = ceil(K/2)) and b = n - a (level KB =
// floor(K/2)). The lane builds its own columns of levels KA and KB, takes
// columns a and b from the lanes that hold them, and multiplies: dim Sym^KA x
// dim Sym^KB complex MACs (100 at K = 6) instead of the 3 dim Sym^(K-1) of
// the last levels' recursion (63 + 45 + 30 at K = 6), and fewer lane
// transfers.
template <int K>
GHX_DEV void sym_column_prod(const Mat<float, 3>& A, int t, int lane0, cx<float> (&c)[symc::sdim(K)]) {
  constexpr int KA = (K + 1) / 2, KB = K / 2, SA = symc::sdim(KA), SB = symc::sdim(KB), SK = symc::sdim(K);
  cx<float> pa[SA], pb[SB];
  if constexpr (KA == KB) {
    sym_column<KA>(A, t, lane0, pa);
    static_for<0, SB>([&](auto m) { pb[int(m)] = pa[int(m)]; });
  } else {
    sym_column<KA, true>(A, t, lane0, pa, pb);
  }
  const int n = t < SK ? t : SK - 1;
  const symc::Mono x = symc::mono(K, n);
  const int a0 = x.a[0] < KA ? x.a[0] : KA;
  const int a1 = x.a[1] < KA - a0 ? x.a[1] : KA - a0;  // a2 = KA - a0 - a1 <= x.a[2]
  const int sa = (lane0 + symc::midx(KA, a0, a1)) * 4, sb = (lane0 + symc::midx(KB, x.a[0] - a0, x.a[1] - a1)) * 4;
  cx<float> ca[SA], cb[SB];
  static_for<0, SA>([&](auto m) { ca[int(m)] = {lane_take(pa[int(m)].r, sa), lane_take(pa[int(m)].i, sa)}; });
  static_for<0, SB>([&](auto m) { cb[int(m)] = {lane_take(pb[int(m)].r, sb), lane_take(pb[int(m)].i, sb)}; });
  static_for<0, SK>([&](auto m) { c[int(m)] = {0.f, 0.f}; });
  static_for<0, SA>([&](auto mac) {
    constexpr symc::Mono ma = symc::mono(KA, int(mac));
    static_for<0, SB>([&](auto mbc) {
      constexpr symc::Mono mb = symc::mono(KB, int(mbc));
      constexpr int m = symc::midx(K, ma.a[0] + mb.a[0], ma.a[1] + mb.a[1]);
      cfma(c[m], ca[int(mac)], cb[int(mbc)]);
    });
  });
}
template <int K>
GHX_DEV void sym_column_top(const Mat<float, 3>& A, int t, int lane0, cx<float> (&c)[symc::sdim(K)]) {
  if constexpr (K >= 4) sym_column_prod<K>(A, t, lane0, c);
  else sym_column<K>(A, t, lane0, c);
}

// W[t][k] for this lane's column c of C1 and the lifted vector V (s1 x s2 in
// LDS, read as broadcasts), C2 = U applied on the right as C2^T (fwd) or on
// the left's adjoint form (adjoint), as in sym_apply_pair_rows.
template <int S1N, int S2N, bool ADJ>
GHX_DEV void sym_column_apply(const cx<float> (&c)[S1N], const cx<float>* V, const Mat<float, 3>& U, cx<float> (&w)[S2N]) {
  cx<float> acc[S2N];
#pragma unroll
  for (int b = 0; b < S2N; ++b) acc[b] = {0.f, 0.f};
  static_for<0, S1N>([&](auto ac) {
    constexpr int a = ac;
#pragma unroll
    for (int b = 0; b < S2N; ++b) cfma(acc[b], c[a], V[a * S2N + b]);
  });
  if constexpr (S2N == 1) {
    w[0] = acc[0];
  } else {
#pragma unroll
    for (int k = 0; k < S2N; ++k) {
      cx<float> z = {0.f, 0.f};
#pragma unroll
      for (int k2 = 0; k2 < S2N; ++k2) cfma(z, ADJ ? U.e[k2][k] : U.e[k][k2], acc[k2]);
      w[k] = z;
    }
  }
}

// Six waves per SIMD (80 VGPRs, 2 spilled) run the {6,1} step fastest: four
// (101 VGPRs, none spilled) 0.2486 ms, five 0.2376, six 0.2303, seven (12
// spilled) 0.2338.
template <int D, int K1, int K2>
__global__ __launch_bounds__(kSymWave, 6) void sym_link_col_kernel(BigArgs<float> a, BigRepTables tb, float* out, int both) {
  using R = float;
  constexpr int NL = kSymPairLanes, S1N = symc::sdim(K1), S2N = K2 == 0 ? 1 : 3, DS = S1N * S2N;
  constexpr int kPre = (2 * DS + NL - 1) / NL;
  static_assert(K2 <= 1 && S1N <= NL, "columns of C_k1 must fit a half-wave");
  const BigRepTables::Sym& st = tb.sym;
  __shared__ cx<R> vs[2][2 * DS];  // per half-wave: V0, V1 (then the two hops before projection)
  const int lane = threadIdx.x, h = lane / NL, t = lane % NL;
  const unsigned g = blockIdx.x, Vp = a.Vp;
  const int r = a.blk_rep[g / kBlock];
  const RepDesc& rd = a.reps[r];
  const unsigned s = g - rd.offset;
  if (s >= rd.vol) return;  // whole block exits together
  const int mu = 2 * int(blockIdx.y) + h;
  const bool act = mu < D;
  const int ml = act ? mu : 0;  // an idle half (odd D) repeats link 0 and stores nothing
  const Nbr<D> nb(rd, s);
  const unsigned gp = g + nb.upv(rd, ml);
  const Mat<R, 3> U = load_link<float, R, 3>(a.U, Vp, ml, g);
  const int d = tb.d, nd = both + 1;
  // V0 = F Yc phi(x+mu) for the forward hop of x, V1 = conj(F Yc phi(x)) for the
  // adjoint hop of x+mu
  const float* lift = static_cast<const float*>(st.lift);
  const R* pre = st.pref + DS;  // F
  cx<R>* V = vs[h];
#pragma unroll
  for (int c = 0; c < kPre; ++c) {
    const int e = t + c * NL;
    if (e < nd * DS) {
      const int dir = e >= DS, q = e - dir * DS;
      const cx<R> v = cscale(ld_cx(lift + size_t(dir ? g : gp) * 2 * DS + 2 * q), pre[q]);
      V[e] = dir ? cconj(v) : v;
    }
  }
  wave_sync();
  // each half's reads of V precede its writes of the results in the wave's
  // instruction stream, so the results overwrite V in place
  if (both) {
    cx<R> c[S1N], w[S2N];
    sym_column_top<K1>(U, t, h * NL, c);
    sym_column_apply<S1N, S2N, true>(c, V + DS, U, w);
    if (t < S1N) {
#pragma unroll
      for (int k = 0; k < S2N; ++k) V[DS + t * S2N + k] = w[k];
    }
  }
  {
    Mat<R, 3> UT;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) UT.e[i][j] = U.e[j][i];
    cx<R> c[S1N], w[S2N];
    sym_column_top<K1>(UT, t, h * NL, c);
    sym_column_apply<S1N, S2N, false>(c, V, U, w);
    if (t < S1N) {
#pragma unroll
      for (int k = 0; k < S2N; ++k) V[t * S2N + k] = w[k];
    }
  }
  wave_sync();
  if (!act) return;
  sym_project<float, R>(st, 1, d, V, out + 2 * ((size_t(g) * D + mu) * d), t, NL);
  if (both) sym_project<float, R>(st, 1, d, V + DS, static_cast<float*>(st.hopsL) + 2 * ((size_t(gp) * D + mu) * d), t, NL);
}

template <typename R>
GHX_DEV R wave_sum(R v) {
#pragma unroll
  for (int o = kSymWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

template <typename T, int D>
GHX_DEV const T* hl_of(const BigRepTables& tb, unsigned g) {
  return static_cast<const T*>(tb.sym.hopsL) + 2 * (size_t(g) * D * tb.d);
}

// The part of a site's finish after its scalar and momentum are loaded: the
// backward hops summed, g^a and the forward hops, the on-site term,
// then the kick and drift into the site's staged rows (or its scalar force).
template <typename T, int N, int D, class HOP>
GHX_DEV void finish_site_body(const BigArgs<T>& a, const BigRepTables& tb, const RepDesc& rd, const T* __restrict__ hl, HOP&& hop,
                              const cx<T> ph, const cx<T> pv0, cx<T> (&psi)[D][kSymWave], cx<T>* phs, T* rph, T* rpi,
                              T* io_g, int si, unsigned g, int nmu) {
  using R = T;
  constexpr int NGN = N * N - 1;
  const unsigned Vp = a.Vp;
  const int d = tb.d, l = threadIdx.x;
  const bool on = l < d;
  // loads that do not depend on the backward hops, issued before them
  cx<R> psf[D], psb[D];
#pragma unroll
  for (int mu = 0; mu < D; ++mu) {
    psf[mu] = mu < nmu ? hop(mu) : cx<R>{R(0), R(0)};
    psb[mu] = mu < nmu && on ? ld_cx(hl + 2 * (mu * d + l)) : cx<R>{R(0), R(0)};
  }
  const R k2 = R(2.0 * rd.kappa);
  const BigRepTables::Sym& st = tb.sym;
  const int lpg = st.glpg;
  const int* glm = st.glm + l;
  const R* gval = (sizeof(R) == 4 ? reinterpret_cast<const R*>(st.gvalf) : reinterpret_cast<const R*>(st.gval)) + 2 * l;
  phs[l] = ph;
  cx<R> f = {R(0), R(0)};
  if (nmu) {  // the backward hops (carried conjugated): chi = conj(sum_mu hl[mu][l])
    cx<R> v = psb[0];
#pragma unroll
    for (int mu = 1; mu < D; ++mu) v = cadd(v, psb[mu]);
    f = {k2 * v.r, -k2 * v.i};
    // the D forward hops at once, then g^a = -2 Im sum_lm conj(phi_l) T^a_lm psi_m
    // = -2 Im sum_m conj(chi_m) psi_m with chi = T^a phi (T^a hermitian): the
    // lanes of generator a form chi row by row and dot it with every direction
#pragma unroll
    for (int mu = 0; mu < D; ++mu) {
      const cx<R> ps = psf[mu];
      psi[mu][l] = ps;
      f.r += k2 * ps.r;
      f.i += k2 * ps.i;
    }
    wave_sync();
    R gp[D];
#pragma unroll
    for (int mu = 0; mu < D; ++mu) gp[mu] = R(0);
    const int nz = st.gK, jl = l % lpg;
#pragma unroll 2
    for (int rr = 0; rr < st.grow; ++rr) {
      const int m = jl + rr * lpg;  // < 64; rows past d are zero in the table and in psi
      cx<R> chi = {R(0), R(0)};
#pragma unroll 4
      for (int k = 0; k < nz; ++k) {
        const int e = (rr * nz + k) * kSymWave;
        cfma(chi, ld_cx(gval + 2 * e), phs[glm[e]]);
      }
#pragma unroll
      for (int mu = 0; mu < D; ++mu) {
        const cx<R> y = psi[mu][m];
        gp[mu] = fma(chi.r, y.i, fma(-chi.i, y.r, gp[mu]));
      }
    }
#pragma unroll
    for (int mu = 0; mu < D; ++mu) {
      R v = gp[mu];
      for (int o = 1; o < lpg; o <<= 1) v += __shfl_xor(v, o);
      if (l % lpg == 0 && l / lpg < NGN) io_g[(mu * NGN + l / lpg) * kFinishSites + si] = T(R(-rd.kappa) * (R(-2) * v));
    }
  }
  // on-site term -2 dV/dphi^* (scalar_higgs.hpp:25-39) or the plugged-in potential's planes
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
    if (on) st_cx(rpi + 2 * l, cx<T>{T(f.r), T(f.i)});
    return;
  }
  const bool abs_eps = a.flags & F_ABS_EPS;
  const R ek = R(abs_eps ? a.ck : a.ck * rd.eps), ed = R(abs_eps ? a.cd : a.cd * rd.eps);
  const bool frozen = a.flags & F_FROZEN;
  cx<R> pv = pv0;
  if (a.flags & F_KICK) {
    pv.r = fma(ek, f.r, pv.r);
    pv.i = fma(ek, f.i, pv.i);
    if (frozen) {  // |phi| = 1 (gauge_higgs_hmc.hpp:61-103): tangent projection
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
  if ((a.flags & F_KICK) || (frozen && (a.flags & F_DRIFT))) st_cx(rpi + 2 * l, cx<T>{T(pv.r), T(pv.i)});
  if (a.flags & F_DRIFT) st_cx(rph + 2 * l, cx<T>{T(pn.r), T(pn.i)});
}

// A site's matter terms from its 2D hops (sym_link_kernel), one wave per site
// with lane l holding component l (d <= 64): the matter link force
// -kappa g^a, g^a = -2 Im phi^dag T^a_R psi_mu (scalar_higgs.hpp:110-128), and
// the scalar force 2 kappa sum(hops) - 2 dV/dphi^* with the pi kick, the
// tangent projection and great-circle drift of a frozen scalar and the phi
// drift (big_matter_kernel's contract), or q1 = sum_mu Re phi^dag psi_mu for
// the action. Sums over components are wave reductions.
//
// phi, pi, G and the scalar force live in planes (one per real component), so
// a wave that holds one site's components would touch a different cache line
// in every lane. A workgroup therefore holds kFinishSites consecutive sites,
// one wave each: all its threads read the sites' plane rows (kFinishSites
// values each) into LDS, every wave finishes its site from LDS, and the
// results leave through LDS as plane rows again.
constexpr int kFinishRow = 2 * kSymWave + 2;  // LDS words per staged site (2d <= 128, even for 8-byte access)

template <typename T, int N, int D>
__global__ __launch_bounds__(kSymWave* kFinishSites) void sym_finish_kernel(BigArgs<T> a, BigRepTables tb, const T* __restrict__ hops) {
  using R = T;
  constexpr int NGN = N * N - 1, S = kFinishSites, NT = kSymWave * kFinishSites;
  constexpr int kScr = D * kSymWave;  // the forward hops
  static_assert(kBlock % S == 0, "a workgroup's sites lie in one replica block");
  __shared__ cx<R> scr_s[S][kScr], phs_s[S][kSymWave];
  __shared__ __align__(16) T io_ph[S * kFinishRow];
  __shared__ __align__(16) T io_pi[S * kFinishRow];
  __shared__ T io_g[D * NGN * S];
  const unsigned g0 = blockIdx.x * S, Vp = a.Vp;
  const int r = a.blk_rep[g0 / kBlock];
  const RepDesc& rd = a.reps[r];
  const int tid = int(threadIdx.y) * kSymWave + int(threadIdx.x);
  const int np = 2 * tb.d;
  const bool stepm = a.mode == BIG_STEP;
  auto live = [&](int si) { return g0 + unsigned(si) - rd.offset < rd.vol; };
  // plane rows in: component plane p of the S sites, S consecutive values
  for (int e = tid; e < np * S; e += NT) {
    const int p = e / S, si = e % S;
    const bool lv = live(si);
    io_ph[si * kFinishRow + p] = lv ? a.phi_in[size_t(p) * Vp + g0 + si] : T(0);
    if (stepm) io_pi[si * kFinishRow + p] = lv ? a.pi[size_t(p) * Vp + g0 + si] : T(0);
  }
  __syncthreads();
  const int si = int(threadIdx.y);
  if (live(si)) {
    const unsigned g = g0 + unsigned(si);
    cx<R>(&psi)[D][kSymWave] = *reinterpret_cast<cx<R>(*)[D][kSymWave]>(scr_s[si]);
    cx<R>(&phs)[kSymWave] = phs_s[si];
    T* rph = io_ph + si * kFinishRow;  // this site's staged phi (then phi_out)
    T* rpi = io_pi + si * kFinishRow;  // its staged pi (then pi, or the scalar force of BIG_FORCE)
    const int d = tb.d, l = threadIdx.x;
    const bool on = l < d;
    const T* hg = hops + 2 * (size_t(g) * D * d);
    auto hop = [&](int mu) { return on ? ld_cx(hg + 2 * (mu * d + l)) : cx<R>{R(0), R(0)}; };
    const cx<R> ph = on ? ld_cx(rph + 2 * l) : cx<R>{R(0), R(0)};
    const int nmu = (a.flags & F_NO_HOP) ? 0 : D;
    if (a.mode == BIG_ACTION) {
      double q = 0.0;
      for (int mu = 0; mu < nmu; ++mu) {
        const cx<R> ps = hop(mu);
        q += double(ph.r) * double(ps.r) + double(ph.i) * double(ps.i);
      }
      q = wave_sum(q);
      if (l == 0) a.q1[g] = q;
    } else {
      finish_site_body<T, N, D>(a, tb, rd, hl_of<T, D>(tb, g), hop, ph, on ? (stepm ? ld_cx(rpi + 2 * l) : cx<R>{R(0), R(0)}) : cx<R>{R(0), R(0)},
                                psi, phs, rph, rpi, io_g, si, g, nmu);
    }
  }
  if (a.mode == BIG_ACTION) return;
  __syncthreads();
  // plane rows out
  const bool put_pi = stepm && ((a.flags & F_KICK) || ((a.flags & F_FROZEN) && (a.flags & F_DRIFT)));
  const bool put_ph = stepm && (a.flags & F_DRIFT);
  for (int e = tid; e < np * S; e += NT) {
    const int p = e / S, sj = e % S;
    if (!live(sj)) continue;
    const size_t o = size_t(p) * Vp + g0 + sj;
    if (a.mode == BIG_FORCE) a.Fphi[o] = io_pi[sj * kFinishRow + p];
    if (put_pi) a.pi[o] = io_pi[sj * kFinishRow + p];
    if (put_ph) a.phi_out[o] = io_ph[sj * kFinishRow + p];
  }
  if (!(a.flags & F_NO_HOP))
    for (int e = tid; e < D * NGN * S; e += NT) {
      const int p = e / S, sj = e % S;
      if (live(sj)) a.G[size_t(p) * Vp + g0 + sj] = io_g[e];
    }
}

}  // namespace ghx
