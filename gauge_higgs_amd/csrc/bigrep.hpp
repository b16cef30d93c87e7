// Large tensor irreps (e.g. SU(3) {6,1}: d = 48, N^n = 3^7 = 2187): the matter
// part of the MD step with one workgroup per site and the tensor in LDS.
//
// Per site x and direction mu the kernel forms, exactly as the reference's
// matrix-free apply (rep_general.hpp:196-220):
//   psi_mu = W U_mu(x)^{(x)n} W^T phi_{x+mu}       (lift, n slot applies, project)
//   chi_mu = W (U_mu(x-mu)^dag)^{(x)n} W^T phi_{x-mu}
// then the scalar force, the matter back-reaction g^a = -2 Im phi_x^dag T^a_R psi_mu
// (scalar_higgs.hpp:114-128) and, in step mode, the pi kick and phi drift.
// W is applied sparsely (it is weight-block sparse: <= 210 of 2187 entries per
// row for {6,1}); T^a_R is dense d x d.
#pragma once
#include "kernels.hpp"

namespace ghx {

constexpr int kBigBlock = 256;
constexpr int kBigMaxD = 64;      // largest irrep dimension of this path
constexpr int kBigMaxSlots = 8;   // largest number of boxes

struct BigRepTables {
  int d, nbox, dimT;
  const int* col_ptr;   // CSC of W by tensor index t: entries [col_ptr[t], col_ptr[t+1])
  const int* col_row;   // basis index l of each entry
  const double* col_val;
  const int* row_ptr;   // CSR of W by basis index l
  const int* row_col;   // tensor index t of each entry
  const double* row_val;
  const float* col_valf;  // the W values again in FP32 (read by the FP32 kernels: half the table traffic)
  const float* row_valf;
  const int* T_ptr;     // CSR of T^a_R by row (a*d + l): entries [T_ptr[r], T_ptr[r+1])
  const int* T_col;     // column m of each entry
  const double* T_val;  // complex value (re, im) of each entry
  // symmetric-power engine (symrep.hpp): D(U) = invc Yc^T (Sym^k1(U) (x) Sym^k2(U)) Yc
  struct Sym {
    int on = 0;           // 1: the engine replaces big_matter_kernel
    int s1 = 0, s2 = 1, dimS = 0, npass = 0;
    int s1p = 0, s2p = 1;                 // LDS row strides of C1 and C2
    const int4* prog = nullptr;       // LDS program: npass x 64 entries {s0, s1, s2, dst | j}, LDS byte offsets
    const double* pre = nullptr;      // [2][dimS]: 1/F (forward hop) and F (adjoint), FP64 and FP32 copies
    const float* pref = nullptr;
    // the projections Yc F / c (forward hop) and Yc F^-1 / c (adjoint) as sparse
    // rows: entry k of component l of direction dir at [(dir ynz + k) d + l]
    const int* ysq = nullptr;         // the lifted index q
    const double* ysv = nullptr;      // the value, FP64 and FP32 copies
    const float* ysvf = nullptr;
    int ynz = 0;
    int ng = 0;                       // groups of 4 lifted components (dim S rounded up): V and W regions are 4 ng apart
    const double* Yl = nullptr;       // the lift table Yc transposed: [d][dimSp], dimSp = dimS rounded up to 16, zero padded
    const float* Ylf = nullptr;
    int dimSp = 0;
    void* lift = nullptr;          // Yc phi per site, site-major: [Vp][dimS] complex, storage type
    void* hops = nullptr;          // forward hops D(U_mu(x)) phi(x+mu), site-major: [Vp][D][d] complex, storage type
    void* hopsL = nullptr;         // backward hops D(U_mu(x))^dag phi(x) of x+mu, conjugated: [Vp][D][d] complex
    int lds_one = 0, lds_S1 = 0, lds_S2 = 0, lds_V = 0, lds_T1 = 0, lds_W = 0, lds_total = 0;  // cx units
    int k1 = 0, k2 = 0, lds_Cb = 0;   // Young rows; the level buffer that alternates with lds_S1
    int pair = 0;                     // 1: sym_link_pair_kernel (SU(3), k1 <= 6), two LDS regions per workgroup; 2: sym_link_col_kernel
    int pair_S2 = 0, pair_V = 0, pair_T1 = 0, pair_total = 0;  // its per-link region (C at 0), cx units
    // g^a by lanes: generator a owns lanes [a lpg, (a+1) lpg); lane j of them
    // takes rows m = j + rr lpg (rr < grow) of T^a, entry k of row slot rr being
    // (l, T^a_ml) at [(rr gK + k) 64 + lane] (gK entries per row, zero padded)
    const int* glm = nullptr;
    const double* gval = nullptr;     // complex, FP64 and FP32 copies
    const float* gvalf = nullptr;
    int gK = 0, glpg = 0, grow = 0;
  } sym;
};

enum BigMode : int { BIG_STEP = 0, BIG_FORCE = 1, BIG_ACTION = 2 };

template <typename R>
GHX_DEV R wval(const double* vd, const float* vf, int e) {
  if constexpr (sizeof(R) == 4) return vf[e];
  else return vd[e];
}

template <typename T>
struct BigArgs {
  const RepDesc* reps;
  const int* site_rep;   // not used: sites map through blk_rep of kBlock-sized blocks
  const int* blk_rep;
  unsigned Vp;
  const T* U;            // links (read)
  const T* phi_in;
  T* phi_out;
  T* pi;
  T* G;                  // matter link force -kappa g^a, planes (mu*(N^2-1)+a)
  T* Fphi;               // full scalar force (BIG_FORCE)
  double* q1;            // per-site sum_mu Re phi_x^dag psi_mu (BIG_ACTION)
  double ck, cd;
  int flags;             // F_KICK / F_DRIFT / F_ABS_EPS as in StepArgs
  int mode;
  const T* Fon = nullptr;  // on-site force planes of a plugged-in potential (else the quartic)
};

// In-place (M on slot s) over the whole tensor: each thread owns whole fibres
// (the N entries that differ only in digit s), so no two threads touch the
// same entry and one LDS buffer suffices.
template <typename R, int N, bool DAG>
GHX_DEV void big_slot(const Mat<R, N>& M, cx<R>* __restrict__ X, int dimT, int blk) {
  const int nfib = dimT / N;
  for (int f = threadIdx.x; f < nfib; f += kBigBlock) {
    const int base = (f / blk) * (N * blk) + (f % blk);
    cx<R> in[N];
#pragma unroll
    for (int k = 0; k < N; ++k) in[k] = X[base + k * blk];
#pragma unroll
    for (int j = 0; j < N; ++j) {
      cx<R> z = {R(0), R(0)};
#pragma unroll
      for (int k = 0; k < N; ++k) {
        if (DAG) cfmac(z, M.e[k][j], in[k]);
        else cfma(z, M.e[j][k], in[k]);
      }
      X[base + j * blk] = z;
    }
  }
}

// out[l] = (W M^{(x)n} W^T v)_l with M = U or U^dag; X is a dimT LDS buffer,
// v and out are d-vectors in LDS.
template <typename R, int N, bool DAG>
GHX_DEV void big_apply(const BigRepTables& tb, const Mat<R, N>& M, const cx<R>* v, cx<R>* out, cx<R>* X) {
  for (int t = threadIdx.x; t < tb.dimT; t += kBigBlock) {  // lift
    cx<R> z = {R(0), R(0)};
    for (int e = tb.col_ptr[t]; e < tb.col_ptr[t + 1]; ++e) {
      const R w = wval<R>(tb.col_val, tb.col_valf, e);
      const cx<R> x = v[tb.col_row[e]];
      z.r = fma(w, x.r, z.r);
      z.i = fma(w, x.i, z.i);
    }
    X[t] = z;
  }
  __syncthreads();
  int blk = tb.dimT / N;
  for (int s = 0; s < tb.nbox; ++s) {
    big_slot<R, N, DAG>(M, X, tb.dimT, blk);
    __syncthreads();
    blk /= N;
  }
  // project: kProjLanes adjacent lanes share a row (a row holds ~dimT/d
  // entries) and combine their partial sums with lane shuffles
  constexpr int kProjLanes = 4;
  for (int idx = threadIdx.x; idx < tb.d * kProjLanes; idx += kBigBlock) {
    const int l = idx / kProjLanes, part = idx % kProjLanes;
    cx<R> z = {R(0), R(0)};
    for (int e = tb.row_ptr[l] + part; e < tb.row_ptr[l + 1]; e += kProjLanes) {
      const R w = wval<R>(tb.row_val, tb.row_valf, e);
      const cx<R> x = X[tb.row_col[e]];
      z.r = fma(w, x.r, z.r);
      z.i = fma(w, x.i, z.i);
    }
#pragma unroll
    for (int o = 1; o < kProjLanes; o <<= 1) {
      z.r += __shfl_xor(z.r, o, 64);
      z.i += __shfl_xor(z.i, o, 64);
    }
    if (part == 0) out[l] = z;
  }
  __syncthreads();
}

// The site's scalar after its hop terms (fph, in LDS, shared by the whole
// workgroup): the on-site term -2 dV/dphi^* (scalar_higgs.hpp:25-39, or the
// Fon planes of a plugged-in potential), then the F_FORCE output or the pi kick
// and phi drift, with the tangent projection and great-circle drift of a
// frozen scalar. vin is workgroup scratch of d entries.
template <typename T>
GHX_DEV void big_phi_finish(const BigArgs<T>& a, const RepDesc& rd, unsigned g, int d, int nthreads, cx<T>* phx, cx<T>* fph,
                            cx<T>* vin) {
  using R = T;
  const unsigned Vp = a.Vp;
  // on-site term -2 dV/dphi^* (scalar_higgs.hpp:25-39)
  __shared__ R n2s;
  if (threadIdx.x == 0) {
    R n2 = R(0);
    for (int l = 0; l < d; ++l) n2 += phx[l].r * phx[l].r + phx[l].i * phx[l].i;
    n2s = n2;
  }
  __syncthreads();
  const R c = (a.flags & F_NO_ONSITE) ? R(0) : R(-2) * (R(1) + R(2.0 * rd.lambda) * (n2s - R(1)));
  const bool abs_eps = a.flags & F_ABS_EPS;
  const R ek = R(abs_eps ? a.ck : a.ck * rd.eps), ed = R(abs_eps ? a.cd : a.cd * rd.eps);
  for (int l = threadIdx.x; l < d; l += nthreads) {
    cx<R> f = fph[l];
    if (!a.Fon) {
      f.r = fma(c, phx[l].r, f.r);
      f.i = fma(c, phx[l].i, f.i);
    } else if (!(a.flags & F_NO_ONSITE)) {
      f.r += R(a.Fon[size_t(2 * l) * Vp + g]);
      f.i += R(a.Fon[size_t(2 * l + 1) * Vp + g]);
    }
    if (a.mode == BIG_FORCE) {
      a.Fphi[size_t(2 * l) * Vp + g] = T(f.r);
      a.Fphi[size_t(2 * l + 1) * Vp + g] = T(f.i);
      continue;
    }
    R pr = R(a.pi[size_t(2 * l) * Vp + g]), pim = R(a.pi[size_t(2 * l + 1) * Vp + g]);
    if (a.flags & F_KICK) {
      pr = fma(ek, f.r, pr);
      pim = fma(ek, f.i, pim);
    }
    vin[l] = {pr, pim};  // the site's momenta, kicked
  }
  if (a.mode == BIG_FORCE) return;
  const bool frozen = a.flags & F_FROZEN;
  __syncthreads();
  if (frozen) {
    // |phi| = 1 (gauge_higgs_hmc.hpp:61-103): project the kicked pi onto the
    // tangent space, then move along the great circle
    __shared__ R red;
    if (a.flags & F_KICK) {
      if (threadIdx.x == 0) {
        R rp = R(0);
        for (int l = 0; l < d; ++l) rp += phx[l].r * vin[l].r + phx[l].i * vin[l].i;
        red = rp;
      }
      __syncthreads();
      const R rp = red;
      for (int l = threadIdx.x; l < d; l += nthreads) { vin[l].r -= rp * phx[l].r; vin[l].i -= rp * phx[l].i; }
      __syncthreads();
    }
    if (a.flags & F_DRIFT) {
      if (threadIdx.x == 0) {
        R r2 = R(0);
        for (int l = 0; l < d; ++l) r2 += vin[l].r * vin[l].r + vin[l].i * vin[l].i;
        red = r2;
      }
      __syncthreads();
      const R rr = sqrt(red);
      const R rmin = sizeof(R) == 8 ? R(1e-300) : R(1e-30);
      if (rr >= rmin) {
        const R cs = cos(rr * ed), sn = sin(rr * ed), q = sn / rr, w = -rr * sn;
        for (int l = threadIdx.x; l < d; l += nthreads) {
          const cx<R> fv = phx[l], pv = vin[l];
          phx[l] = {cs * fv.r + q * pv.r, cs * fv.i + q * pv.i};
          vin[l] = {w * fv.r + cs * pv.r, w * fv.i + cs * pv.i};
        }
      }
    }
  }
  for (int l = threadIdx.x; l < d; l += nthreads) {
    if ((a.flags & F_KICK) || (frozen && (a.flags & F_DRIFT))) {
      a.pi[size_t(2 * l) * Vp + g] = T(vin[l].r);
      a.pi[size_t(2 * l + 1) * Vp + g] = T(vin[l].i);
    }
    if (a.flags & F_DRIFT) {
      const cx<R> ph = frozen ? phx[l] : cx<R>{fma(ed, vin[l].r, phx[l].r), fma(ed, vin[l].i, phx[l].i)};
      a.phi_out[size_t(2 * l) * Vp + g] = T(ph.r);
      a.phi_out[size_t(2 * l + 1) * Vp + g] = T(ph.i);
    }
  }
}

template <typename T, int N, int D>
__global__ __launch_bounds__(kBigBlock) void big_matter_kernel(BigArgs<T> a, BigRepTables tb) {
  using R = T;
  constexpr int NGN = N * N - 1;
  extern __shared__ unsigned char smem_raw[];
  cx<R>* X = reinterpret_cast<cx<R>*>(smem_raw);
  __shared__ cx<R> phx[kBigMaxD], vin[kBigMaxD], psi[kBigMaxD], fph[kBigMaxD];
  __shared__ R gred[NGN][kBigMaxD];
  __shared__ double q1s;
  // one block per site; sites are numbered like the kBlock grid of the step kernels
  const unsigned gsite = blockIdx.x;
  const int r = a.blk_rep[gsite / kBlock];
  const RepDesc& rd = a.reps[r];
  const unsigned s = gsite - rd.offset;
  if (s >= rd.vol) return;  // whole block exits together
  const unsigned g = gsite, Vp = a.Vp;
  const Nbr<D> nb(rd, s);
  const int d = tb.d;
  const R k2 = R(2.0 * rd.kappa);
  for (int l = threadIdx.x; l < d; l += kBigBlock) {
    phx[l] = {R(a.phi_in[size_t(2 * l) * Vp + g]), R(a.phi_in[size_t(2 * l + 1) * Vp + g])};
    fph[l] = {R(0), R(0)};
  }
  if (threadIdx.x == 0) q1s = 0.0;
  __syncthreads();
  const int nmu = (a.flags & F_NO_HOP) ? 0 : D;
  for (int mu = 0; mu < nmu; ++mu) {
    // forward: psi = D(U_mu(x)) phi_{x+mu}
    const unsigned gpm = g + nb.upv(rd, mu);
    for (int l = threadIdx.x; l < d; l += kBigBlock)
      vin[l] = {R(a.phi_in[size_t(2 * l) * Vp + gpm]), R(a.phi_in[size_t(2 * l + 1) * Vp + gpm])};
    const Mat<R, N> Ux = load_link<T, R, N>(a.U, Vp, mu, g);
    __syncthreads();
    big_apply<R, N, false>(tb, Ux, vin, psi, X);
    if (a.mode == BIG_ACTION) {
      if (threadIdx.x == 0) {
        double q = 0.0;
        for (int l = 0; l < d; ++l) q += double(phx[l].r) * double(psi[l].r) + double(phx[l].i) * double(psi[l].i);
        q1s += q;
      }
      __syncthreads();
      continue;
    }
    // g^a = -2 Im sum_{l,m} conj(phx_l) T^a_{lm} psi_m ; link force -kappa g^a
    for (int idx = threadIdx.x; idx < NGN * d; idx += kBigBlock) {
      const int ga = idx / d, l = idx % d;
      cx<R> t = {R(0), R(0)};
      for (int e = tb.T_ptr[idx]; e < tb.T_ptr[idx + 1]; ++e)
        cfma(t, cx<R>{R(tb.T_val[2 * e]), R(tb.T_val[2 * e + 1])}, psi[tb.T_col[e]]);
      cx<R> z = {R(0), R(0)};
      cfmac(z, phx[l], t);
      gred[ga][l] = z.i;
    }
    __syncthreads();
    if (threadIdx.x < NGN) {
      R sum = R(0);
      for (int l = 0; l < d; ++l) sum += gred[threadIdx.x][l];
      const R fm = R(-rd.kappa) * (R(-2) * sum);
      a.G[size_t(mu * NGN + threadIdx.x) * Vp + g] = T(fm);
    }
    for (int l = threadIdx.x; l < d; l += kBigBlock) { fph[l].r += k2 * psi[l].r; fph[l].i += k2 * psi[l].i; }
    __syncthreads();
    // backward: chi = D(U_mu(x-mu))^dag phi_{x-mu}
    const unsigned gmm = g + nb.dnv(rd, mu);
    for (int l = threadIdx.x; l < d; l += kBigBlock)
      vin[l] = {R(a.phi_in[size_t(2 * l) * Vp + gmm]), R(a.phi_in[size_t(2 * l + 1) * Vp + gmm])};
    const Mat<R, N> Ub = load_link<T, R, N>(a.U, Vp, mu, gmm);
    __syncthreads();
    big_apply<R, N, true>(tb, Ub, vin, psi, X);
    for (int l = threadIdx.x; l < d; l += kBigBlock) { fph[l].r += k2 * psi[l].r; fph[l].i += k2 * psi[l].i; }
    __syncthreads();
  }
  if (a.mode == BIG_ACTION) {
    if (threadIdx.x == 0) a.q1[g] = q1s;
    return;
  }
  big_phi_finish<T>(a, rd, g, d, kBigBlock, phx, fph, vin);
}

// On-site part alone (inner steps of the multi-timescale split, F_NO_HOP):
// -2 dV/dphi^* (scalar_higgs.hpp:25-39), pi kick and phi drift; the same
// arithmetic as big_matter_kernel with no hops. A block covers 64 sites with
// kOnsiteGroups waves, wave w taking the components l = w, w + G, ... of every
// site, so that a small lattice still fills the GPU; the sums over a site's
// components (|phi|^2 and, for a frozen scalar, Re phi^dag pi and |pi|^2) meet
// in LDS.
constexpr int kOnsiteGroups = 4;
template <typename T>
struct OnsiteArgs {
  const RepDesc* reps;
  const int* blk_rep;
  unsigned Vp;
  const T* phi_in;
  T* phi_out;
  T* pi;
  int d;
  double ck, cd;
  int flags;
  const T* Fon;
};
constexpr int kOnsiteK = (64 + kOnsiteGroups - 1) / kOnsiteGroups;  // components per thread (d <= 64)

// A thread's components of its site: l = grp + k kOnsiteGroups, k < kOnsiteK.
template <typename T>
struct OnsiteRegs {
  T xr[kOnsiteK], xi[kOnsiteK], pr[kOnsiteK], pm[kOnsiteK];
};

// One on-site sub-step on a thread's registers: the kick of pi with
// -2 dV/dphi^* (or the plugged-in potential's planes Fon), the phi drift, and
// for a frozen scalar the tangent projection and great-circle drift. Products
// are contracted within expressions only, so every caller rounds alike.
template <typename T, class SUM>
GHX_DEV void onsite_substep(OnsiteRegs<T>& v, int d, int grp, bool live, const RepDesc& rd, double ck, double cd, int flags,
                            const T* Fon, unsigned Vp, unsigned g, SUM&& site_sum) {
#pragma clang fp contract(on)
  using R = T;
  constexpr int G = kOnsiteGroups;
  R n2 = R(0);
#pragma unroll
  for (int k = 0; k < kOnsiteK; ++k)
    if (live && grp + k * G < d) n2 += v.xr[k] * v.xr[k] + v.xi[k] * v.xi[k];
  n2 = site_sum(n2);
  const R c = (flags & F_NO_ONSITE) ? R(0) : R(-2) * (R(1) + R(2.0 * rd.lambda) * (n2 - R(1)));
  const bool abs_eps = flags & F_ABS_EPS;
  const R ek = R(abs_eps ? ck : ck * rd.eps), ed = R(abs_eps ? cd : cd * rd.eps);
  const bool frozen = flags & F_FROZEN;
  R rp = R(0);
#pragma unroll
  for (int k = 0; k < kOnsiteK; ++k) {
    const int l = grp + k * G;
    if (!live || l >= d) continue;
    R fr = fma(c, v.xr[k], R(0)), fi = fma(c, v.xi[k], R(0));
    if (Fon) {
      const bool on = !(flags & F_NO_ONSITE);
      fr = on ? R(Fon[size_t(2 * l) * Vp + g]) : R(0);
      fi = on ? R(Fon[size_t(2 * l + 1) * Vp + g]) : R(0);
    }
    if (flags & F_KICK) {
      v.pr[k] = fma(ek, fr, v.pr[k]);
      v.pm[k] = fma(ek, fi, v.pm[k]);
    }
    if (frozen) {
      rp += v.xr[k] * v.pr[k] + v.xi[k] * v.pm[k];
      continue;
    }
    if (flags & F_DRIFT) {
      v.xr[k] = fma(ed, v.pr[k], v.xr[k]);
      v.xi[k] = fma(ed, v.pm[k], v.xi[k]);
    }
  }
  if (!frozen) return;  // uniform over the block
  // |phi| = 1: tangent projection of the kicked pi, then the great-circle drift
  // (gauge_higgs_hmc.hpp:61-103)
  rp = site_sum(rp);
  R r2 = R(0);
#pragma unroll
  for (int k = 0; k < kOnsiteK; ++k) {
    if (!live || grp + k * G >= d) continue;
    if (flags & F_KICK) {
      v.pr[k] -= rp * v.xr[k];
      v.pm[k] -= rp * v.xi[k];
    }
    r2 += v.pr[k] * v.pr[k] + v.pm[k] * v.pm[k];
  }
  r2 = site_sum(r2);
  if (!(flags & F_DRIFT) || !live) return;
  const R rr = sqrt(r2);
  const R rmin = sizeof(R) == 8 ? R(1e-300) : R(1e-30);
  const bool move = rr >= rmin;
  const R cs = move ? cos(rr * ed) : R(1), sn = move ? sin(rr * ed) : R(0), q = move ? sn / rr : R(0), w = -rr * sn;
#pragma unroll
  for (int k = 0; k < kOnsiteK; ++k) {
    if (grp + k * G >= d) continue;
    const R fr = v.xr[k], fi = v.xi[k], pr = v.pr[k], pm = v.pm[k];
    v.xr[k] = move ? cs * fr + q * pr : fr;
    v.xi[k] = move ? cs * fi + q * pm : fi;
    v.pr[k] = move ? w * fr + cs * pr : pr;
    v.pm[k] = move ? w * fi + cs * pm : pm;
  }
}

// The 64 sites of block blk: the thread's components in, sub-steps, out.
template <typename T, class BODY>
GHX_DEV void onsite_block(const OnsiteArgs<T>& o, unsigned blk, BODY&& body) {
  using R = T;
  constexpr int G = kOnsiteGroups;
  __shared__ R red[G][64];
  const int lane = int(threadIdx.x) % 64, grp = int(threadIdx.x) / 64;
  const unsigned g = blk * 64 + unsigned(lane), Vp = o.Vp;
  const RepDesc& rd = o.reps[o.blk_rep[blk * 64 / kBlock]];
  const bool live = g - rd.offset < rd.vol;  // the block stays whole for its barriers
  auto site_sum = [&](R x) {
    red[grp][lane] = x;
    __syncthreads();
    R t = red[0][lane];
#pragma unroll
    for (int k = 1; k < G; ++k) t += red[k][lane];
    __syncthreads();
    return t;
  };
  OnsiteRegs<T> v;
#pragma unroll
  for (int k = 0; k < kOnsiteK; ++k) {
    const int l = grp + k * G;
    const bool on = live && l < o.d;
    v.xr[k] = on ? R(o.phi_in[size_t(2 * l) * Vp + g]) : R(0);
    v.xi[k] = on ? R(o.phi_in[size_t(2 * l + 1) * Vp + g]) : R(0);
    v.pr[k] = on ? R(o.pi[size_t(2 * l) * Vp + g]) : R(0);
    v.pm[k] = on ? R(o.pi[size_t(2 * l + 1) * Vp + g]) : R(0);
  }
  const bool put = body(v, grp, live, rd, g, site_sum);
  if (!live || !put) return;
#pragma unroll
  for (int k = 0; k < kOnsiteK; ++k) {
    const int l = grp + k * G;
    if (l >= o.d) continue;
    o.pi[size_t(2 * l) * Vp + g] = T(v.pr[k]);
    o.pi[size_t(2 * l + 1) * Vp + g] = T(v.pm[k]);
    o.phi_out[size_t(2 * l) * Vp + g] = T(v.xr[k]);
    o.phi_out[size_t(2 * l + 1) * Vp + g] = T(v.xi[k]);
  }
}

// One workgroup's part of one sub-step: the 64 sites of block blk. pi is
// written back unchanged where the sub-step does not kick, and phi_out gets
// phi where it does not drift (then the caller keeps phi_in as the scalar).
template <typename T>
GHX_DEV void onsite_step_block(const OnsiteArgs<T>& o, unsigned blk) {
  onsite_block<T>(o, blk, [&](OnsiteRegs<T>& v, int grp, bool live, const RepDesc& rd, unsigned g, auto&& site_sum) {
    onsite_substep<T>(v, o.d, grp, live, rd, o.ck, o.cd, o.flags, o.Fon, o.Vp, g, site_sum);
    return true;
  });
}

// All on-site sub-steps of one inner-scale block at once (OnsiteSeq: the
// sub-steps' coefficients and flags), each thread's components held in
// registers between them: the arithmetic of onsite_step_block sub-step by
// sub-step, without the round trips through memory. The host uses it where
// the scalar's inner-scale update does not depend on the links (no surrogate
// force, no plugged-in potential), so that it can run apart from the gauge
// sub-steps. phi_out receives the scalar after the last sub-step.
constexpr int kOnsiteMaxSteps = 48;
struct OnsiteSeq {
  int n;
  int flags[kOnsiteMaxSteps];
  double ck[kOnsiteMaxSteps], cd[kOnsiteMaxSteps];
};
template <typename T>
__global__ __launch_bounds__(64 * kOnsiteGroups) void onsite_seq_kernel(OnsiteArgs<T> o, OnsiteSeq q) {
  onsite_block<T>(o, blockIdx.x, [&](OnsiteRegs<T>& v, int grp, bool live, const RepDesc& rd, unsigned g, auto&& site_sum) {
    for (int s = 0; s < q.n; ++s) onsite_substep<T>(v, o.d, grp, live, rd, q.ck[s], q.cd[s], q.flags[s], nullptr, o.Vp, g, site_sum);
    return true;
  });
}
template <typename T>
__global__ __launch_bounds__(64 * kOnsiteGroups) void onsite_step_kernel(OnsiteArgs<T> o) {
  onsite_step_block<T>(o, blockIdx.x);
}

// The inner-scale step of a large irrep in one launch: workgroups [0, nlink)
// run the link-parallel gauge update (U and P), the rest the on-site update of
// the scalar (phi and pi). The two parts read and write disjoint fields, so
// their latencies overlap instead of adding up in two launches of a few waves
// per CU each; the arithmetic is that of the two separate kernels.
template <typename T, int N, int D>
__global__ __launch_bounds__(kLinkSites * D) void md_link_onsite_kernel(StepArgs<T> a, unsigned nlink, OnsiteArgs<T> o) {
  static_assert(kLinkSites * D == 64 * kOnsiteGroups, "both parts run workgroups of the same size");
  if (blockIdx.x < nlink) md_link_block<T, N, D, RepExt<T, N>>(with_store_policy<false>(a), NoParams{}, xcd_block(blockIdx.x, nlink));
  else onsite_step_block<T>(o, blockIdx.x - nlink);
}

// Per-site FP64 contributions for H with a large rep: q1 from big_matter_kernel
// (BIG_ACTION), the rest streamed plane by plane.
template <typename T, int N, int D>
__global__ __launch_bounds__(kRedBlock) void big_action_kernel(ActArgs a, const T* __restrict__ U, const T* __restrict__ P,
                                                              const T* __restrict__ phi, const T* __restrict__ pi, const double* q1, int d,
                                                              const double* Von) {
  using R = double;
  constexpr int NGN = N * N - 1;
  const int r = a.red_rep[blockIdx.x];
  const RepDesc& rd = a.reps[r];
  const unsigned s = a.red_first[blockIdx.x] - rd.offset + threadIdx.x;
  double q[kNQ] = {0, 0, 0, 0, 0, 0};
  if (s < rd.vol) {
    const unsigned g = rd.offset + s, Vp = a.Vp;
    const Nbr<D> nb(rd, s);
    for (int mu = 0; mu < D; ++mu) {
      const Mat<R, N> Ux = load_link<T, R, N>(U, Vp, mu, g);
      const unsigned gpm = g + nb.upv(rd, mu);
      for (int nu = mu + 1; nu < D; ++nu) {
        const Mat<R, N> A = mul_nn(Ux, load_link<T, R, N>(U, Vp, nu, gpm));
        const Mat<R, N> B = mul_nn(load_link<T, R, N>(U, Vp, nu, g), load_link<T, R, N>(U, Vp, mu, g + nb.upv(rd, nu)));
        q[0] += retr_nd(A, B);
      }
      for (int c = 0; c < NGN; ++c) { const double v = double(P[size_t(mu * NGN + c) * Vp + g]); q[4] += v * v; }
    }
    q[1] = q1[g];
    double n2 = 0.0, p2 = 0.0;
    for (int k = 0; k < 2 * d; ++k) {
      const double x = double(phi[size_t(k) * Vp + g]), y = double(pi[size_t(k) * Vp + g]);
      n2 += x * x;
      p2 += y * y;
    }
    q[2] = Von ? Von[g] : n2 + rd.lambda * (n2 - 1.0) * (n2 - 1.0);
    q[3] = n2;
    q[5] = p2;
  }
  block_reduce_store<kNQ>(q, a.partials + size_t(blockIdx.x) * kNQ);
}

// Multi-invariant on-site term for a runtime irrep dimension d <= kMultiMaxD
// (the large-irrep path): the Hermitian-basis form of multi_onsite_kernel
// (A_R real, d^2 x d^2, on m_k = Tr(E_k phi phi^dag)) with the per-site
// vectors in LDS. A block holds SB sites, one per thread: m (d^2 x SB), phi
// and the accumulated g (d x SB each). t_k = sum_l A_R[k][l] m_l reads A_R
// with uniform loads; each t_k is applied at once (E_k has at most two
// entries), so t is never stored.
constexpr int kMultiMaxD = 16;
template <typename T, int SB>
__global__ __launch_bounds__(SB) void multi_onsite_lds_kernel(const RepDesc* reps, const int* blk_rep, unsigned Vp, int d,
                                                             const T* __restrict__ phi, const T* __restrict__ A, double mu2,
                                                             T* Fon, double* Von) {
  using R = T;
  extern __shared__ unsigned char smem_raw[];
  const int dd = d * d;
  R* m = reinterpret_cast<R*>(smem_raw);                 // [dd][SB]
  cx<R>* ph = reinterpret_cast<cx<R>*>(m + size_t(dd) * SB);  // [d][SB]
  cx<R>* gv = ph + size_t(d) * SB;                      // [d][SB]
  const unsigned g0 = blockIdx.x * SB;                  // SB divides kBlock: one replica per block
  const int r = blk_rep[g0 / kBlock];
  const RepDesc& rd = reps[r];
  const int t = threadIdx.x;
  const unsigned g = g0 + t;
  const bool live = g - rd.offset < rd.vol;
  const R h = R(0.70710678118654752440);
  const R m2 = R(mu2);
  R n2 = R(0);
  for (int i = 0; i < d; ++i) {
    cx<R> x = {R(0), R(0)};
    if (live) x = {R(phi[size_t(2 * i) * Vp + g]), R(phi[size_t(2 * i + 1) * Vp + g])};
    ph[i * SB + t] = x;
    gv[i * SB + t] = {-m2 * x.r, -m2 * x.i};
    m[i * SB + t] = x.r * x.r + x.i * x.i;
    n2 += m[i * SB + t];
  }
  {
    int k = d;
    for (int i = 0; i < d; ++i)
      for (int j = i + 1; j < d; ++j, k += 2) {
        const cx<R> z = cmul(ph[i * SB + t], cconj(ph[j * SB + t]));
        m[k * SB + t] = R(2) * h * z.r;
        m[(k + 1) * SB + t] = R(2) * h * z.i;
      }
  }
  R v2 = R(0);
  for (int k = 0; k < d; ++k) {  // diagonal units
    R tk = R(0);
    const T* row = A + size_t(k) * dd;
    for (int l = 0; l < dd; ++l) tk = fma(R(row[l]), m[l * SB + t], tk);
    v2 = fma(tk, m[k * SB + t], v2);
    cx<R>& gk = gv[k * SB + t];
    gk.r = fma(tk, ph[k * SB + t].r, gk.r);
    gk.i = fma(tk, ph[k * SB + t].i, gk.i);
  }
  {
    int k = d;
    for (int i = 0; i < d; ++i)
      for (int j = i + 1; j < d; ++j, k += 2) {
        R tS = R(0), tK = R(0);
        const T* rS = A + size_t(k) * dd;
        const T* rK = rS + dd;
        for (int l = 0; l < dd; ++l) {
          const R ml = m[l * SB + t];
          tS = fma(R(rS[l]), ml, tS);
          tK = fma(R(rK[l]), ml, tK);
        }
        v2 = fma(tS, m[k * SB + t], fma(tK, m[(k + 1) * SB + t], v2));
        const cx<R> Tij = {h * tS, h * tK};
        cfma(gv[i * SB + t], Tij, ph[j * SB + t]);
        cfma(gv[j * SB + t], cconj(Tij), ph[i * SB + t]);
      }
  }
  if (!live) return;
  if (Fon)
    for (int i = 0; i < d; ++i) {
      Fon[size_t(2 * i) * Vp + g] = T(R(-2) * gv[i * SB + t].r);
      Fon[size_t(2 * i + 1) * Vp + g] = T(R(-2) * gv[i * SB + t].i);
    }
  if (Von) Von[g] = double(-m2 * n2) + 0.5 * double(v2);
}

}  // namespace ghx
