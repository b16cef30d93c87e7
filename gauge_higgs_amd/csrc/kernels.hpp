// gfx950 kernels of the MD step and its reductions, templated on storage type T,
// colours N, dimension D and representation policy REP.
//
// HBM layout (structure of arrays, one plane per real component, each plane
// holds every site of every replica, stride Vp):
//   U   plane (mu*2N*NR + 2(iN+j) + re/im)    links, row-major, rows i < NR
//                                              (NR = N in FP64, N-1 in FP32)
//   P   plane (mu*(N^2-1) + a)                  link momenta (generator order)
//   phi plane (2k + re/im)                      scalar field, W-basis components
// Replicas are laid back to back; each starts on a block boundary so a block
// never straddles two lattices (blk_rep maps block -> replica).
#pragma once
#include <type_traits>

#include "ops.hpp"
#include "reps.hpp"

namespace ghx {

constexpr int kBlock = 128;     // threads per block of the per-site kernels
constexpr int kRedBlock = 256;  // threads per block of the reduction kernels
constexpr int kNQ = 6;          // per-site reduction quantities
#ifndef GHX_NU_UNROLL
#define GHX_NU_UNROLL 0  // unroll the staple loop (more loads in flight, more registers)
#endif
#ifndef GHX_MD_WAVES
#define GHX_MD_WAVES 3  // minimum waves per SIMD requested for the MD-step kernel (FP32)
#endif
// Minimum waves per SIMD of the thread-per-site step. FP64 matrices take twice
// the registers: two waves (256 VGPRs) keep the SU(3) step free of spills. The
// FP32 SU(3) sextet step fits 128 registers without spills and runs 2 % faster
// at four waves; the other FP32 reps spill there and keep three.
template <class REP>
struct md_waves_rep { static constexpr int value = GHX_MD_WAVES; };
template <typename R, int N>
struct md_waves_rep<RepSym2<R, N>> { static constexpr int value = 4; };
template <typename T, class REP = RepNone<T, 3>>
constexpr int md_waves() { return sizeof(T) == 8 ? 2 : md_waves_rep<REP>::value; }

struct RepDesc {
  int L[4];
  unsigned stride[4];
  unsigned offset;  // first global site of this replica (multiple of kBlock)
  unsigned vol;
  double beta, kappa, lambda, tau, eps;
  unsigned long long seed;
  unsigned blk_begin, blk_end;  // reduction blocks of this replica
};

// F_NO_GAUGE / F_NO_HOP / F_NO_ONSITE drop force parts (multi-timescale
// splitting); F_SURR adds surr_sign * S (a position-dependent surrogate link
// force) to the link force.
// F_FROZEN: |phi_x| = 1 scalar (gauge_higgs_hmc.hpp:61-103): pi is projected
// onto the tangent space after each kick and the phi drift is the geodesic.
// F_EXT_ONSITE: the on-site force -2 dV/dphi^* comes from the Fon planes (a
// plugged-in potential, e.g. the multi-invariant one) instead of the quartic.
enum StepFlags : int { F_KICK = 1, F_DRIFT = 2, F_FORCE = 4, F_ABS_EPS = 8, F_NO_GAUGE = 16, F_NO_HOP = 32, F_NO_ONSITE = 64,
                       F_SURR = 128, F_FROZEN = 256, F_EXT_ONSITE = 512,
                       F_STORE_CACHED = 1024 };  // the step's stores with the default cache policy instead of streamed

template <typename T>
struct StepArgs {
  const RepDesc* reps;
  const int* blk_rep;
  unsigned Vp;
  const T* U_in;
  T* U_out;
  T* P;
  const T* phi_in;
  T* phi_out;
  T* pi;
  T* F_out;
  T* Fphi_out;
  double ck, cd;  // kick / drift coefficient (times eps of the replica unless F_ABS_EPS)
  int flags;
  const T* G;     // external matter link force planes (REP_EXT)
  const T* S;     // surrogate link force planes (F_SURR)
  double surr_sign;
  const T* Fon;   // on-site scalar force planes (F_EXT_ONSITE)
  BlockWalk walk{};  // block order of md_step_kernel
};

// Link storage. FP64 (parity mode) stores all N rows. FP32 MD stores the first
// N-1 rows only and rebuilds the last one on load, which is exact for SU(N):
//   SU(2): row 1 = (-conj(U01), conj(U00));  SU(3): row 2 = conj(row 0 x row 1).
// This cuts link traffic by 1/3 (SU(3)) or 1/2 (SU(2)) in the bandwidth-bound
// MD step; the reconstructed matrix is unitary to the rounding of rows 0..N-2.
#ifndef GHX_COMPRESS_LINKS
#define GHX_COMPRESS_LINKS 1  // FP32 links keep N-1 rows
#endif
template <typename T, int N>
constexpr int link_rows() { return (sizeof(T) == 4 && GHX_COMPRESS_LINKS) ? N - 1 : N; }
template <typename T, int N>
constexpr int link_reals() { return 2 * N * link_rows<T, N>(); }

template <typename R, int N>
GHX_DEV void rebuild_last_row(Mat<R, N>& m) {
  if constexpr (N == 2) {
    m.e[1][0] = {-m.e[0][1].r, m.e[0][1].i};
    m.e[1][1] = {m.e[0][0].r, -m.e[0][0].i};
  } else {
    static_assert(N == 3, "compressed links are defined for SU(2) and SU(3)");
    // conj(a_i b_j - a_j b_i), a = row 0, b = row 1: one product and three
    // fused multiply-adds per real part (the rebuild runs on every link load)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int i = (k + 1) % 3, j = (k + 2) % 3;
      const cx<R> ai = m.e[0][i], aj = m.e[0][j], bi = m.e[1][i], bj = m.e[1][j];
      R re = ai.r * bj.r;
      re = fma(-ai.i, bj.i, re);
      re = fma(-aj.r, bi.r, re);
      re = fma(aj.i, bi.i, re);
      R im = aj.r * bi.i;
      im = fma(aj.i, bi.r, im);
      im = fma(-ai.r, bj.i, im);
      im = fma(-ai.i, bj.r, im);
      m.e[2][k] = {re, im};
    }
  }
}

template <typename T, typename R, int N>
GHX_DEV Mat<R, N> load_link(const T* __restrict__ U, unsigned Vp, int mu, unsigned g) {
  constexpr int NR = link_rows<T, N>();
  Mat<R, N> m;
  const T* p = U + size_t(mu) * link_reals<T, N>() * Vp + g;
#pragma unroll
  for (int i = 0; i < NR; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      m.e[i][j].r = R(p[size_t(2 * (i * N + j)) * Vp]);
      m.e[i][j].i = R(p[size_t(2 * (i * N + j) + 1) * Vp]);
    }
  if constexpr (NR < N) rebuild_last_row<R, N>(m);
  return m;
}
template <typename T, typename R, int N>
GHX_DEV void store_link(T* __restrict__ U, unsigned Vp, int mu, unsigned g, const Mat<R, N>& m) {
  constexpr int NR = link_rows<T, N>();
  T* p = U + size_t(mu) * link_reals<T, N>() * Vp + g;
#pragma unroll
  for (int i = 0; i < NR; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      p[size_t(2 * (i * N + j)) * Vp] = T(m.e[i][j].r);
      p[size_t(2 * (i * N + j) + 1) * Vp] = T(m.e[i][j].i);
    }
}
template <typename T, typename R, int DR>
GHX_DEV void load_vec(const T* __restrict__ f, unsigned Vp, unsigned g, cx<R>* v) {
#pragma unroll
  for (int k = 0; k < DR; ++k) { v[k].r = R(f[size_t(2 * k) * Vp + g]); v[k].i = R(f[size_t(2 * k + 1) * Vp + g]); }
}
template <typename T, typename R, int DR>
GHX_DEV void store_vec(T* __restrict__ f, unsigned Vp, unsigned g, const cx<R>* v) {
#pragma unroll
  for (int k = 0; k < DR; ++k) { f[size_t(2 * k) * Vp + g] = T(v[k].r); f[size_t(2 * k + 1) * Vp + g] = T(v[k].i); }
}

// ---- buffer-resource access for the MD-step kernel ------------------------
// Every field is a set of planes; a load is (vector offset = site, scalar
// offset = plane * Vp), so one VGPR addresses all planes of a site and the
// hardware bounds check turns any stray offset into a zero read.
using rsrc_t = __amdgpu_buffer_rsrc_t;
GHX_DEV rsrc_t make_rsrc(const void* p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
// POL is the cache policy of the load: 0, or kLoadNT for data a step reads
// once (a site's own momenta, the external and surrogate link forces): such
// loads stream through L2 and leave it to the neighbour links, which the
// staples read several times. The x-lane step kernels (large SU(3) lattices)
// use it; on the small lattices of the other kernels, where every field fits
// the caches, the default policy is faster (6-9 % per MD step on SU(2)).
constexpr int kLoadNT = 2;
template <typename T, int POL = 0> GHX_DEV T bload(rsrc_t r, unsigned voff, unsigned soff) {
  if constexpr (sizeof(T) == 4) return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, POL));
  else return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, POL));
}
// Stores of the MD step are streamed by default (non-temporal, cache policy
// bit NT = 2): nothing written in a step is read again in that step, so they
// should not evict the neighbour links that the staples re-read through L2.
// Where the next step's reads find the written fields in cache (small
// lattices), the default policy can be faster: F_STORE_CACHED selects it. The
// step kernels fix it per instance (with_store_policy), so the branch below
// folds away; the autotuner tries both.
constexpr int kStoreNT = 2;
GHX_DEV void bstore(float v, rsrc_t r, unsigned voff, unsigned soff, bool nt) {
  if (nt) __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)voff, (int)soff, kStoreNT);
  else __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)voff, (int)soff, 0);
}
GHX_DEV void bstore(double v, rsrc_t r, unsigned voff, unsigned soff, bool nt) {
  using v2 = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
  if (nt) __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2, v), r, (int)voff, (int)soff, kStoreNT);
  else __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2, v), r, (int)voff, (int)soff, 0);
}
// Planes of one field: plane c of site g at byte offset (c*Vp + g)*sizeof(T).
template <typename T>
struct Planes {
  rsrc_t r;
  unsigned pstride;  // Vp * sizeof(T)
  bool nt;           // streamed stores
  GHX_DEV Planes(const T* p, unsigned Vp, unsigned nplanes, bool stream_stores = true)
      : r(make_rsrc(p, nplanes * Vp * unsigned(sizeof(T)))), pstride(Vp * unsigned(sizeof(T))), nt(stream_stores) {}
  GHX_DEV T ld(unsigned g, unsigned c) const { return bload<T>(r, g * unsigned(sizeof(T)), c * pstride); }
  template <int POL> GHX_DEV T ldp(unsigned g, unsigned c) const { return bload<T, POL>(r, g * unsigned(sizeof(T)), c * pstride); }
  GHX_DEV void st(unsigned g, unsigned c, T v) const { bstore(v, r, g * unsigned(sizeof(T)), c * pstride, nt); }
};
// FULL = false leaves the last row of a compressed link unset (for a left
// factor whose product keeps only its first rows).
template <typename T, typename R, int N, bool FULL = true>
GHX_DEV Mat<R, N> bload_link(const Planes<T>& U, int mu, unsigned g) {
  constexpr int NR = link_rows<T, N>(), LR = link_reals<T, N>();
  Mat<R, N> m;
#pragma unroll
  for (int i = 0; i < NR; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      m.e[i][j].r = R(U.ld(g, unsigned(mu * LR + 2 * (i * N + j))));
      m.e[i][j].i = R(U.ld(g, unsigned(mu * LR + 2 * (i * N + j) + 1)));
    }
  if constexpr (NR < N && FULL) rebuild_last_row<R, N>(m);
  return m;
}
template <typename T, typename R, int N>
GHX_DEV void bstore_link(const Planes<T>& U, int mu, unsigned g, const Mat<R, N>& m) {
  constexpr int NR = link_rows<T, N>(), LR = link_reals<T, N>();
#pragma unroll
  for (int i = 0; i < NR; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      U.st(g, unsigned(mu * LR + 2 * (i * N + j)), T(m.e[i][j].r));
      U.st(g, unsigned(mu * LR + 2 * (i * N + j) + 1), T(m.e[i][j].i));
    }
}
template <typename T, typename R, int DR>
GHX_DEV void bload_vec(const Planes<T>& f, unsigned g, cx<R>* v) {
#pragma unroll
  for (int k = 0; k < DR; ++k) { v[k].r = R(f.ld(g, 2 * k)); v[k].i = R(f.ld(g, 2 * k + 1)); }
}
template <typename T, typename R, int DR>
GHX_DEV void bstore_vec(const Planes<T>& f, unsigned g, const cx<R>* v) {
#pragma unroll
  for (int k = 0; k < DR; ++k) { f.st(g, 2 * k, T(v[k].r)); f.st(g, 2 * k + 1, T(v[k].i)); }
}

// XCD-aware block order: the hardware deals consecutive workgroups round-robin
// over the 8 XCDs (each with its own L2). Renumbering so XCD k gets the k-th
// contiguous chunk of blocks keeps every +-mu neighbour block (1, L0*L1/128,
// ... blocks away) in the same L2. Bijective on [0, n).
#ifndef GHX_XCD_REMAP
#define GHX_XCD_REMAP 1
#endif
GHX_DEV unsigned xcd_block(unsigned b, unsigned n) {
#if GHX_XCD_REMAP
  constexpr unsigned kXcd = 8;
  const unsigned q = n / kXcd;
  if (b < q * kXcd) return (b % kXcd) * q + b / kXcd;
#endif
  return b;
}

// Storage-order block of the p-th block in the walk w (see BlockWalk).
GHX_DEV unsigned walk_block(unsigned p, const BlockWalk& w) {
  if (w.order == 0) return p;
  const unsigned span = w.nbt * w.tt;
  const unsigned chunk = p / span, rem = p % span;
  unsigned j = rem / w.tt;
  const unsigned t = chunk * w.tt + rem % w.tt;
  if (w.order == 2) j = (j % w.lz) * w.ngy + j / w.lz;
  return t * w.nbt + j;
}

// Orders LDS accesses across the lanes of one wave: a wave's LDS operations
// complete in issue order, so a wave-scope fence (no s_barrier) suffices.
GHX_DEV void wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Site -> neighbour offsets for a periodic lattice (geometry.hpp:20-38 order:
// direction 0 fastest). up[m]/dn[m] are signed offsets to add to the site.
template <int D>
struct Nbr {
  // Neighbour offsets of one site. They are kept as one mask of boundary flags
  // and computed per direction from the lattice's strides (scalar loads): an
  // array indexed with a runtime direction would live in scratch memory.
  unsigned wrap = 0;  // bit m: x_m = L_m - 1; bit D + m: x_m = 0
  GHX_DEV Nbr(const RepDesc& r, unsigned s) {
#pragma unroll
    for (int m = 0; m < D; ++m) {
      const unsigned xm = (s / r.stride[m]) % unsigned(r.L[m]);
      wrap |= (int(xm) == r.L[m] - 1 ? 1u << m : 0u) | (xm == 0 ? 1u << (D + m) : 0u);
    }
  }
  GHX_DEV int upv(const RepDesc& r, int m) const {  // offset of x + m
    const int st = int(r.stride[m]);
    return (wrap >> m & 1u) ? -(r.L[m] - 1) * st : st;
  }
  GHX_DEV int dnv(const RepDesc& r, int m) const {  // offset of x - m
    const int st = int(r.stride[m]);
    return (wrap >> (D + m) & 1u) ? (r.L[m] - 1) * st : -st;
  }
};

// |phi| = 1 helpers on one site's d components (gauge_higgs_hmc.hpp:63-104).
template <typename R, int DR>
GHX_DEV void tangent_project(const cx<R>* ph, cx<R>* pv) {
  R rp = R(0);
#pragma unroll
  for (int k = 0; k < DR; ++k) rp += ph[k].r * pv[k].r + ph[k].i * pv[k].i;  // Re(phi^dag pi)
#pragma unroll
  for (int k = 0; k < DR; ++k) { pv[k].r -= rp * ph[k].r; pv[k].i -= rp * ph[k].i; }
}
// Great-circle flow by time e: phi' = cos(r e) phi + sin(r e)/r pi, pi' = -r sin(r e) phi + cos(r e) pi.
template <typename R, int DR>
GHX_DEV void geodesic_drift(cx<R>* ph, cx<R>* pv, R e) {
  R r2 = R(0);
#pragma unroll
  for (int k = 0; k < DR; ++k) r2 += pv[k].r * pv[k].r + pv[k].i * pv[k].i;
  const R r = sqrt(r2);
  const R rmin = sizeof(R) == 8 ? R(1e-300) : R(1e-30);
  if (r < rmin) return;
  const R c = cos(r * e), sn = sin(r * e), q = sn / r, w = -r * sn;
#pragma unroll
  for (int k = 0; k < DR; ++k) {
    const cx<R> f = ph[k], p = pv[k];
    ph[k] = {c * f.r + q * p.r, c * f.i + q * p.i};
    pv[k] = {w * f.r + c * p.r, w * f.i + c * p.i};
  }
}

// ---------------------------------------------------------------------------
// Fused MD step: kick (P -= eps F_link, pi += eps F_phi) then drift
// (U' = exp(i eps P) U, phi' = phi + eps pi) for one site's D links and phi.
// With F_FORCE the forces are written out instead (parity / FD tests).
// Reads U_in/phi_in (all neighbours), writes U_out/phi_out (double buffer),
// updates P and pi in place. Gauge force: gauge_wilson.hpp:12-27,71-83;
// matter forces: scalar_higgs.hpp:67-128; drift: gauge_higgs_hmc.hpp:136-147.
// ---------------------------------------------------------------------------
// The site's scalar: phi_x into phx and the on-site term -2 dV/dphi^* into fph
// (the quartic, or the Fon planes of a plugged-in potential).
template <typename T, int DR>
GHX_DEV void site_phi_begin(const StepArgs<T>& a, const RepDesc& rd, unsigned g, cx<T>* phx, cx<T>* fph) {
  using R = T;
  const Planes<T> Ph(a.phi_in, a.Vp, 2 * DR);
  bload_vec<T, R, DR>(Ph, g, phx);
  R n2 = R(0);
#pragma unroll
  for (int k = 0; k < DR; ++k) n2 += phx[k].r * phx[k].r + phx[k].i * phx[k].i;
  if ((a.flags & F_EXT_ONSITE) && !(a.flags & F_NO_ONSITE)) {
    bload_vec<T, R, DR>(Planes<T>(a.Fon, a.Vp, 2 * DR), g, fph);
  } else {
    const R c = (a.flags & F_NO_ONSITE) ? R(0) : R(-2) * (R(1) + R(2.0 * rd.lambda) * (n2 - R(1)));  // -2 dV/dphi^*
#pragma unroll
    for (int k = 0; k < DR; ++k) fph[k] = cscale(phx[k], c);
  }
}

// One link of the fused MD step: link mu at site g. Adds the site's hop terms
// through this link to fph (the scalar force of the site), writes the link
// force (F_FORCE) or kicks P_mu and drifts U_mu.
//
// With a plaquette stash (pst, one thread's column of [D][N^2-1][kBlock] floats
// in LDS, slots zeroed before link 0, links updated in order mu = 0..D-1), the
// forward staples are shared between the two links of each plane: for nu > mu
// the plaquette P = U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag gives link mu
// the force term (beta/N) im_tr(P) and link nu, whose forward staple closes the
// same loop as P^dag, the term -(beta/N) im_tr(P), stashed in slot nu. Link mu
// then skips its forward staples for nu < mu: 6 plaquettes per site instead of
// 12 forward staples, and 18 fewer link loads.
//
// The exact factor a representation policy leaves on the M of fwd_M (reps.hpp).
template <class REP, class = void>
struct m_scale_of { static constexpr int value = 1; };
template <class REP>
struct m_scale_of<REP, std::void_t<decltype(REP::M_SCALE)>> { static constexpr int value = REP::M_SCALE; };

// XL (x-lines in lanes: every wave holds whole lines of direction 0, so the
// sites x - 0 and x + 0 sit in lanes xl.src and xl.dst of the same wave; links
// run in the order 1, ..., D-1, 0): link mu >= 1 forms the plaquette of plane
// (0, mu), P = a b c^dag d^dag with a = U_0(x), b = U_mu(x+0) (rows 0, 1 taken
// from lane x+0, whose own link it is), c = U_0(x+mu), d = U_mu(x), as
// a (b (d c)^dag). Its lower staple (d c)^dag a is the backward staple of link
// mu at x+0 in direction 0, so every lane forms it once for its x+0 neighbour,
// parks rows 0, 1 in its LDS column (the stash slot that link 1 does not use
// in this order, plus four words), and at the end of the staple sum reads its
// own from the column of lane x-0 instead of forming it from three more links.
// A wave's LDS operations complete in issue order, so a wave-scope fence
// orders the lanes' writes before the reads. Link 0, updated last, has all of
// its forward plaquette terms in the stash; its hop terms take phi(x+0) from
// lane x+0 and the backward hop D(U_0(x-0))^dag phi(x-0) as lane x-0 formed it
// from its own link and scalar.
struct XLanes {
  int src = 0, dst = 0;  // lanes of x - 0 and x + 0 (byte addresses for ds_bpermute)
  int src_thread = 0;    // thread of this block that holds x - 0
};
GHX_DEV float lane_take(float v, int src_byte) {
  return __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(src_byte, __builtin_bit_cast(int, v)));
}
GHX_DEV double lane_take(double v, int src_byte) {  // two 32-bit halves (XL runs in FP32; this keeps any caller correct)
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = unsigned(__builtin_amdgcn_ds_bpermute(src_byte, int(unsigned(u))));
  const unsigned hi = unsigned(__builtin_amdgcn_ds_bpermute(src_byte, int(unsigned(u >> 32))));
  return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}
template <typename R, int DR>
GHX_DEV void lane_take_vec(const cx<R>* v, int src_byte, cx<R>* o) {
#pragma unroll
  for (int k = 0; k < DR; ++k) o[k] = {lane_take(v[k].r, src_byte), lane_take(v[k].i, src_byte)};
}
// Stash word of element k (0..11) of the parked lower staple: slot 1, then four
// words past the D slots.
template <int D, int NGN>
constexpr int xl_word(int k) { return k < NGN ? NGN + k : D * NGN + k - NGN; }
template <typename T, int N>
constexpr bool xl_able() { return N == 3 && link_rows<T, N>() == 2; }

template <typename T, int N, int D, class REP, bool XL = false>
GHX_DEV void link_update(const StepArgs<T>& a, const typename REP::template Params<T>& rp, const RepDesc& rd, unsigned g,
                         const Nbr<D>& nb, int mu, const cx<T>* phx, cx<T>* fph, T* pst = nullptr, const XLanes& xl = {}) {
  using R = T;
  constexpr int NC = link_reals<T, N>();
  constexpr int NGN = N * N - 1;
  constexpr int DR = REP::DR;
  const unsigned Vp = a.Vp;
  const bool abs_eps = a.flags & F_ABS_EPS;
  const R ek = R(abs_eps ? a.ck : a.ck * rd.eps);
  const R ed = R(abs_eps ? a.cd : a.cd * rd.eps);
  const R bN = R(rd.beta / N);
  const R k2 = R(2.0 * rd.kappa);
  const R hbN = R(0.5) * bN, hk2 = R(0.5) * k2;  // coefficients of im_tr_gen2 (exact halves)
  const bool snt = !(a.flags & F_STORE_CACHED);
  const Planes<T> Uin(a.U_in, Vp, D * NC), Uout(a.U_out, Vp, D * NC, snt), Pp(a.P, Vp, D * NGN, snt);
  const bool do_gauge = !(a.flags & F_NO_GAUGE), do_hop = !(a.flags & F_NO_HOP);
  const Mat<R, N> Ux = bload_link<T, R, N>(Uin, mu, g);
  const unsigned gpm = g + nb.upv(rd, mu);
  // matter first (its temporaries die before the staple sum is built)
  // force terms of this link besides beta/N U Sigma, in generator components:
  // the matter link force (k2 times that of M) and the shared plaquettes; kept
  // as 8 numbers rather than a 3 x 3 matrix across the staple loop
  R fa[NGN];
#pragma unroll
  for (int c = 0; c < NGN; ++c) fa[c] = R(0);
  constexpr bool kXL = XL && xl_able<T, N>();
  if constexpr (DR > 0) if (do_hop) {
    const Planes<T> Ph(a.phi_in, Vp, 2 * DR);
    {
      cx<R> phy[DR], psi[DR];
      if (kXL && mu == 0) lane_take_vec<R, DR>(phx, xl.dst, phy);  // phi(x+0) of lane x+0
      else bload_vec<T, R, DR>(Ph, gpm, phy);
      Mat<R, N> M;
      REP::fwd_M(Ux, phy, phx, psi, M, rp);
      R gm[NGN];
      im_tr_gen2<R, N>(M, gm);
#pragma unroll
      for (int c = 0; c < NGN; ++c) fa[c] = (hk2 * R(m_scale_of<REP>::value)) * gm[c];
#pragma unroll
      for (int k = 0; k < DR; ++k) { fph[k].r = fma(k2, psi[k].r, fph[k].r); fph[k].i = fma(k2, psi[k].i, fph[k].i); }
    }
    {
      cx<R> chi[DR];
      if (kXL && mu == 0) {  // D(U_0(x))^dag phi_x is the backward hop of x+0: lane x-0 formed this one
        cx<R> own[DR];
        REP::apply_dag(Ux, phx, own, rp);
        lane_take_vec<R, DR>(own, xl.src, chi);
      } else {
        const unsigned gmm = g + nb.dnv(rd, mu);
        cx<R> phb[DR];
        bload_vec<T, R, DR>(Ph, gmm, phb);
        const Mat<R, N> Ub = bload_link<T, R, N>(Uin, mu, gmm);
        REP::apply_dag(Ub, phb, chi, rp);
      }
#pragma unroll
      for (int k = 0; k < DR; ++k) { fph[k].r = fma(k2, chi[k].r, fph[k].r); fph[k].i = fma(k2, chi[k].i, fph[k].i); }
    }
  }
  Mat<R, N> Sig = mzero<R, N>();
  // Compressed links: a product of two SU(N) links is SU(N), so its last row
  // follows from the others (SU(3): conj(row 0 x row 1); SU(2): (-conj(c01),
  // conj(c00))). Forming rows 0..N-2 only skips the left factor's rebuild and
  // 1/N of the product. The SU(2) rule is linear, so it also holds for sums of
  // SU(2) matrices: the staple sum keeps row 0 only and the gauge part of
  // U Sigma as well (half the complex MACs of the gauge force).
  constexpr bool kTopRows = (N == 3 || N == 2) && link_rows<T, N>() == N - 1;
  constexpr bool kQuat = N == 2 && kTopRows;
  constexpr bool kShare = N == 3 && kTopRows;
  if constexpr (kXL) {
    if (do_gauge && mu != 0) {  // plane (0, mu), see above
      const Mat<R, N> A0 = bload_link<T, R, N>(Uin, 0, g);
      Mat<R, N> DC = mul_nn_top(Ux, bload_link<T, R, N>(Uin, 0, gpm));
      rebuild_last_row<R, N>(DC);
      {
        const Mat<R, N> Lo = mul_dn_top(DC, A0);  // (d c)^dag a, for the x+0 neighbour
#pragma unroll
        for (int i = 0; i < N - 1; ++i)
#pragma unroll
          for (int j = 0; j < N; ++j) {
            pst[xl_word<D, NGN>(2 * (i * N + j)) * kBlock] = Lo.e[i][j].r;
            pst[xl_word<D, NGN>(2 * (i * N + j) + 1) * kBlock] = Lo.e[i][j].i;
          }
      }
      Mat<R, N> Bx;  // rows 0, 1 of U_mu(x+0), lane x+0's own link
#pragma unroll
      for (int i = 0; i < N - 1; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) Bx.e[i][j] = {lane_take(Ux.e[i][j].r, xl.dst), lane_take(Ux.e[i][j].i, xl.dst)};
      Mat<R, N> BT = mul_nd_top(Bx, DC);
      rebuild_last_row<R, N>(BT);
      Mat<R, N> Pl = mul_nn_top(A0, BT);
      rebuild_last_row<R, N>(Pl);
      R fp[NGN];
      im_tr_gen2<R, N>(Pl, fp);
#pragma unroll
      for (int c = 0; c < NGN; ++c) {
        fa[c] = fma(-hbN, fp[c], fa[c]);
        pst[c * kBlock] += hbN * fp[c];
      }
    }
  }
#if GHX_NU_UNROLL
#pragma unroll
#else
#pragma unroll 1
#endif
  for (int nu = 0; nu < (do_gauge ? D : 0); ++nu) {
    if (nu == mu) continue;
    if (kXL && mu != 0 && nu == 0) continue;  // plane (0, mu) done above
    if (kShare && pst) {
      if (nu > mu && !(kXL && mu == 0)) {  // the plaquette of plane (mu, nu), for both of its links at x
        Mat<R, N> AB = mul_nd_top(bload_link<T, R, N, false>(Uin, nu, gpm), bload_link<T, R, N>(Uin, mu, g + nb.upv(rd, nu)));
        rebuild_last_row<R, N>(AB);
        Mat<R, N> S = mul_nd_top(AB, bload_link<T, R, N>(Uin, nu, g));
        rebuild_last_row<R, N>(S);
        Mat<R, N> Pl = mul_nn_top(Ux, S);
        rebuild_last_row<R, N>(Pl);
        R fp[NGN];
        im_tr_gen2<R, N>(Pl, fp);
#pragma unroll
        for (int c = 0; c < NGN; ++c) {
          fa[c] = fma(hbN, fp[c], fa[c]);
          pst[(nu * NGN + c) * kBlock] -= hbN * fp[c];
        }
      }
    } else {  // forward staple U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag
      Mat<R, N> AB;
      if constexpr (kTopRows) {  // rows 0, 1 of the product from rows 0, 1 of U_nu(x+mu); row 2 by the SU(3) cross product
        AB = mul_nd_top(bload_link<T, R, N, false>(Uin, nu, gpm), bload_link<T, R, N>(Uin, mu, g + nb.upv(rd, nu)));
        rebuild_last_row<R, N>(AB);
      } else {
        AB = mul_nd(bload_link<T, R, N>(Uin, nu, gpm), bload_link<T, R, N>(Uin, mu, g + nb.upv(rd, nu)));
      }
      if constexpr (kQuat) mac_nd_top(Sig, AB, bload_link<T, R, N>(Uin, nu, g));
      else mac_nd(Sig, AB, bload_link<T, R, N>(Uin, nu, g));
    }
    // backward staple U_nu(x+mu-nu)^dag U_mu(x-nu)^dag U_nu(x-nu) = (U_mu(x-nu) U_nu(x+mu-nu))^dag U_nu(x-nu)
    {
      const unsigned gmn = g + nb.dnv(rd, nu);
      if constexpr (kTopRows) {
        Mat<R, N> GF = mul_nn_top(bload_link<T, R, N, false>(Uin, mu, gmn), bload_link<T, R, N>(Uin, nu, gpm + nb.dnv(rd, nu)));
        rebuild_last_row<R, N>(GF);
        if constexpr (kQuat) mac_dn_top(Sig, GF, bload_link<T, R, N>(Uin, nu, gmn));
        else mac_dn(Sig, GF, bload_link<T, R, N>(Uin, nu, gmn));
      } else {
        const Mat<R, N> AB = mul_dd(bload_link<T, R, N>(Uin, nu, gpm + nb.dnv(rd, nu)), bload_link<T, R, N>(Uin, mu, gmn));
        mac_nn(Sig, AB, bload_link<T, R, N>(Uin, nu, gmn));
      }
    }
  }
  if constexpr (kXL) {
    if (do_gauge && mu != 0) {  // this link's backward staple in direction 0, from lane x-0
      wave_fence();
      const T* src = pst - int(threadIdx.x) + xl.src_thread;
      Mat<R, N> Lb;
#pragma unroll
      for (int i = 0; i < N - 1; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) {
          Lb.e[i][j].r = src[xl_word<D, NGN>(2 * (i * N + j)) * kBlock];
          Lb.e[i][j].i = src[xl_word<D, NGN>(2 * (i * N + j) + 1) * kBlock];
        }
      wave_fence();
      rebuild_last_row<R, N>(Lb);
#pragma unroll
      for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) Sig.e[i][j] = cadd(Sig.e[i][j], Lb.e[i][j]);
    }
  }
  Mat<R, N> X;
  if constexpr (kQuat) {  // Sigma and U Sigma are sums of SU(2) matrices
    rebuild_last_row<R, N>(Sig);
    X = mul_nn_top(Ux, Sig);
    rebuild_last_row<R, N>(X);
  } else {
    X = mul_nn(Ux, Sig);
  }
  R f[NGN];
  im_tr_gen2<R, N>(X, f);
#pragma unroll
  for (int c = 0; c < NGN; ++c) f[c] = fma(hbN, f[c], fa[c]);
  if (kShare && pst && !(kXL && mu == 1)) {  // (with XL, slot 1 holds the parked lower staple)
#pragma unroll
    for (int c = 0; c < NGN; ++c) f[c] += pst[(mu * NGN + c) * kBlock];
  }
  if constexpr (REP::KIND == REP_EXT) {
    if (do_hop) {
      const Planes<T> Gp(a.G, Vp, D * NGN);
#pragma unroll
      for (int c = 0; c < NGN; ++c) f[c] += R(Gp.template ldp<XL ? kLoadNT : 0>(g, unsigned(mu * NGN + c)));
    }
  }
  if (a.flags & F_SURR) {
    const Planes<T> Sp(a.S, Vp, D * NGN);
    const R sg = R(a.surr_sign);
#pragma unroll
    for (int c = 0; c < NGN; ++c) f[c] = fma(sg, R(Sp.template ldp<XL ? kLoadNT : 0>(g, unsigned(mu * NGN + c))), f[c]);
  }
  if (a.flags & F_FORCE) {
    const Planes<T> Fp(a.F_out, Vp, D * NGN);
#pragma unroll
    for (int c = 0; c < NGN; ++c) Fp.st(g, unsigned(mu * NGN + c), T(f[c]));
    return;
  }
  R p[NGN];
#pragma unroll
  for (int c = 0; c < NGN; ++c) p[c] = R(Pp.template ldp<XL ? kLoadNT : 0>(g, unsigned(mu * NGN + c)));
  if (a.flags & F_KICK) {
#pragma unroll
    for (int c = 0; c < NGN; ++c) { p[c] = fma(-ek, f[c], p[c]); Pp.st(g, unsigned(mu * NGN + c), T(p[c])); }
  }
  if (a.flags & F_DRIFT) {
    const Mat<R, N> H = alg_to_mat<R, N>(p, ed);
    const Mat<R, N> E = expi<R, N>(H);
    bstore_link<T, R, N>(Uout, mu, g, mul_nn(E, Ux));
  }
}

// The site's scalar after its links: the F_FORCE output, or the pi kick (with
// the tangent projection of a frozen scalar) and the phi drift. NT: pi is
// loaded with the streaming policy (x-lane kernels).
template <typename T, int DR, bool NT = false>
GHX_DEV void site_phi_end(const StepArgs<T>& a, const RepDesc& rd, unsigned g, cx<T>* phx, const cx<T>* fph) {
  using R = T;
  const unsigned Vp = a.Vp;
  const bool abs_eps = a.flags & F_ABS_EPS;
  const R ek = R(abs_eps ? a.ck : a.ck * rd.eps);
  const R ed = R(abs_eps ? a.cd : a.cd * rd.eps);
  if (a.flags & F_FORCE) { bstore_vec<T, R, DR>(Planes<T>(a.Fphi_out, Vp, 2 * DR), g, fph); return; }
  const bool snt = !(a.flags & F_STORE_CACHED);
  const Planes<T> Pi(a.pi, Vp, 2 * DR, snt);
  cx<R> pv[DR];
#pragma unroll
  for (int k = 0; k < DR; ++k) {
    pv[k].r = R(Pi.template ldp<NT ? kLoadNT : 0>(g, 2 * k));
    pv[k].i = R(Pi.template ldp<NT ? kLoadNT : 0>(g, 2 * k + 1));
  }
  const bool frozen = a.flags & F_FROZEN;
  if (a.flags & F_KICK) {
#pragma unroll
    for (int k = 0; k < DR; ++k) { pv[k].r = fma(ek, fph[k].r, pv[k].r); pv[k].i = fma(ek, fph[k].i, pv[k].i); }
    if (frozen) tangent_project<R, DR>(phx, pv);
  }
  if (a.flags & F_DRIFT) {
    if (frozen) {
      geodesic_drift<R, DR>(phx, pv, ed);
    } else {
#pragma unroll
      for (int k = 0; k < DR; ++k) { phx[k].r = fma(ed, pv[k].r, phx[k].r); phx[k].i = fma(ed, pv[k].i, phx[k].i); }
    }
    bstore_vec<T, R, DR>(Planes<T>(a.phi_out, Vp, 2 * DR, snt), g, phx);
  }
  if ((a.flags & F_KICK) || (frozen && (a.flags & F_DRIFT))) bstore_vec<T, R, DR>(Pi, g, pv);
}

// XL: the caller guarantees that every wave holds whole x-lines (L0 divides 64,
// site = lane mod 64) and that all 64 lanes of a wave are live.
template <typename T, int N, int D, class REP, bool XL = false>
GHX_DEV void site_update(const StepArgs<T>& a, const typename REP::template Params<T>& rp, const RepDesc& rd, unsigned s) {
  constexpr int DR = REP::DR;
  const unsigned g = rd.offset + s;
  const Nbr<D> nb(rd, s);
  cx<T> phx[DR > 0 ? DR : 1], fph[DR > 0 ? DR : 1];
  if constexpr (DR > 0) site_phi_begin<T, DR>(a, rd, g, phx, fph);
  T* pst = nullptr;
  if constexpr (xl_able<T, N>()) {  // shared plaquettes (link_update); XL parks a lower staple in slot 1 and 4 more words
    constexpr int kWords = D * (N * N - 1) + (XL ? 2 * (N - 1) * N - (N * N - 1) : 0);
    __shared__ T stash[kWords * kBlock];
    pst = stash + threadIdx.x;
#pragma unroll
    for (int c = 0; c < D * (N * N - 1); ++c) pst[c * kBlock] = T(0);
  }
  XLanes xl;
  if constexpr (XL && xl_able<T, N>()) {
    const int lane = int(threadIdx.x) & 63, l0 = rd.L[0], line = lane & ~(l0 - 1);
    const int src = line | ((lane + l0 - 1) & (l0 - 1)), dst = line | ((lane + 1) & (l0 - 1));
    xl.src = 4 * src;
    xl.dst = 4 * dst;
    xl.src_thread = (int(threadIdx.x) & ~63) | src;
  }
#pragma unroll 1
  for (int i = 0; i < D; ++i) {
    const int mu = XL ? (i + 1 < D ? i + 1 : 0) : i;
    link_update<T, N, D, REP, XL>(a, rp, rd, g, nb, mu, phx, fph, pst, xl);
  }
  if constexpr (DR > 0) site_phi_end<T, DR, XL>(a, rd, g, phx, fph);
}

// The store policy of a kernel instance (CACHED: F_STORE_CACHED) as a
// constant, so that the stores' policy branch folds away.
template <bool CACHED, typename T>
GHX_DEV StepArgs<T> with_store_policy(StepArgs<T> a) {
  a.flags = CACHED ? (a.flags | F_STORE_CACHED) : (a.flags & ~int(F_STORE_CACHED));
  return a;
}

// One site per thread; blocks renumbered XCD-aware. XL: launched only in
// storage-order or XCD-slice walks of lattices whose x-lines fill waves.
template <typename T, int N, int D, class REP, bool XL = false, bool CACHED = false>
__global__ __launch_bounds__(kBlock, (md_waves<T, REP>())) void md_step_kernel(StepArgs<T> a0, typename REP::template Params<T> rp) {
  const StepArgs<T> a = with_store_policy<CACHED>(a0);
  if constexpr (!XL) {
    if (a.walk.order >= 3) {  // one lattice in 4D tiles (BlockWalk)
      const BlockWalk& w = a.walk;
      unsigned p = xcd_block(blockIdx.x, gridDim.x), i = threadIdx.x;
      const unsigned bx = p % w.nbx, by = (p / w.nbx) % w.nby, bz = (p / (w.nbx * w.nby)) % w.nbz, bt = p / (w.nbx * w.nby * w.nbz);
      const unsigned x = bx * w.tx + i % w.tx, y = by * w.ty + (i / w.tx) % w.ty, z = bz * w.tz + (i / (w.tx * w.ty)) % w.tz;
      const unsigned t = bt * (kBlock / (w.tx * w.ty * w.tz)) + i / (w.tx * w.ty * w.tz);
      site_update<T, N, D, REP>(a, rp, a.reps[0], x + w.L0 * (y + w.L1 * (z + w.L2 * t)));
      return;
    }
  }
  const unsigned bid = walk_block(xcd_block(blockIdx.x, gridDim.x), a.walk);
  const int r = a.blk_rep[bid];
  const RepDesc& rd = a.reps[r];
  const unsigned s = bid * kBlock + threadIdx.x - rd.offset;
  if (s >= rd.vol) return;
  site_update<T, N, D, REP, XL>(a, rp, rd, s);
}

// Streaming variant for a single lattice: a block owns kBlock consecutive
// sites of the (D-1)-dimensional cross-section and walks `seg` consecutive
// slices of the last direction, so the +-t neighbours it needs were touched
// by the same block one or two iterations earlier (cache-resident) instead of
// being re-fetched from HBM. Block b handles cross-section tile (b % ntile),
// segment (b / ntile); ntile % 8 == 0 keeps all segments of a tile on one XCD.
template <typename T, int N, int D, class REP, bool XL = false, bool CACHED = false>
__global__ __launch_bounds__(kBlock, md_waves<T>()) void md_stream_kernel(StepArgs<T> a0, typename REP::template Params<T> rp, unsigned ntile,
                                                                       int seg) {
  const StepArgs<T> a = with_store_policy<CACHED>(a0);
  const RepDesc& rd = a.reps[0];
  const unsigned V3 = rd.stride[D - 1];
  const unsigned tile = blockIdx.x % ntile, sg = blockIdx.x / ntile;
  const unsigned s3 = tile * kBlock + threadIdx.x;
  if (s3 >= V3) return;
  const int Lt = rd.L[D - 1];
  const int t0 = int(sg) * seg;
  const int t1 = t0 + seg < Lt ? t0 + seg : Lt;
#pragma unroll 1
  for (int t = t0; t < t1; ++t) site_update<T, N, D, REP, XL>(a, rp, rd, s3 + unsigned(t) * V3);
}

// Link-parallel variant for lattices too small to fill the GPU with one thread
// per site: a workgroup of D waves covers 64 sites, wave mu runs link_update
// for link mu of every site (so mu is uniform within a wave and every access
// stays a 256-byte row of a plane). The waves' hop terms of the scalar force
// meet in LDS; wave 0 adds them in direction order, then the on-site term,
// and updates the site's scalar. D times the waves of md_step_kernel to hide
// load latency; the same arithmetic except for the order in which the scalar
// force terms are summed.
constexpr int kLinkSites = 64;  // sites per workgroup (one wave per direction)
// One workgroup's part: the kLinkSites sites of block bid (storage order).
template <typename T, int N, int D, class REP>
GHX_DEV void md_link_block(const StepArgs<T>& a, const typename REP::template Params<T>& rp, unsigned bid) {
  constexpr int DR = REP::DR;
  const int mu = __builtin_amdgcn_readfirstlane(int(threadIdx.x) / kLinkSites);
  const int lane = int(threadIdx.x) % kLinkSites;
  const unsigned gsite = bid * kLinkSites + unsigned(lane);
  const int r = a.blk_rep[gsite / kBlock];
  const RepDesc& rd = a.reps[r];
  const unsigned s = gsite - rd.offset;
  const bool live = s < rd.vol;
  cx<T> phx[DR > 0 ? DR : 1], fph[DR > 0 ? DR : 1];
#pragma unroll
  for (int k = 0; k < (DR > 0 ? DR : 1); ++k) fph[k] = {T(0), T(0)};
  if (live) {
    const Nbr<D> nb(rd, s);
    if constexpr (DR > 0) bload_vec<T, T, DR>(Planes<T>(a.phi_in, a.Vp, 2 * DR), gsite, phx);
    link_update<T, N, D, REP>(a, rp, rd, gsite, nb, mu, phx, fph);
  }
  if constexpr (DR > 0) {
    __shared__ cx<T> part[D][DR][kLinkSites];
#pragma unroll
    for (int k = 0; k < DR; ++k) part[mu][k][lane] = fph[k];
    __syncthreads();
    if (mu != 0 || !live) return;
    cx<T> on[DR];
    site_phi_begin<T, DR>(a, rd, gsite, phx, on);
#pragma unroll
    for (int k = 0; k < DR; ++k) {
      cx<T> t = part[0][k][lane];
#pragma unroll
      for (int m = 1; m < D; ++m) { t.r += part[m][k][lane].r; t.i += part[m][k][lane].i; }
      fph[k] = {on[k].r + t.r, on[k].i + t.i};
    }
    site_phi_end<T, DR>(a, rd, gsite, phx, fph);
  }
}
template <typename T, int N, int D, class REP, bool CACHED = false>
__global__ __launch_bounds__(kLinkSites * D) void md_link_kernel(StepArgs<T> a, typename REP::template Params<T> rp) {
  md_link_block<T, N, D, REP>(with_store_policy<CACHED>(a), rp, xcd_block(blockIdx.x, gridDim.x));
}

// Link-parallel step with the staples split over planes, for small lattices
// and no scalar components in the kernel (pure gauge, or the external matter
// link force of a large irrep): lane p of link (x, mu) forms the forward and
// backward staples of the p-th plane nu != mu, so a workgroup of kLinkSites
// sites runs D (D - 1) waves, (D - 1) times the waves of md_link_kernel. The
// planes' staple sums meet in LDS and the wave of plane 0 finishes every link
// as link_update does (U Sigma, the force with G and S, the P kick and the
// drift). Each plane's two staples are summed first and the planes then in
// order, so the rounding differs from md_link_kernel's.
template <typename T, int N, int D, class REP, bool CACHED = false>
__global__ __launch_bounds__(kLinkSites * D * (D - 1)) void md_link3_kernel(StepArgs<T> a0, typename REP::template Params<T> rp) {
  static_assert(REP::DR == 0, "the plane-split step carries no scalar");
  using R = T;
  constexpr int NC = link_reals<T, N>(), NGN = N * N - 1, NP = D - 1;
  constexpr bool kTopRows = (N == 3 || N == 2) && link_rows<T, N>() == N - 1;
  constexpr bool kQuat = N == 2 && kTopRows;
  const StepArgs<T> a = with_store_policy<CACHED>(a0);
  const int w = __builtin_amdgcn_readfirstlane(int(threadIdx.x) / kLinkSites);
  const int mu = w / NP, p = w % NP, nu = p < mu ? p : p + 1;
  const int lane = int(threadIdx.x) % kLinkSites;
  const unsigned g = xcd_block(blockIdx.x, gridDim.x) * kLinkSites + unsigned(lane), Vp = a.Vp;
  const RepDesc& rd = a.reps[a.blk_rep[g / kBlock]];
  const unsigned s = g - rd.offset;
  const bool live = s < rd.vol;
  const bool do_gauge = !(a.flags & F_NO_GAUGE);
  __shared__ R part[D][NP > 1 ? NP - 1 : 1][2 * N * N][kLinkSites];
  const Planes<T> Uin(a.U_in, Vp, D * NC);
  Mat<R, N> Sig = mzero<R, N>();
  if (live && do_gauge) {
    const Nbr<D> nb(rd, s);
    const unsigned gpm = g + nb.upv(rd, mu);
    // forward staple U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag
    Mat<R, N> AB;
    if constexpr (kTopRows) {
      AB = mul_nd_top(bload_link<T, R, N, false>(Uin, nu, gpm), bload_link<T, R, N>(Uin, mu, g + nb.upv(rd, nu)));
      rebuild_last_row<R, N>(AB);
    } else {
      AB = mul_nd(bload_link<T, R, N>(Uin, nu, gpm), bload_link<T, R, N>(Uin, mu, g + nb.upv(rd, nu)));
    }
    if constexpr (kQuat) mac_nd_top(Sig, AB, bload_link<T, R, N>(Uin, nu, g));
    else mac_nd(Sig, AB, bload_link<T, R, N>(Uin, nu, g));
    // backward staple (U_mu(x-nu) U_nu(x+mu-nu))^dag U_nu(x-nu)
    const unsigned gmn = g + nb.dnv(rd, nu);
    if constexpr (kTopRows) {
      Mat<R, N> GF = mul_nn_top(bload_link<T, R, N, false>(Uin, mu, gmn), bload_link<T, R, N>(Uin, nu, gpm + nb.dnv(rd, nu)));
      rebuild_last_row<R, N>(GF);
      if constexpr (kQuat) mac_dn_top(Sig, GF, bload_link<T, R, N>(Uin, nu, gmn));
      else mac_dn(Sig, GF, bload_link<T, R, N>(Uin, nu, gmn));
    } else {
      const Mat<R, N> B2 = mul_dd(bload_link<T, R, N>(Uin, nu, gpm + nb.dnv(rd, nu)), bload_link<T, R, N>(Uin, mu, gmn));
      mac_nn(Sig, B2, bload_link<T, R, N>(Uin, nu, gmn));
    }
  }
  if (p > 0) {
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) {
        part[mu][p - 1][2 * (i * N + j)][lane] = Sig.e[i][j].r;
        part[mu][p - 1][2 * (i * N + j) + 1][lane] = Sig.e[i][j].i;
      }
  }
  __syncthreads();
  if (p > 0 || !live) return;
#pragma unroll
  for (int q = 0; q < NP - 1; ++q)
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) {
        Sig.e[i][j].r += part[mu][q][2 * (i * N + j)][lane];
        Sig.e[i][j].i += part[mu][q][2 * (i * N + j) + 1][lane];
      }
  // the rest of link_update for a link without scalar terms
  const bool abs_eps = a.flags & F_ABS_EPS;
  const R ek = R(abs_eps ? a.ck : a.ck * rd.eps);
  const R ed = R(abs_eps ? a.cd : a.cd * rd.eps);
  const R hbN = R(0.5) * R(rd.beta / N);
  const bool snt = !(a.flags & F_STORE_CACHED);
  const Planes<T> Uout(a.U_out, Vp, D * NC, snt), Pp(a.P, Vp, D * NGN, snt);
  const Mat<R, N> Ux = bload_link<T, R, N>(Uin, mu, g);
  Mat<R, N> X;
  if constexpr (kQuat) {
    rebuild_last_row<R, N>(Sig);
    X = mul_nn_top(Ux, Sig);
    rebuild_last_row<R, N>(X);
  } else {
    X = mul_nn(Ux, Sig);
  }
  R f[NGN];
  im_tr_gen2<R, N>(X, f);
#pragma unroll
  for (int c = 0; c < NGN; ++c) f[c] = hbN * f[c];
  if constexpr (REP::KIND == REP_EXT) {
    if (!(a.flags & F_NO_HOP)) {
      const Planes<T> Gp(a.G, Vp, D * NGN);
#pragma unroll
      for (int c = 0; c < NGN; ++c) f[c] += R(Gp.template ldp<0>(g, unsigned(mu * NGN + c)));
    }
  }
  if (a.flags & F_SURR) {
    const Planes<T> Sp(a.S, Vp, D * NGN);
    const R sg = R(a.surr_sign);
#pragma unroll
    for (int c = 0; c < NGN; ++c) f[c] = fma(sg, R(Sp.template ldp<0>(g, unsigned(mu * NGN + c))), f[c]);
  }
  if (a.flags & F_FORCE) {
    const Planes<T> Fp(a.F_out, Vp, D * NGN);
#pragma unroll
    for (int c = 0; c < NGN; ++c) Fp.st(g, unsigned(mu * NGN + c), T(f[c]));
    return;
  }
  R pm[NGN];
#pragma unroll
  for (int c = 0; c < NGN; ++c) pm[c] = R(Pp.template ldp<0>(g, unsigned(mu * NGN + c)));
  if (a.flags & F_KICK) {
#pragma unroll
    for (int c = 0; c < NGN; ++c) { pm[c] = fma(-ek, f[c], pm[c]); Pp.st(g, unsigned(mu * NGN + c), T(pm[c])); }
  }
  if (a.flags & F_DRIFT) {
    const Mat<R, N> H = alg_to_mat<R, N>(pm, ed);
    const Mat<R, N> E = expi<R, N>(H);
    bstore_link<T, R, N>(Uout, mu, g, mul_nn(E, Ux));
  }
}

// Drift alone (U' = exp(i eps P) U, phi' = phi + eps pi) — for the kick/drift API.
template <typename T, int N, int D, int DR>
__global__ __launch_bounds__(kBlock) void drift_kernel(StepArgs<T> a) {
  using R = T;
  constexpr int NGN = N * N - 1;
  const int r = a.blk_rep[blockIdx.x];
  const RepDesc& rd = a.reps[r];
  const unsigned s = blockIdx.x * kBlock + threadIdx.x - rd.offset;
  if (s >= rd.vol) return;
  const unsigned g = rd.offset + s, Vp = a.Vp;
  const R ed = R((a.flags & F_ABS_EPS) ? a.cd : a.cd * rd.eps);
#pragma unroll 1
  for (int mu = 0; mu < D; ++mu) {
    R p[NGN];
    const T* Pm = a.P + size_t(mu) * NGN * Vp + g;
#pragma unroll
    for (int c = 0; c < NGN; ++c) p[c] = R(Pm[size_t(c) * Vp]);
    const Mat<R, N> Ux = load_link<T, R, N>(a.U_in, Vp, mu, g);
    store_link<T, R, N>(a.U_out, Vp, mu, g, mul_nn(expi<R, N>(alg_to_mat<R, N>(p, ed)), Ux));
  }
  if constexpr (DR > 0) {
    cx<R> ph[DR], pv[DR];
    load_vec<T, R, DR>(a.phi_in, Vp, g, ph);
    load_vec<T, R, DR>(a.pi, Vp, g, pv);
    if (a.flags & F_FROZEN) {
      geodesic_drift<R, DR>(ph, pv, ed);
      store_vec<T, R, DR>(a.pi, Vp, g, pv);
    } else {
#pragma unroll
      for (int k = 0; k < DR; ++k) { ph[k].r = fma(ed, pv[k].r, ph[k].r); ph[k].i = fma(ed, pv[k].i, ph[k].i); }
    }
    store_vec<T, R, DR>(a.phi_out, Vp, g, ph);
  }
}

// |phi_x| = 1 scalar, runtime d (gauge_higgs_hmc.hpp:63-84): mode 0 projects pi
// onto the tangent space at phi, mode 1 normalises phi.
template <typename T>
__global__ __launch_bounds__(kBlock) void frozen_site_kernel(const RepDesc* reps, const int* blk_rep, unsigned Vp, T* phi, T* pi, int d,
                                                            int mode) {
  const int r = blk_rep[blockIdx.x];
  const RepDesc& rd = reps[r];
  const unsigned s = blockIdx.x * kBlock + threadIdx.x - rd.offset;
  if (s >= rd.vol) return;
  const unsigned g = rd.offset + s;
  T acc = T(0);
  for (int k = 0; k < d; ++k) {
    const T xr = phi[size_t(2 * k) * Vp + g], xi = phi[size_t(2 * k + 1) * Vp + g];
    if (mode == 0) acc += xr * pi[size_t(2 * k) * Vp + g] + xi * pi[size_t(2 * k + 1) * Vp + g];
    else acc += xr * xr + xi * xi;
  }
  if (mode == 0) {
    for (int k = 0; k < d; ++k) {
      pi[size_t(2 * k) * Vp + g] -= acc * phi[size_t(2 * k) * Vp + g];
      pi[size_t(2 * k + 1) * Vp + g] -= acc * phi[size_t(2 * k + 1) * Vp + g];
    }
  } else if (acc > T(0)) {
    const T inv = T(1) / sqrt(acc);
    for (int k = 0; k < 2 * d; ++k) phi[size_t(k) * Vp + g] *= inv;
  }
}

// Multi-invariant on-site term (scalar_invariants.hpp:62-87, 119-146):
// g = -mu2 phi + A(M) phi with M = phi phi^dag, V = -mu2 |phi|^2 + Tr(M A(M)) / 2.
// A maps Hermitian matrices to Hermitian ones, so it acts as a real d^2 x d^2
// matrix A_R on the coordinates m_k = Tr(E_k M) in an orthonormal Hermitian
// basis (k < d: diagonal units; then per pair i < j, row-major, S and K with
// m_S = sqrt2 Re M_ij, m_K = sqrt2 Im M_ij). t = A_R m gives T = A(M):
// T_ii = t_i, T_ij = (t_S + i t_K)/sqrt2, T_ji = conj(T_ij). That is d^4 real
// FMA per site instead of d^4 complex ones; A_R is read with uniform (scalar)
// loads. Writes Fon = -2 g (the on-site force planes) and/or Von = V (FP64).
template <typename T, int DR>
__global__ __launch_bounds__(kBlock) void multi_onsite_kernel(const RepDesc* reps, const int* blk_rep, unsigned Vp,
                                                             const T* __restrict__ phi, const T* __restrict__ A, double mu2,
                                                             T* Fon, double* Von) {
  using R = T;
  constexpr int DD = DR * DR;
  const int r = blk_rep[blockIdx.x];
  const RepDesc& rd = reps[r];
  const unsigned s = blockIdx.x * kBlock + threadIdx.x - rd.offset;
  if (s >= rd.vol) return;
  const unsigned g = rd.offset + s;
  const R h = R(0.70710678118654752440);  // 1/sqrt2
  cx<R> ph[DR];
  load_vec<T, R, DR>(phi, Vp, g, ph);
  R m[DD];
  R n2 = R(0);
#pragma unroll
  for (int i = 0; i < DR; ++i) {
    m[i] = ph[i].r * ph[i].r + ph[i].i * ph[i].i;
    n2 += m[i];
  }
  {
    int k = DR;
#pragma unroll
    for (int i = 0; i < DR; ++i)
#pragma unroll
      for (int j = i + 1; j < DR; ++j, k += 2) {
        const cx<R> z = cmul(ph[i], cconj(ph[j]));
        m[k] = R(2) * h * z.r;
        m[k + 1] = R(2) * h * z.i;
      }
  }
  const R m2 = R(mu2);
  cx<R> gv[DR];
#pragma unroll
  for (int i = 0; i < DR; ++i) gv[i] = {-m2 * ph[i].r, -m2 * ph[i].i};
  R v2 = R(0);
#pragma unroll
  for (int k = 0; k < DR; ++k) {  // diagonal units
    R t = R(0);
#pragma unroll
    for (int l = 0; l < DD; ++l) t = fma(A[k * DD + l], m[l], t);
    v2 = fma(t, m[k], v2);
    gv[k].r = fma(t, ph[k].r, gv[k].r);
    gv[k].i = fma(t, ph[k].i, gv[k].i);
  }
  {
    int k = DR;
#pragma unroll
    for (int i = 0; i < DR; ++i)
#pragma unroll
      for (int j = i + 1; j < DR; ++j, k += 2) {
        R tS = R(0), tK = R(0);
#pragma unroll
        for (int l = 0; l < DD; ++l) {
          tS = fma(A[k * DD + l], m[l], tS);
          tK = fma(A[(k + 1) * DD + l], m[l], tK);
        }
        v2 = fma(tS, m[k], fma(tK, m[k + 1], v2));
        const cx<R> Tij = {h * tS, h * tK};
        cfma(gv[i], Tij, ph[j]);
        cfma(gv[j], cconj(Tij), ph[i]);
      }
  }
  if (Fon) {
#pragma unroll
    for (int i = 0; i < DR; ++i) {
      Fon[size_t(2 * i) * Vp + g] = T(R(-2) * gv[i].r);
      Fon[size_t(2 * i + 1) * Vp + g] = T(R(-2) * gv[i].i);
    }
  }
  if (Von) Von[g] = double(-m2 * n2) + 0.5 * double(v2);
}

// phi' = phi + eps pi for a runtime number of planes (large reps).
template <typename T>
__global__ __launch_bounds__(kBlock) void phi_drift_kernel(const RepDesc* reps, const int* blk_rep, unsigned Vp, const T* phi_in,
                                                          T* phi_out, T* pi, int nplanes, double cd, int abs_eps, int frozen) {
  const int r = blk_rep[blockIdx.x];
  const RepDesc& rd = reps[r];
  const unsigned s = blockIdx.x * kBlock + threadIdx.x - rd.offset;
  if (s >= rd.vol) return;
  const unsigned g = rd.offset + s;
  const T e = T(abs_eps ? cd : cd * rd.eps);
  if (!frozen) {
    for (int c = 0; c < nplanes; ++c) phi_out[size_t(c) * Vp + g] = phi_in[size_t(c) * Vp + g] + e * pi[size_t(c) * Vp + g];
    return;
  }
  // |phi| = 1: great-circle drift of phi and pi (gauge_higgs_hmc.hpp:84-103)
  T r2 = T(0);
  for (int c = 0; c < nplanes; ++c) r2 += pi[size_t(c) * Vp + g] * pi[size_t(c) * Vp + g];
  const T rr = sqrt(r2);
  const T rmin = sizeof(T) == 8 ? T(1e-300) : T(1e-30);
  const bool move = rr >= rmin;
  const T cs = move ? cos(rr * e) : T(1), sn = move ? sin(rr * e) : T(0), q = move ? sn / rr : T(0), w = -rr * sn;
  for (int c = 0; c < nplanes; ++c) {
    const T fv = phi_in[size_t(c) * Vp + g], pv = pi[size_t(c) * Vp + g];
    phi_out[size_t(c) * Vp + g] = move ? cs * fv + q * pv : fv;
    pi[size_t(c) * Vp + g] = move ? w * fv + cs * pv : pv;
  }
}

// ---------------------------------------------------------------------------
// Deterministic FP64 block reduction of kNQ quantities per thread.
// ---------------------------------------------------------------------------
template <int NQ>
GHX_DEV void block_reduce_store(double (&q)[NQ], double* __restrict__ out) {
  __shared__ double sh[NQ][kRedBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    double v = q[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (lane == 0) sh[k][wid] = v;
  }
  __syncthreads();
  if (threadIdx.x < NQ) {
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < kRedBlock / 64; ++w) v += sh[threadIdx.x][w];
    out[threadIdx.x] = v;
  }
}

struct ActArgs {
  const RepDesc* reps;
  const int* red_rep;  // reduction block -> replica
  const unsigned* red_first;  // reduction block -> first global site
  unsigned Vp;
  double* partials;  // [nblocks][kNQ]
  const double* Von = nullptr;  // per-site potential of a plugged-in potential (else the quartic)
};

// Per-site FP64 contributions to H and the observables:
//  q0 sum_{mu<nu} Re Tr U_pl (gauge_wilson.hpp:32-65)
//  q1 sum_mu Re phi_x^dag D(U_mu(x)) phi_{x+mu} (scalar_higgs.hpp:42-58, 130-144)
//  q2 V(phi_x), q3 |phi_x|^2 (scalar_higgs.hpp:25-39, 146-154)
//  q4 sum P^2 (fields.hpp:69-74), q5 sum |pi|^2 (scalar_field.hpp:53-57)
template <typename T, int N, int D, class REPD>
__global__ __launch_bounds__(kRedBlock) void action_kernel(ActArgs a, const T* __restrict__ U, const T* __restrict__ P,
                                                          const T* __restrict__ phi, const T* __restrict__ pi,
                                                          typename REPD::template Params<double> rp) {
  using R = double;
  constexpr int NGN = N * N - 1;
  constexpr int DR = REPD::DR;
  const int r = a.red_rep[blockIdx.x];
  const RepDesc& rd = a.reps[r];
  const unsigned s = a.red_first[blockIdx.x] - rd.offset + threadIdx.x;
  double q[kNQ] = {0, 0, 0, 0, 0, 0};
  if (s < rd.vol) {
    const unsigned g = rd.offset + s, Vp = a.Vp;
    const Nbr<D> nb(rd, s);
#pragma unroll 1
    for (int mu = 0; mu < D; ++mu) {
      const Mat<R, N> Ux = load_link<T, R, N>(U, Vp, mu, g);
      const unsigned gpm = g + nb.upv(rd, mu);
#pragma unroll 1
      for (int nu = mu + 1; nu < D; ++nu) {
        const Mat<R, N> A = mul_nn(Ux, load_link<T, R, N>(U, Vp, nu, gpm));
        const Mat<R, N> B = mul_nn(load_link<T, R, N>(U, Vp, nu, g), load_link<T, R, N>(U, Vp, mu, g + nb.upv(rd, nu)));
        q[0] += retr_nd(A, B);  // Re Tr(U_mu U_nu(x+mu) (U_nu U_mu(x+nu))^dag)
      }
      const T* Pm = P + size_t(mu) * NGN * Vp + g;
#pragma unroll
      for (int c = 0; c < NGN; ++c) { const double v = double(Pm[size_t(c) * Vp]); q[4] += v * v; }
      if constexpr (DR > 0) {
        cx<R> px[DR], py[DR], dy[DR];
        load_vec<T, R, DR>(phi, Vp, g, px);
        load_vec<T, R, DR>(phi, Vp, gpm, py);
        REPD::apply(Ux, py, dy, rp);
#pragma unroll
        for (int k = 0; k < DR; ++k) q[1] += px[k].r * dy[k].r + px[k].i * dy[k].i;
      }
    }
    if constexpr (DR > 0) {
      cx<R> px[DR], pv[DR];
      load_vec<T, R, DR>(phi, Vp, g, px);
      load_vec<T, R, DR>(pi, Vp, g, pv);
      double n2 = 0.0, p2 = 0.0;
#pragma unroll
      for (int k = 0; k < DR; ++k) { n2 += px[k].r * px[k].r + px[k].i * px[k].i; p2 += pv[k].r * pv[k].r + pv[k].i * pv[k].i; }
      q[2] = a.Von ? a.Von[g] : n2 + rd.lambda * (n2 - 1.0) * (n2 - 1.0);
      q[3] = n2;
      q[5] = p2;
    }
  }
  block_reduce_store<kNQ>(q, a.partials + size_t(blockIdx.x) * kNQ);
}

// Sum the partials of each replica in a fixed order: one block per replica.
static __global__ __launch_bounds__(kRedBlock) void finalize_kernel(const RepDesc* reps, const double* partials, int nq, double* sums) {
  const RepDesc& rd = reps[blockIdx.x];
  __shared__ double sh[kRedBlock];
  for (int k = 0; k < nq; ++k) {
    double v = 0.0;
    for (unsigned b = rd.blk_begin + threadIdx.x; b < rd.blk_end; b += kRedBlock) v += partials[size_t(b) * nq + k];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = kRedBlock / 2; o > 0; o >>= 1) {
      if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) sums[size_t(blockIdx.x) * nq + k] = sh[0];
    __syncthreads();
  }
}

// Hamiltonian parts from the sums: [KE_P, KE_pi, S_g, S_H, H] (gauge_higgs_hmc.hpp:105-110, hmc.hpp:32)
static __global__ void hamiltonian_kernel(const RepDesc* reps, int nrep, int N, int D, int has_phi, const double* sums, double* H) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrep) return;
  const RepDesc& rd = reps[r];
  const double* q = sums + size_t(r) * kNQ;
  const double nplaq = double(rd.vol) * (D * (D - 1) / 2);
  const double KP = 0.5 * q[4];
  const double Kpi = has_phi ? 0.5 * q[5] : 0.0;
  const double Sg = (rd.beta / N) * (double(N) * nplaq - q[0]);
  const double SH = has_phi ? (q[2] - 2.0 * rd.kappa * q[1]) : 0.0;
  double* h = H + size_t(r) * 5;
  h[0] = KP; h[1] = Kpi; h[2] = Sg; h[3] = SH; h[4] = KP + Kpi + Sg + SH;
}

// Metropolis accept/reject per replica, FP64 (gauge_higgs_hmc.hpp:183-191, hmc.hpp:83-90).
struct McState {
  unsigned long long traj_count, accept_count;
  double last_dH;
  int accepted;
  int pad;
};
static __global__ void metropolis_kernel(const RepDesc* reps, int nrep, const double* Hi, const double* Hf, McState* st,
                                  unsigned long long tag) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrep) return;
  McState& m = st[r];
  const double dH = Hf[size_t(r) * 5 + 4] - Hi[size_t(r) * 5 + 4];
  m.last_dH = dH;
  m.traj_count += 1;
  const double u = draw_uniform(reps[r].seed, stream_key(tag, m.traj_count));
  const bool acc = (dH <= 0.0) || (u < exp(-dH));
  m.accepted = acc ? 1 : 0;
  if (acc) m.accept_count += 1;
}

// Replica-exchange decisions for one step (replica_tempering.hpp:157-194),
// all on one thread in the reference order: even pass then odd pass, pairs in
// list order. Delta = S_i(C_j)+S_j(C_i)-S_i(C_i)-S_j(C_j) from the per-replica
// sums q0 = sum Re Tr U_pl and q1 = sum Re phi^dag D phi' (beta and kappa
// ladders; lambda equal): (beta_i-beta_j)/N (q0_i-q0_j) + 2(kappa_i-kappa_j)(q1_i-q1_j).
// Accept if Delta <= 0 (the coin counter still advances), else with
// uniform(seed, key(0x57A9, ++counter)) < exp(-Delta).
struct SwapState {
  unsigned long long seed, counter;
};
static __global__ void swap_decide_kernel(const RepDesc* reps, int N, const double* sums, int nrep, const int* pairs, int npairs,
                                          int nfirst, int* cfg, double* q0, double* q1, SwapState* st,
                                          unsigned long long* attempts, unsigned long long* accepts, int* flags) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int r = 0; r < nrep; ++r) { cfg[r] = r; q0[r] = sums[size_t(r) * kNQ + 0]; q1[r] = sums[size_t(r) * kNQ + 1]; }
  for (int k = 0; k < npairs; ++k) {
    const int i = pairs[2 * k], j = pairs[2 * k + 1];
    const int ci = cfg[i], cj = cfg[j];
    const double Delta = (reps[i].beta - reps[j].beta) / N * (q0[ci] - q0[cj]) + 2.0 * (reps[i].kappa - reps[j].kappa) * (q1[ci] - q1[cj]);
    attempts[k] += 1;
    bool acc = Delta <= 0.0;
    if (!acc) {
      st->counter += 1;
      acc = draw_uniform(st->seed, stream_key(0x57A9, st->counter)) < exp(-Delta);
    } else {
      st->counter += 1;
    }
    flags[k] = acc ? 1 : 0;
    if (acc) { accepts[k] += 1; cfg[i] = cj; cfg[j] = ci; }
    (void)nfirst;
  }
}
// Exchange U and phi site ranges of the accepted pairs in [k0, k1) of the pair list.
template <typename T>
__global__ __launch_bounds__(kBlock) void swap_apply_kernel(const RepDesc* reps, const int* pairs, const int* flags, int k0, unsigned Vp,
                                                           T* U, int nU, T* phi, int nphi) {
  const int k = k0 + blockIdx.y;
  if (!flags[k]) return;
  const RepDesc &ri = reps[pairs[2 * k]], &rj = reps[pairs[2 * k + 1]];
  const unsigned s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= ri.vol) return;
  const unsigned gi = ri.offset + s, gj = rj.offset + s;
  for (int c = 0; c < nU; ++c) { const T t = U[size_t(c) * Vp + gi]; U[size_t(c) * Vp + gi] = U[size_t(c) * Vp + gj]; U[size_t(c) * Vp + gj] = t; }
  for (int c = 0; c < nphi; ++c) { const T t = phi[size_t(c) * Vp + gi]; phi[size_t(c) * Vp + gi] = phi[size_t(c) * Vp + gj]; phi[size_t(c) * Vp + gj] = t; }
}

// Restore U, phi from the saved copies for rejected replicas.
template <typename T>
__global__ __launch_bounds__(kBlock) void restore_kernel(const RepDesc* reps, const int* blk_rep, const McState* st, unsigned Vp,
                                                        T* U, const T* Us, int nU, T* phi, const T* phis, int nphi) {
  const int r = blk_rep[blockIdx.x];
  if (st[r].accepted) return;
  const RepDesc& rd = reps[r];
  const unsigned s = blockIdx.x * kBlock + threadIdx.x - rd.offset;
  if (s >= rd.vol) return;
  const unsigned g = rd.offset + s;
  for (int c = 0; c < nU; ++c) U[size_t(c) * Vp + g] = Us[size_t(c) * Vp + g];
  for (int c = 0; c < nphi; ++c) phi[size_t(c) * Vp + g] = phis[size_t(c) * Vp + g];
}

// Momentum heatbath (fields.hpp:59-66, scalar_field.hpp:40-51): P^a from
// key(stream, s, mu, a), pi_k from key(stream', s, 2k[+1]); FP64 Gaussians.
template <typename T, int N, int D>
__global__ __launch_bounds__(kBlock) void refresh_kernel(const RepDesc* reps, const int* blk_rep, unsigned Vp, T* P, T* pi, int d,
                                                        int real_rep, const McState* st, unsigned long long tagP,
                                                        unsigned long long tagPi) {
  constexpr int NGN = N * N - 1;
  const int r = blk_rep[blockIdx.x];
  const RepDesc& rd = reps[r];
  const unsigned s = blockIdx.x * kBlock + threadIdx.x - rd.offset;
  if (s >= rd.vol) return;
  const unsigned g = rd.offset + s;
  const unsigned long long t = st[r].traj_count;
  const unsigned long long sP = stream_key(tagP, t);
  // FP32 storage takes the float draw directly (equal to T(draw_gauss), cheaper)
  auto draw = [&](unsigned long long key) -> T {
    if constexpr (sizeof(T) == 4) return draw_gauss_f32(rd.seed, key);
    else return draw_gauss(rd.seed, key);
  };
  for (int mu = 0; mu < D; ++mu)
    for (int c = 0; c < NGN; ++c)
      P[size_t(mu * NGN + c) * Vp + g] = draw(stream_key(sP, s, mu, c));
  if (d > 0) {
    const unsigned long long sPi = stream_key(tagPi, t);
    for (int k = 0; k < d; ++k) {
      pi[size_t(2 * k) * Vp + g] = draw(stream_key(sPi, s, 2 * k));
      pi[size_t(2 * k + 1) * Vp + g] = real_rep ? T(0) : draw(stream_key(sPi, s, 2 * k + 1));
    }
  }
}

// Start configurations: mode 0 cold (U = 1), 1 reference hot start
// U = exp(i sigma g_a T^a), g_a = gauss(key(1, s, mu, a)) (fields.hpp:25-34).
template <typename T, int N, int D>
__global__ __launch_bounds__(kBlock) void init_links_kernel(const RepDesc* reps, const int* blk_rep, unsigned Vp, T* U, int replica,
                                                           int mode, double sigma) {
  constexpr int NGN = N * N - 1;
  const int r = blk_rep[blockIdx.x];
  if (r != replica) return;
  const RepDesc& rd = reps[r];
  const unsigned s = blockIdx.x * kBlock + threadIdx.x - rd.offset;
  if (s >= rd.vol) return;
  const unsigned g = rd.offset + s;
  for (int mu = 0; mu < D; ++mu) {
    Mat<double, N> m;
    if (mode == 0) m = meye<double, N>();
    else {
      double v[NGN];
      for (int c = 0; c < NGN; ++c) v[c] = sigma * draw_gauss(rd.seed, stream_key(1, s, mu, c));
      m = expi<double, N>(alg_to_mat<double, N>(v, 1.0));
    }
    store_link<T, double, N>(U, Vp, mu, g, m);
  }
}
// phi: mode 0 cold (phi_0 = c), 1 Gaussian with sigma from key(stream, s, 2k[+1]) (scalar_field.hpp:33-51)
template <typename T>
__global__ __launch_bounds__(kBlock) void init_phi_kernel(const RepDesc* reps, const int* blk_rep, unsigned Vp, T* phi, int d, int real_rep,
                                                         int replica, int mode, unsigned long long stream, double sigma) {
  const int r = blk_rep[blockIdx.x];
  if (r != replica) return;
  const RepDesc& rd = reps[r];
  const unsigned s = blockIdx.x * kBlock + threadIdx.x - rd.offset;
  if (s >= rd.vol) return;
  const unsigned g = rd.offset + s;
  for (int k = 0; k < d; ++k) {
    double re, im;
    if (mode == 0) { re = (k == 0) ? sigma : 0.0; im = 0.0; }
    else {
      re = sigma * draw_gauss(rd.seed, stream_key(stream, s, 2 * k));
      im = real_rep ? 0.0 : sigma * draw_gauss(rd.seed, stream_key(stream, s, 2 * k + 1));
    }
    phi[size_t(2 * k) * Vp + g] = T(re);
    phi[size_t(2 * k + 1) * Vp + g] = T(im);
  }
}

// Reunitarize in FP64 (sun.hpp:49-72), reading U_in and writing U_out.
template <typename T, int N, int D>
__global__ __launch_bounds__(kBlock) void reunit_kernel(const RepDesc* reps, const int* blk_rep, unsigned Vp, const T* U_in, T* U_out) {
  const int r = blk_rep[blockIdx.x];
  const RepDesc& rd = reps[r];
  const unsigned s = blockIdx.x * kBlock + threadIdx.x - rd.offset;
  if (s >= rd.vol) return;
  const unsigned g = rd.offset + s;
  for (int mu = 0; mu < D; ++mu) {
    Mat<double, N> m = load_link<T, double, N>(U_in, Vp, mu, g);
    reunitarize<double, N>(m);
    store_link<T, double, N>(U_out, Vp, mu, g, m);
  }
}

// Max ||U^dag U - 1|| per reduction block (fields.hpp:38-42).
template <typename T, int N, int D>
__global__ __launch_bounds__(kRedBlock) void unitarity_kernel(ActArgs a, const T* U) {
  const int r = a.red_rep[blockIdx.x];
  const RepDesc& rd = a.reps[r];
  const unsigned s = a.red_first[blockIdx.x] - rd.offset + threadIdx.x;
  double w = 0.0;
  if (s < rd.vol) {
    const unsigned g = rd.offset + s;
    for (int mu = 0; mu < D; ++mu) w = fmax(w, unitarity_violation<double, N>(load_link<T, double, N>(U, a.Vp, mu, g)));
  }
  __shared__ double sh[kRedBlock];
  sh[threadIdx.x] = w;
  __syncthreads();
  for (int o = kRedBlock / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) a.partials[blockIdx.x] = sh[0];
}

// Polyakov loop in direction D-1 (observables.hpp:18-34): one thread per
// transverse site; per-reduction-block partial sums of Re Tr.
template <typename T, int N, int D>
__global__ __launch_bounds__(kRedBlock) void polyakov_kernel(ActArgs a, const T* U) {
  const int r = a.red_rep[blockIdx.x];
  const RepDesc& rd = a.reps[r];
  const unsigned s = a.red_first[blockIdx.x] - rd.offset + threadIdx.x;
  double q[1] = {0.0};
  const unsigned Lt = unsigned(rd.L[D - 1]);
  const unsigned nt = rd.vol / Lt;  // transverse sites are s < nt (x_{D-1} = 0)
  if (s < nt) {
    const unsigned st = rd.stride[D - 1];
    Mat<double, N> Pl = load_link<T, double, N>(U, a.Vp, D - 1, rd.offset + s);
    for (unsigned t = 1; t < Lt; ++t) Pl = mul_nn(Pl, load_link<T, double, N>(U, a.Vp, D - 1, rd.offset + s + t * st));
    q[0] = retrace(Pl);
  }
  block_reduce_store<1>(q, a.partials + blockIdx.x);
}

}  // namespace ghx
