// Launch table: one set of launchers per compiled (precision, N, D, rep)
// combination. Instantiations are spread over the inst_*.hip translation units.
#pragma once
#include <hip/hip_runtime.h>

namespace ghx {

struct RepDesc;
struct McState;

// Order in which the MD-step kernel walks the blocks of a single lattice
// (D >= 3, extents L_0..L_{D-1}, blocks of 128 sites in storage order).
//  0: storage order (the last direction slowest);
//  1: each XCD's contiguous share of Lt/8 slices of the last direction is
//     walked with that direction fastest, so a block's +-t neighbour blocks run
//     1-3 blocks earlier or later on the same XCD (still in its L2) instead of
//     a whole slice of blocks apart;
//  2: as 1, and inside a slice direction D-2 runs before the lower directions.
// nbt = blocks per slice, tt = Lt/8, ngy = blocks per (D-2)-plane, lz = L_{D-2}.
//  3-6: a workgroup takes a 4D tile of kBlock sites (tx x ty x tz x tt: 3 16x4x2x1,
//  4 8x4x2x2, 5 16x2x2x2, 6 32x2x2x1), tiles numbered x fastest and dealt to
//  the XCDs in contiguous chunks; nb* = tiles per direction, L* = extents.
struct BlockWalk {
  int order = 0;
  unsigned nbt = 1, tt = 1, ngy = 1, lz = 1;
  unsigned tx = 1, ty = 1, tz = 1, nbx = 1, nby = 1, nbz = 1, L0 = 1, L1 = 1, L2 = 1;
};

struct LaunchCtx {
  hipStream_t stream;
  unsigned nblk, nred, Vp;
  const RepDesc* reps;
  const int* blk_rep;
  const int* red_rep;
  const unsigned* red_first;
  double* partials;
  const double* W;  // tensor basis on the host, d x dimT (tensor reps)
  int d, real_rep;
  const void* G;    // external matter link force planes (large reps)
  int stream_seg;   // > 0: single-lattice streaming kernel with this many t-slices per block
  unsigned V3, Lt;  // cross-section volume and last extent of replica 0 (streaming kernel)
  BlockWalk walk;   // block order of the MD-step kernel
  int link_lanes;   // 1: link-parallel MD step (one lane per link)
  int xlane;        // 1: every wave of the MD step holds whole x-lines (x-neighbour lanes share products)
  int fuse_inner;   // 1: a large irrep's inner-scale step runs its on-site and gauge updates in one launch
  int store_cached; // 1: the MD step's stores take the default cache policy (F_STORE_CACHED) instead of streaming
  int link_split;   // 1: link lanes with no scalar in the kernel split each link's staples over its planes (md_link3_kernel)
  const void* S;    // surrogate link force planes
  double surr_sign;
  int frozen;       // |phi_x| = 1 scalar: adds F_FROZEN to every MD launch
  int multi;        // multi-invariant potential in use (on-site force from Fon)
  const void* Aq;   // its superoperator (complex, storage type)
  void* Fon;        // on-site force planes
  double* Von;      // per-site potential (FP64)
  double mu2;
};

struct BigRepTables;
// Extra launchers of the large-irrep path.
struct BigOps {
  void (*matter)(const LaunchCtx&, const BigRepTables&, const void* U, const void* phi_in, void* phi_out, void* pi, void* G,
                 void* Fphi, double* q1, double ck, double cd, int flags, int mode);
  void (*action)(const LaunchCtx&, const void* U, const void* P, const void* phi, const void* pi, const double* q1, int d);
  void (*phi_drift)(const LaunchCtx&, const void* phi_in, void* phi_out, const void* pi, int nplanes, double cd, int abs_eps);
  void (*onsite)(const LaunchCtx&, const void* phi_in, void* phi_out, void* pi, int d, double ck, double cd, int flags);
  // the inner-scale step (F_NO_HOP): on-site update and gauge update, fused where the links run link-parallel
  void (*onsite_gauge)(const LaunchCtx&, const void* U_in, void* U_out, void* P, const void* phi_in, void* phi_out, void* pi, int d,
                       double ck, double cd, int flags);
  // static LDS bytes of the tensor-apply matter kernel (its dynamic LDS holds the dimT-entry tensor)
  size_t (*matter_static_lds)();
  // the inner scale split: every on-site sub-step of an inner block in one launch
  // (seq: a const OnsiteSeq*), and the gauge update of one sub-step alone
  void (*onsite_seq)(const LaunchCtx&, const void* phi_in, void* phi_out, void* pi, int d, const void* seq);
  void (*gauge_step)(const LaunchCtx&, const void* U_in, void* U_out, void* P, const void* phi_in, void* phi_out, void* pi, double ck,
                     double cd, int flags);
};

struct Ops {
  void (*md_step)(const LaunchCtx&, const void* U_in, void* U_out, void* P, const void* phi_in, void* phi_out, void* pi,
                  void* F, void* Fphi, double ck, double cd, int flags);
  void (*drift)(const LaunchCtx&, const void* U_in, void* U_out, const void* P, const void* phi_in, void* phi_out,
                const void* pi, double cd, int flags);
  void (*action)(const LaunchCtx&, const void* U, const void* P, const void* phi, const void* pi);
  void (*refresh)(const LaunchCtx&, void* P, void* pi, const McState* st, unsigned long long tagP, unsigned long long tagPi);
  void (*init_links)(const LaunchCtx&, void* U, int replica, int mode, double sigma);
  void (*reunit)(const LaunchCtx&, const void* U_in, void* U_out);
  void (*unitarity)(const LaunchCtx&, const void* U);
  void (*polyakov)(const LaunchCtx&, const void* U);
  void (*frozen_site)(const LaunchCtx&, void* phi, void* pi, int mode);
};

// Returns false when the combination is not compiled. kind: RepKind; for
// tensor reps (nbox, d) select the instantiation.
bool ops_none(int prec, int N, int D, Ops* o);
bool ops_fund(int prec, int N, int D, Ops* o);
bool ops_adj(int prec, int N, int D, Ops* o);
bool ops_tensor(int prec, int N, int D, int nbox, int d, Ops* o);
bool ops_sym2(int prec, int N, int D, Ops* o);
bool ops_big(int prec, int N, int D, Ops* o, BigOps* b);

}  // namespace ghx
