// Context of the C ABI (shared by the API and the learned-force units).
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <string>
#include <vector>

#include "../../include/ghx.h"
#include "bigrep.hpp"
#include "kernels.hpp"
#include "measure.hpp"
#include "ops.hpp"
#include "rep_build.hpp"

using namespace ghx;

struct SurrogateState;

struct ghx_ctx {
  int D = 4, N = 2, nrep = 1, prec = GHX_FP64;
  size_t tsize = 8;
  std::vector<std::array<int, 4>> L;
  std::vector<RepDesc> hreps;
  std::vector<int> hblk_rep, hred_rep;
  std::vector<unsigned> hred_first;
  unsigned nblk = 0, nred = 0, Vp = 0;
  // representation
  RepInfo rep;
  bool ops_ok = false;
  Ops ops{};
  MeasOps mops{};  // Wilson / Polyakov loops in a probe irrep
  // integrator
  int nmd = 20, integ = GHX_OMELYAN;
  bool reunit = true;
  // device state
  RepDesc* dreps = nullptr;
  int *dblk_rep = nullptr, *dred_rep = nullptr;
  unsigned* dred_first = nullptr;
  void *U[2] = {nullptr, nullptr}, *Usave = nullptr, *P = nullptr;
  void *phi[2] = {nullptr, nullptr}, *phisave = nullptr, *pi = nullptr;
  void *F = nullptr, *Fphi = nullptr;
  double *partials = nullptr, *sums = nullptr, *Hi = nullptr, *Hf = nullptr, *Hwork = nullptr;
  McState* mc = nullptr;
  int cur = 0;
  hipStream_t stream = nullptr;
  hipGraphExec_t graph = nullptr;
  int graph_cur = -1;
  std::string err;
  const char* kernel_path = "";
  int stream_seg = 0;  // streaming MD kernel (0 = one site per thread)
  int block_order = 0;  // BlockWalk order of the MD-step kernel (single lattice)
  int link_lanes = 0;   // link-parallel MD step (one lane per link) for small lattices
  bool xlane_enabled = true;  // GHX_TUNE_X_LANES
  bool fuse_inner = true;     // GHX_TUNE_FUSE_INNER
  bool split_inner = true;    // GHX_TUNE_SPLIT_INNER
  int store_cached = 0;       // GHX_TUNE_STORE_POLICY
  bool link_split = true;     // GHX_TUNE_LINK_SPLIT
  // multi-timescale integrator (0 = single scale)
  int n_outer = 0, m_inner = 1, inner_integ = GHX_LEAPFROG;
  // surrogate link force (learned); S planes are filled by surr_fill before a kick
  void* S = nullptr;
  bool use_surr = false;
  bool frozen = false;  // |phi_x| = 1 scalar
  // multi-invariant potential (scalar_invariants.hpp): superoperator, on-site force and value buffers
  bool multi = false;
  double mu2 = 0.0;
  void* dAq = nullptr;
  void* dFon = nullptr;
  double* dVon = nullptr;
  void (*surr_fill)(ghx_ctx*, int buf) = nullptr;
  SurrogateState* surr_state = nullptr;
  // large-irrep path
  bool big = false;
  BigOps bops{};
  BigRepTables tb{};
  void* dtab[11] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  std::vector<void*> dsym;   // symmetric-power engine tables and lifted scalar
  bool sym_valid = false;    // the engine applies to this irrep (host-checked)
  bool sym_enabled = true;   // GHX_TUNE_SYM_ENGINE
  int sym_pair_mode = 2;         // GHX_TUNE_SYM_PAIRS: 0 one link per wave, 1 paired in LDS, 2 paired in registers (FP32, k2 <= 1)
  void* G = nullptr;
  double* q1 = nullptr;
  // replica exchange
  bool temper = false;
  std::vector<int> hpairs;  // (i,j) list: even pass pairs then odd pass pairs
  int neven = 0, npairs = 0;
  int *dpairs = nullptr, *dflags = nullptr, *dcfg = nullptr;
  double *dq0 = nullptr, *dq1 = nullptr;
  unsigned long long *dattempts = nullptr, *daccepts = nullptr;
  SwapState* dswap = nullptr;
  unsigned maxvol = 0;

  // link planes per site: FP32 MD stores rows 0..N-2 (the last row is rebuilt on load)
  int link_rows() const { return (tsize == 4 && GHX_COMPRESS_LINKS) ? N - 1 : N; }
  size_t nU() const { return size_t(D) * 2 * N * link_rows(); }
  size_t nP() const { return size_t(D) * (N * N - 1); }
  size_t nphi() const { return size_t(2) * rep.d; }
  // The requested block walk when it applies (one lattice whose slices and
  // (D-2)-planes are whole blocks and whose last extent splits over the 8 XCDs),
  // else storage order.
  BlockWalk block_walk() const {
    BlockWalk w;
    if (block_order == 0 || nrep != 1) return w;
    const RepDesc& r = hreps[0];
    if (block_order >= 3) {  // 4D tiles of kBlock sites
      static const unsigned shape[4][3] = {{16, 4, 2}, {8, 4, 2}, {16, 2, 2}, {32, 2, 2}};
      const unsigned* sh = shape[block_order - 3];
      const unsigned tt = unsigned(kBlock) / (sh[0] * sh[1] * sh[2]);
      if (D != 4 || nblk * unsigned(kBlock) != unsigned(r.vol)) return w;
      if (unsigned(r.L[0]) % sh[0] || unsigned(r.L[1]) % sh[1] || unsigned(r.L[2]) % sh[2] || unsigned(r.L[3]) % tt) return w;
      w.order = block_order;
      w.tx = sh[0]; w.ty = sh[1]; w.tz = sh[2];
      w.nbx = unsigned(r.L[0]) / sh[0]; w.nby = unsigned(r.L[1]) / sh[1]; w.nbz = unsigned(r.L[2]) / sh[2];
      w.L0 = unsigned(r.L[0]); w.L1 = unsigned(r.L[1]); w.L2 = unsigned(r.L[2]);
      return w;
    }
    const unsigned Lt = unsigned(r.L[D - 1]), V3 = r.stride[D - 1], plane = r.stride[D - 2];
    if (V3 % kBlock || plane % kBlock || Lt % 8 || nblk != V3 / kBlock * Lt) return w;
    w.order = block_order;
    w.nbt = V3 / kBlock;
    w.tt = Lt / 8;
    w.ngy = plane / kBlock;
    w.lz = unsigned(r.L[D - 2]);
    return w;
  }
  // Every wave of the one-site-per-thread MD step holds whole x-lines and is
  // either wholly live or wholly past the end of its lattice: SU(3) with two
  // stored rows, L0 a power of two dividing 64, replica offsets and volumes (and
  // the cross-section of a single lattice, for the streaming kernel) multiples
  // of 64 sites.
  bool xlane_ok() const {
    if (tsize != 4 || N != 3 || link_rows() != 2) return false;
    for (int r = 0; r < nrep; ++r) {
      const RepDesc& h = hreps[r];
      const unsigned l0 = unsigned(h.L[0]);
      if (l0 < 2 || l0 > 64 || (l0 & (l0 - 1)) || h.offset % 64 || h.vol % 64) return false;
    }
    return nrep != 1 || hreps[0].stride[D - 1] % 64 == 0;
  }
  LaunchCtx lc() const {
    LaunchCtx c;
    c.stream = stream; c.nblk = nblk; c.nred = nred; c.Vp = Vp;
    c.reps = dreps; c.blk_rep = dblk_rep; c.red_rep = dred_rep; c.red_first = dred_first;
    c.partials = partials; c.W = rep.W.data(); c.d = rep.d; c.real_rep = rep.real ? 1 : 0;
    c.G = G;
    c.S = S;
    c.surr_sign = 0.0;
    c.frozen = frozen ? 1 : 0;
    c.multi = multi ? 1 : 0;
    c.Aq = dAq;
    c.Fon = dFon;
    c.Von = dVon;
    c.mu2 = mu2;
    c.stream_seg = (nrep == 1) ? stream_seg : 0;
    c.V3 = hreps[0].stride[D - 1];
    c.Lt = unsigned(hreps[0].L[D - 1]);
    c.walk = block_walk();
    c.link_lanes = link_lanes;
    c.xlane = (xlane_enabled && xlane_ok()) ? 1 : 0;
    c.fuse_inner = fuse_inner ? 1 : 0;
    c.store_cached = store_cached;
    c.link_split = link_split ? 1 : 0;
    return c;
  }
};


// Helpers defined in ghx_api.hip, used by the learned-force unit.
namespace ghx {
int ctx_fail(ghx_ctx* c, int code, const std::string& msg);
int ctx_check_ready(ghx_ctx* ctx);
// Link force restricted to `parts` (F_NO_* flags) of buffer b into ctx->F (and
// ctx->Fphi); big reps run their matter kernel first.
void ctx_launch_forces(ghx_ctx* c, int b, int parts);
void surrogate_free(ghx_ctx* c);  // learned.hip
}  // namespace ghx
