// C-ABI implementation: context, device buffers, trajectory driver.
// The trajectory (refresh, save, H_i, MD, reunitarize, H_f, Metropolis,
// restore) runs entirely on the context's stream with no host round trip, so
// it can be captured once into a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "ctx.hpp"
#include "rng_kernels.hpp"
#include "potential.hpp"
#include "symrep.hpp"

namespace {
constexpr double kOmelyanLambda = 0.1931833275037836;  // config.hpp:25
constexpr unsigned long long kTagP_pure = 0xA1, kTagMC_pure = 0xB2;        // hmc.hpp:74,86
constexpr unsigned long long kTagP = 0xC3, kTagPi = 0xD4, kTagMC = 0xE5;   // gauge_higgs_hmc.hpp:167-188
}  // namespace

namespace {

int fail(ghx_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}
#define HIPCHK(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) return fail(ctx, GHX_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

void free_fields(ghx_ctx* c) {
  for (void** p : {&c->U[0], &c->U[1], &c->Usave, &c->P, &c->phi[0], &c->phi[1], &c->phisave, &c->pi, &c->F, &c->Fphi, &c->G,
                   (void**)&c->q1})
    if (*p) { (void)hipFree(*p); *p = nullptr; }
}
void free_multi(ghx_ctx* c) {
  for (void** p : {&c->dAq, &c->dFon, (void**)&c->dVon})
    if (*p) { (void)hipFree(*p); *p = nullptr; }
  c->multi = false;
}
void free_tables(ghx_ctx* c) {
  for (auto& p : c->dtab)
    if (p) { (void)hipFree(p); p = nullptr; }
  for (void* p : c->dsym) (void)hipFree(p);
  c->dsym.clear();
  c->tb.sym = BigRepTables::Sym{};
  c->sym_valid = false;
}

int alloc_zero(ghx_ctx* ctx, void** p, size_t bytes) {
  if (bytes == 0) { *p = nullptr; return GHX_OK; }
  HIPCHK(hipMalloc(p, bytes));
  HIPCHK(hipMemsetAsync(*p, 0, bytes, ctx->stream));
  return GHX_OK;
}

int alloc_fields(ghx_ctx* ctx) {
  free_fields(ctx);
  const size_t V = ctx->Vp, ts = ctx->tsize;
  int rc;
  for (int b = 0; b < 2; ++b) if ((rc = alloc_zero(ctx, &ctx->U[b], ctx->nU() * V * ts))) return rc;
  if ((rc = alloc_zero(ctx, &ctx->Usave, ctx->nU() * V * ts))) return rc;
  if ((rc = alloc_zero(ctx, &ctx->P, ctx->nP() * V * ts))) return rc;
  if ((rc = alloc_zero(ctx, &ctx->F, ctx->nP() * V * ts))) return rc;
  if (ctx->rep.d > 0) {
    for (int b = 0; b < 2; ++b) if ((rc = alloc_zero(ctx, &ctx->phi[b], ctx->nphi() * V * ts))) return rc;
    if ((rc = alloc_zero(ctx, &ctx->phisave, ctx->nphi() * V * ts))) return rc;
    if ((rc = alloc_zero(ctx, &ctx->pi, ctx->nphi() * V * ts))) return rc;
    if ((rc = alloc_zero(ctx, &ctx->Fphi, ctx->nphi() * V * ts))) return rc;
  }
  if (ctx->big) {
    if ((rc = alloc_zero(ctx, &ctx->G, ctx->nP() * V * ts))) return rc;
    if ((rc = alloc_zero(ctx, (void**)&ctx->q1, V * sizeof(double)))) return rc;
  }
  ctx->cur = 0;
  return GHX_OK;
}

int upload_reps(ghx_ctx* ctx) {
  HIPCHK(hipMemcpyAsync(ctx->dreps, ctx->hreps.data(), sizeof(RepDesc) * ctx->nrep, hipMemcpyHostToDevice, ctx->stream));
  return GHX_OK;
}

// True when W is the single-row n=2 basis in odometer order (RepSym2's layout).
bool sym2_canonical(const RepInfo& r) {
  if (r.kind != REP_TENSOR || r.rows.size() != 1 || r.rows[0] != 2) return false;
  const int N = r.N;
  if (r.d != N * (N + 1) / 2 || r.dimT != long(N) * N) return false;
  int k = 0;
  for (int i = 0; i < N; ++i)
    for (int j = i; j < N; ++j, ++k)
      for (long t = 0; t < r.dimT; ++t) {
        const int a = int(t / N), b = int(t % N);
        double want = 0.0;
        if (i == j && a == i && b == i) want = 1.0;
        else if (i != j && ((a == i && b == j) || (a == j && b == i))) want = std::sqrt(0.5);
        if (std::fabs(r.W[size_t(k) * r.dimT + t] - want) > 1e-15) return false;
      }
  return true;
}

// Sparse W (CSC by tensor index, CSR by basis index) and sparse T^a_R (CSR by row) on the device.
int upload_tables(ghx_ctx* ctx) {
  free_tables(ctx);
  const RepInfo& r = ctx->rep;
  const long dimT = r.dimT;
  std::vector<int> cptr(dimT + 1, 0), crow, rptr(r.d + 1, 0), rcol;
  std::vector<double> cval, rval;
  for (long t = 0; t < dimT; ++t) {
    for (int l = 0; l < r.d; ++l) {
      const double w = r.W[size_t(l) * dimT + t];
      if (w != 0.0) { crow.push_back(l); cval.push_back(w); }
    }
    cptr[t + 1] = int(crow.size());
  }
  for (int l = 0; l < r.d; ++l) {
    for (long t = 0; t < dimT; ++t) {
      const double w = r.W[size_t(l) * dimT + t];
      if (w != 0.0) { rcol.push_back(int(t)); rval.push_back(w); }
    }
    rptr[l + 1] = int(rcol.size());
  }
  // T^a_R by rows (a, l), nonzeros only (a zero entry adds exactly nothing)
  const int ngn = ctx->N * ctx->N - 1;
  std::vector<int> tptr(size_t(ngn) * r.d + 1, 0), tcol;
  std::vector<double> tval;
  for (int a = 0; a < ngn; ++a)
    for (int l = 0; l < r.d; ++l) {
      for (int m = 0; m < r.d; ++m) {
        const std::complex<double> z = r.T[(size_t(a) * r.d + l) * r.d + m];
        if (z != 0.0) { tcol.push_back(m); tval.push_back(z.real()); tval.push_back(z.imag()); }
      }
      tptr[size_t(a) * r.d + l + 1] = int(tcol.size());
    }
  const std::vector<float> cvalf(cval.begin(), cval.end()), rvalf(rval.begin(), rval.end());
  const void* src[11] = {cptr.data(), crow.data(), cval.data(), rptr.data(), rcol.data(), rval.data(), cvalf.data(), rvalf.data(),
                         tptr.data(), tcol.data(), tval.data()};
  const size_t bytes[11] = {cptr.size() * 4, crow.size() * 4, cval.size() * 8, rptr.size() * 4, rcol.size() * 4, rval.size() * 8,
                            cvalf.size() * 4, rvalf.size() * 4, tptr.size() * 4, tcol.size() * 4, tval.size() * 8};
  for (int k = 0; k < 11; ++k) {
    HIPCHK(hipMalloc(&ctx->dtab[k], std::max<size_t>(bytes[k], 8)));
    HIPCHK(hipMemcpy(ctx->dtab[k], src[k], bytes[k], hipMemcpyHostToDevice));
  }
  ctx->tb = BigRepTables{r.d, r.nbox, int(dimT), (const int*)ctx->dtab[0], (const int*)ctx->dtab[1], (const double*)ctx->dtab[2],
                         (const int*)ctx->dtab[3], (const int*)ctx->dtab[4], (const double*)ctx->dtab[5], (const float*)ctx->dtab[6],
                         (const float*)ctx->dtab[7], (const int*)ctx->dtab[8], (const int*)ctx->dtab[9], (const double*)ctx->dtab[10]};
  // symmetric-power engine, when the irrep admits it and passes the host checks
  const SymBuild sb = build_sym(r);
  const SymProgram sp = sb.ok ? sym_program(sb) : SymProgram{};
  ctx->sym_valid = sb.ok && sp.ok;
  if (ctx->sym_valid) {
    const int dimSp = (sb.dimS + 15) / 16 * 16;
    std::vector<double> Yl(size_t(r.d) * dimSp, 0.0);
    for (int q = 0; q < sb.dimS; ++q)
      for (int l = 0; l < r.d; ++l) Yl[size_t(l) * dimSp + q] = sb.Yc[size_t(q) * r.d + l];
    const std::vector<float> Ylf(Yl.begin(), Yl.end()), pref(sp.pre.begin(), sp.pre.end()), ysvf(sp.ysv.begin(), sp.ysv.end());
    // the program in LDS byte offsets of this precision's complex slots: sources
    // as offsets, the destination's offset with the column j in its low bits
    const int csz = 2 * int(ctx->tsize);
    std::vector<int> progb(sp.prog.size());
    for (size_t e = 0; e < sp.prog.size(); e += 4) {
      for (int i = 0; i < 3; ++i) progb[e + i] = sp.prog[e + i] * csz;
      progb[e + 3] = (sp.prog[e + 3] >> 2) * csz | (sp.prog[e + 3] & 3);
    }
    const void* ssrc[8] = {progb.data(), sp.pre.data(), pref.data(), sp.ysv.data(), ysvf.data(), Yl.data(), Ylf.data(), sp.ysq.data()};
    const size_t sbytes[8] = {sp.prog.size() * 4, sp.pre.size() * 8, pref.size() * 4, sp.ysv.size() * 8, ysvf.size() * 4,
                              Yl.size() * 8, Ylf.size() * 4, sp.ysq.size() * 4};
    void* dp[8] = {};
    for (int k = 0; k < 8; ++k) {
      HIPCHK(hipMalloc(&dp[k], std::max<size_t>(sbytes[k], 16)));
      ctx->dsym.push_back(dp[k]);
      HIPCHK(hipMemcpy(dp[k], ssrc[k], sbytes[k], hipMemcpyHostToDevice));
    }
    void* lift = nullptr;
    HIPCHK(hipMalloc(&lift, size_t(2) * sb.dimS * ctx->Vp * ctx->tsize));
    ctx->dsym.push_back(lift);
    void* hops = nullptr;
    HIPCHK(hipMalloc(&hops, size_t(2) * ctx->D * r.d * ctx->Vp * ctx->tsize));
    ctx->dsym.push_back(hops);
    void* hopsL = nullptr;
    HIPCHK(hipMalloc(&hopsL, size_t(2) * ctx->D * r.d * ctx->Vp * ctx->tsize));
    ctx->dsym.push_back(hopsL);
    BigRepTables::Sym& st = ctx->tb.sym;
    st.s1 = sb.s1; st.s2 = sb.s2; st.dimS = sb.dimS; st.npass = sp.npass; st.ng = sp.ng;
    st.s1p = sp.s1p; st.s2p = sp.s2p;
    st.prog = (const int4*)dp[0];
    st.pre = (const double*)dp[1]; st.pref = (const float*)dp[2];
    st.ysv = (const double*)dp[3]; st.ysvf = (const float*)dp[4]; st.ysq = (const int*)dp[7]; st.ynz = sp.ynz;
    st.Yl = (const double*)dp[5]; st.Ylf = (const float*)dp[6]; st.dimSp = dimSp;
    st.lift = lift;
    st.hops = hops;
    st.hopsL = hopsL;
    st.lds_one = sp.lds_one; st.lds_S1 = sp.lds_S1; st.lds_S2 = sp.lds_S2;
    st.lds_V = sp.lds_V; st.lds_T1 = sp.lds_T1; st.lds_W = sp.lds_W; st.lds_total = sp.lds_total;
    st.k1 = sb.k1; st.k2 = sb.k2; st.lds_Cb = sp.lds_Cb;
    st.pair_S2 = sp.pair_S2; st.pair_V = sp.pair_V; st.pair_T1 = sp.pair_T1; st.pair_total = sp.pair_total;
    // rows of T^a_R for g^a: lane j of generator a (lanes a lpg .. a lpg + lpg - 1)
    // takes the rows m = j, j + lpg, ... of T^a; entry k of its row slot rr is
    // (l, T^a_ml) at [(rr nz + k) 64 + lane], rows padded with zeros to nz entries
    const int lpg = ngn <= 4 ? 16 : 8;
    const int nrow = (r.d + lpg - 1) / lpg;
    size_t K = 1;
    for (int a = 0; a < ngn; ++a)
      for (int m = 0; m < r.d; ++m) {
        size_t c = 0;
        for (int l = 0; l < r.d; ++l) c += r.T[(size_t(a) * r.d + m) * r.d + l] != 0.0;
        K = std::max(K, c);
      }
    std::vector<int> glm(size_t(nrow) * K * 64, 0);
    std::vector<double> gval(glm.size() * 2, 0.0);
    for (int a = 0; a < ngn; ++a)
      for (int j = 0; j < lpg; ++j)
        for (int rr = 0; rr < nrow; ++rr) {
          const int m = j + rr * lpg, lane = a * lpg + j;
          if (m >= r.d) continue;
          size_t k = 0;
          for (int l = 0; l < r.d; ++l) {
            const std::complex<double> z = r.T[(size_t(a) * r.d + m) * r.d + l];
            if (z == 0.0) continue;
            const size_t e = (size_t(rr) * K + k++) * 64 + size_t(lane);
            glm[e] = l;
            gval[2 * e] = z.real();
            gval[2 * e + 1] = z.imag();
          }
        }
    const std::vector<float> gvalf(gval.begin(), gval.end());
    const void* gsrc[3] = {glm.data(), gval.data(), gvalf.data()};
    const size_t gbytes[3] = {glm.size() * 4, gval.size() * 8, gvalf.size() * 4};
    void* gp[3] = {};
    for (int k = 0; k < 3; ++k) {
      HIPCHK(hipMalloc(&gp[k], gbytes[k]));
      ctx->dsym.push_back(gp[k]);
      HIPCHK(hipMemcpy(gp[k], gsrc[k], gbytes[k], hipMemcpyHostToDevice));
    }
    st.glm = (const int*)gp[0]; st.gval = (const double*)gp[1]; st.gvalf = (const float*)gp[2];
    st.grow = nrow;
    st.gK = int(K); st.glpg = lpg;
  }
  ctx->tb.sym.on = (ctx->sym_valid && ctx->sym_enabled) ? 1 : 0;
  return GHX_OK;
}

// LDS that the large-irrep path asks for at launch, against the device's
// per-workgroup limit: the engine's LDS program, or the tensor-apply kernel's
// dimT-entry tensor (dynamic) plus its static arrays. Irreps that fit neither
// are refused here instead of failing at their first launch.
// The engine's link kernel for the chosen mode: 0 one link per wave (LDS
// program), 1 two links per wave with Sym^k(U) in LDS, 2 two links per wave with
// the columns of Sym^k(U) in registers (FP32 with a second Young row <= 1; else 1).
int sym_pair_applies(const ghx_ctx* ctx) {
  if (!ctx->sym_valid || ctx->sym_pair_mode == 0 || ctx->N != 3 || ctx->tb.sym.k1 < 2 || ctx->tb.sym.k1 > kSymPairMaxK) return 0;
  return ctx->sym_pair_mode == 2 && ctx->tsize == 4 && ctx->tb.sym.k2 <= 1 ? 2 : 1;
}
size_t big_lds_need(const ghx_ctx* ctx, bool engine) {
  const size_t cs = 2 * size_t(ctx->tsize);
  if (engine && ctx->tb.sym.pair == 2) return 2 * 2 * size_t(ctx->tb.sym.dimS) * cs;  // static: the two lifted vectors per link
  if (engine) return ctx->tb.sym.pair ? 2 * size_t(ctx->tb.sym.pair_total) * cs : size_t(ctx->tb.sym.lds_total) * cs;
  const size_t st = ctx->bops.matter_static_lds ? ctx->bops.matter_static_lds() : size_t(-1);
  return st == size_t(-1) ? st : size_t(ctx->rep.dimT) * cs + st;
}
int lds_limit(ghx_ctx* ctx, size_t* lim) {
  int dev = 0, v = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  *lim = size_t(v);
  return GHX_OK;
}
int check_big_lds(ghx_ctx* ctx) {
  if (!ctx->big) return GHX_OK;
  ctx->tb.sym.pair = sym_pair_applies(ctx);
  size_t lim = 0;
  int rc = lds_limit(ctx, &lim);
  if (rc) return rc;
  if (ctx->tb.sym.on && ctx->tb.sym.pair && big_lds_need(ctx, true) > lim) ctx->tb.sym.pair = 0;  // one link per workgroup
  if (ctx->tb.sym.on && big_lds_need(ctx, true) > lim) {  // the tensor path may still fit
    ctx->sym_valid = false;
    ctx->tb.sym.on = 0;
  }
  if (!ctx->tb.sym.on) {
    const size_t need = big_lds_need(ctx, false);
    if (need > lim)
      return fail(ctx, GHX_E_UNSUPPORTED, "irrep " + ctx->rep.name + ": the tensor apply needs " + std::to_string(need) +
                                               " bytes of LDS per workgroup, the device allows " + std::to_string(lim));
  }
  return GHX_OK;
}

int select_ops(ghx_ctx* ctx) {
  const int p = ctx->prec == GHX_FP32MD ? 1 : 0;
  const RepInfo& r = ctx->rep;
  bool ok = false;
  ctx->big = false;
  switch (r.kind) {
    case REP_NONE: ok = ops_none(p, ctx->N, ctx->D, &ctx->ops); break;
    case REP_FUND: ok = ops_fund(p, ctx->N, ctx->D, &ctx->ops); break;
    case REP_ADJ: ok = ops_adj(p, ctx->N, ctx->D, &ctx->ops); break;
    case REP_TENSOR:
      if (sym2_canonical(r) && !std::getenv("GHX_GENERIC_TENSOR")) { ok = ops_sym2(p, ctx->N, ctx->D, &ctx->ops); ctx->kernel_path = "sym2"; }
      else if (ops_tensor(p, ctx->N, ctx->D, r.nbox, r.d, &ctx->ops)) { ok = true; ctx->kernel_path = "tensor"; }
      else if (r.d <= kBigMaxD && ops_big(p, ctx->N, ctx->D, &ctx->ops, &ctx->bops)) {
        ok = true; ctx->big = true; ctx->kernel_path = "big";
        const int rc2 = upload_tables(ctx);
        if (rc2) return rc2;
      }
      break;
  }
  if (ok) ok = meas_ops(p, ctx->N, ctx->D, &ctx->mops);
  ctx->ops_ok = ok;
  if (!ok)
    return fail(ctx, GHX_E_UNSUPPORTED, "no compiled kernels for N=" + std::to_string(ctx->N) + " D=" + std::to_string(ctx->D) +
                                            " rep=" + r.name);
  if (int rc = check_big_lds(ctx)) {
    ctx->ops_ok = false;
    return rc;
  }
  return GHX_OK;
}

int check_ready(ghx_ctx* ctx) {
  if (!ctx) return GHX_E_INVALID;
  if (!ctx->ops_ok) return fail(ctx, GHX_E_INVALID, "representation not set (call ghx_set_rep)");
  return GHX_OK;
}

// The swap rule leaves the on-site potential out of Delta, which holds only
// while the two replicas of a pair share lambda. ghx_set_couplings may change
// lambda after the ladders were set, so every entry point that swaps checks it.
int check_ladders(ghx_ctx* ctx) {
  if (!ctx->temper || ctx->rep.d == 0) return GHX_OK;
  for (int p = 0; p < ctx->npairs; ++p) {
    const int a = ctx->hpairs[2 * p], b = ctx->hpairs[2 * p + 1];
    if (ctx->hreps[a].lambda != ctx->hreps[b].lambda)
      return fail(ctx, GHX_E_INVALID, "replicas " + std::to_string(a) + " and " + std::to_string(b) +
                                          " of one ladder have different lambda (swaps need equal lambda)");
  }
  return GHX_OK;
}

bool has_phi(const ghx_ctx* c) { return c->rep.d > 0; }

// Sum per-site quantities into ctx->sums [nrep][kNQ] for buffer `b`.
void launch_action(ghx_ctx* c, int b) {
  const LaunchCtx l = c->lc();
  if (c->big) {
    c->bops.matter(l, c->tb, c->U[b], c->phi[b], nullptr, c->pi, c->G, nullptr, c->q1, 0.0, 0.0, 0, BIG_ACTION);
    c->bops.action(l, c->U[b], c->P, c->phi[b], c->pi, c->q1, c->rep.d);
  } else {
    c->ops.action(l, c->U[b], c->P, c->phi[b], c->pi);
  }
  hipLaunchKernelGGL(finalize_kernel, dim3(c->nrep), dim3(kRedBlock), 0, c->stream, c->dreps, c->partials, kNQ, c->sums);
}
void launch_hamiltonian(ghx_ctx* c, int b, double* H) {
  launch_action(c, b);
  hipLaunchKernelGGL(hamiltonian_kernel, dim3((c->nrep + 63) / 64), dim3(64), 0, c->stream, c->dreps, c->nrep, c->N, c->D,
                     has_phi(c) ? 1 : 0, c->sums, H);
}

// A kick (+ optional drift) with part flags; surr_sign multiplies the
// surrogate force S (filled here from the pre-kick fields) when non-zero.
void launch_part_step(ghx_ctx* c, double ck, double cd, bool drift, int parts, double surr_sign) {
  LaunchCtx l = c->lc();
  const int in = c->cur, out = c->cur ^ 1;
  int flags = F_KICK | (drift ? F_DRIFT : 0) | parts;
  if (surr_sign != 0.0 && c->use_surr && c->surr_fill) {
    c->surr_fill(c, in);
    flags |= F_SURR;
    l.surr_sign = surr_sign;
  }
  if (c->big && (flags & F_NO_HOP)) {  // the inner scale of a large irrep: on-site and gauge updates
    c->bops.onsite_gauge(l, c->U[in], c->U[out], c->P, c->phi[in], c->phi[out], c->pi, c->rep.d, ck, cd, flags);
    if (drift) c->cur = out;
    return;
  }
  if (c->big) c->bops.matter(l, c->tb, c->U[in], c->phi[in], c->phi[out], c->pi, c->G, nullptr, nullptr, ck, cd, flags, BIG_STEP);
  c->ops.md_step(l, c->U[in], c->U[out], c->P, c->phi[in], c->phi[out], c->pi, c->F, c->Fphi, ck, cd, flags);
  if (drift) c->cur = out;
}

// One fused kick+drift (or kick-only when cd == 0 and !drift) on buffer cur.
void launch_step(ghx_ctx* c, double ck, double cd, bool drift) {
  const LaunchCtx l = c->lc();
  const int in = c->cur, out = c->cur ^ 1;
  int flags = F_KICK | (drift ? F_DRIFT : 0);
  if (c->big) c->bops.matter(l, c->tb, c->U[in], c->phi[in], c->phi[out], c->pi, c->G, nullptr, nullptr, ck, cd, flags, BIG_STEP);
  c->ops.md_step(l, c->U[in], c->U[out], c->P, c->phi[in], c->phi[out], c->pi, c->F, c->Fphi, ck, cd, flags);
  if (drift) c->cur = out;
}

// md_evolve as a sequence of fused launches (gauge_higgs_hmc.hpp:149-165):
// leapfrog  K(e/2)D(e) [K(e)D(e)]^{nmd-1} K(e/2)
// Omelyan   K(le)D(e/2)K((1-2l)e)D(e/2)K(2le)D(e/2)...K((1-2l)e)D(e/2)K(le)
// Multi-timescale leapfrog (Sexton-Weingarten, theory_notes.md:129-134), eps = tau/n_outer:
//   [K_out(eps/2) [K_in(d/2) D(d) K_in(d/2)]^m K_out(eps/2)]^n_outer,  d = eps/m,
// K_in: gauge force + on-site scalar force (+ surrogate S), K_out: the hopping
// forces (matter link force and the D(U) phi hops of the scalar force, - S).
// Adjacent inner half kicks are merged and fused with the drifts. With the
// Omelyan integrator selected the outer scale is the minimum-norm scheme
//   [K_out(l eps) I(eps/2) K_out((1-2l) eps) I(eps/2) K_out(l eps)]^n_outer
// (adjacent K_out(l eps) merged), I(t) = m inner steps over time t: leapfrog,
// or with GHX_OMELYAN chosen for the inner scale the minimum-norm step
//   K_in(l h) D(h/2) K_in((1-2l) h) D(h/2) K_in(l h),  h = t/m,
// adjacent K_in(l h) merged (the inner_fast scheme of u1.hpp:300-317): 2m+1
// fused launches against m+1 for leapfrog, with a far smaller error constant.
constexpr int kInParts = F_NO_HOP, kOutParts = F_NO_GAUGE | F_NO_ONSITE;
// The inner sub-steps of a large irrep, (kick, drift, drifts) each, as
// launch_inner issues them.
struct InnerStep {
  double ck, cd;
  bool drift;
};
std::vector<InnerStep> inner_steps(const ghx_ctx* c, double t) {
  const int m = c->m_inner;
  const double h = t / m;
  std::vector<InnerStep> v;
  if (c->inner_integ == GHX_OMELYAN) {
    const double lam = kOmelyanLambda;
    v.push_back({lam * h, 0.5 * h, true});
    for (int k = 0; k < m; ++k) {
      v.push_back({(1.0 - 2.0 * lam) * h, 0.5 * h, true});
      if (k != m - 1) v.push_back({2.0 * lam * h, 0.5 * h, true});
    }
    v.push_back({lam * h, 0.0, false});
  } else {
    v.push_back({0.5 * h, h, true});
    for (int k = 1; k < m; ++k) v.push_back({h, h, true});
    v.push_back({0.5 * h, 0.0, false});
  }
  return v;
}
// Without a surrogate force or a plugged-in potential, the inner scale of a
// large irrep moves (U, P) by the gauge force alone and (phi, pi) by the
// on-site force alone: the two updates do not depend on each other. Then all
// of a block's on-site sub-steps run in one launch (onsite_seq_kernel, the
// site's components in registers throughout) and the links take one
// gauge-only launch per sub-step; the fields are those of the interleaved
// launches bit for bit.
bool inner_split_applies(const ghx_ctx* c) {
  return c->big && c->split_inner && !(c->use_surr && c->surr_fill) && !c->multi && c->rep.d <= 64 &&
         2 * c->m_inner + 1 <= kOnsiteMaxSteps;
}
void launch_inner_split(ghx_ctx* c, double t) {
  const std::vector<InnerStep> v = inner_steps(c, t);
  LaunchCtx l = c->lc();
  OnsiteSeq q{};
  q.n = int(v.size());
  int ndrift = 0;
  for (int s = 0; s < q.n; ++s) {
    q.ck[s] = v[s].ck;
    q.cd[s] = v[s].cd;
    q.flags[s] = F_KICK | (v[s].drift ? F_DRIFT : 0) | kInParts;
    ndrift += v[s].drift ? 1 : 0;
  }
  const int in = c->cur, fin = c->cur ^ (ndrift & 1);
  c->bops.onsite_seq(l, c->phi[in], c->phi[fin], c->pi, c->rep.d, &q);
  for (const InnerStep& s : v) {
    const int a = c->cur, b = c->cur ^ 1;
    c->bops.gauge_step(l, c->U[a], c->U[b], c->P, c->phi[fin], c->phi[fin], c->pi, s.ck, s.cd,
                       F_KICK | (s.drift ? F_DRIFT : 0) | kInParts);
    if (s.drift) c->cur = b;
  }
}
void launch_inner(ghx_ctx* c, double t) {  // t in units of eps
  if (inner_split_applies(c)) { launch_inner_split(c, t); return; }
  const int m = c->m_inner;
  const double h = t / m;
  if (c->inner_integ == GHX_OMELYAN) {
    const double lam = kOmelyanLambda;
    launch_part_step(c, lam * h, 0.5 * h, true, kInParts, +1.0);
    for (int k = 0; k < m; ++k) {
      launch_part_step(c, (1.0 - 2.0 * lam) * h, 0.5 * h, true, kInParts, +1.0);
      if (k != m - 1) launch_part_step(c, 2.0 * lam * h, 0.5 * h, true, kInParts, +1.0);
    }
    launch_part_step(c, lam * h, 0.0, false, kInParts, +1.0);
    return;
  }
  launch_part_step(c, 0.5 * h, h, true, kInParts, +1.0);
  for (int k = 1; k < m; ++k) launch_part_step(c, h, h, true, kInParts, +1.0);
  launch_part_step(c, 0.5 * h, 0.0, false, kInParts, +1.0);
}
void launch_md_nested(ghx_ctx* c) {
  const int n = c->n_outer;
  if (c->integ == GHX_OMELYAN) {
    const double lam = kOmelyanLambda;
    for (int i = 0; i < n; ++i) {
      launch_part_step(c, i == 0 ? lam : 2.0 * lam, 0.0, false, kOutParts, -1.0);
      launch_inner(c, 0.5);
      launch_part_step(c, 1.0 - 2.0 * lam, 0.0, false, kOutParts, -1.0);
      launch_inner(c, 0.5);
    }
    launch_part_step(c, lam, 0.0, false, kOutParts, -1.0);
    return;
  }
  for (int i = 0; i < n; ++i) {
    launch_part_step(c, i == 0 ? 0.5 : 1.0, 0.0, false, kOutParts, -1.0);
    launch_inner(c, 1.0);
  }
  launch_part_step(c, 0.5, 0.0, false, kOutParts, -1.0);
}

void launch_md(ghx_ctx* c) {
  if (c->n_outer > 0) { launch_md_nested(c); return; }
  if (c->integ == GHX_LEAPFROG) {
    launch_step(c, 0.5, 1.0, true);
    for (int i = 1; i < c->nmd; ++i) launch_step(c, 1.0, 1.0, true);
    launch_step(c, 0.5, 0.0, false);
  } else {
    const double lam = kOmelyanLambda;
    launch_step(c, lam, 0.5, true);
    for (int i = 0; i < c->nmd; ++i) {
      launch_step(c, 1.0 - 2.0 * lam, 0.5, true);
      if (i != c->nmd - 1) launch_step(c, 2.0 * lam, 0.5, true);
    }
    launch_step(c, lam, 0.0, false);
  }
}

// Bring the current buffer back to buffer 0 (reunitarizing if requested).
// Device-to-device copy of `bytes` (a multiple of 16: every field is whole
// planes of Vp = 128 k sites) as a kernel, so that a captured trajectory graph
// holds kernel nodes only (no memcpy nodes).
__global__ __launch_bounds__(256) void copy16_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
  for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) dst[i] = src[i];
}
void stream_copy(ghx_ctx* c, void* dst, const void* src, size_t bytes) {
  const size_t n = bytes / 16;
  const unsigned nb = unsigned(std::min<size_t>((n + 255) / 256, 4096));
  if (n) hipLaunchKernelGGL(copy16_kernel, dim3(nb), dim3(256), 0, c->stream, static_cast<const uint4*>(src), static_cast<uint4*>(dst), n);
}

void launch_settle(ghx_ctx* c, bool reunit) {
  const LaunchCtx l = c->lc();
  if (reunit) {
    c->ops.reunit(l, c->U[c->cur], c->U[0]);
  } else if (c->cur != 0) {
    stream_copy(c, c->U[0], c->U[1], c->nU() * c->Vp * c->tsize);
  }
  if (c->cur != 0 && has_phi(c)) stream_copy(c, c->phi[0], c->phi[1], c->nphi() * c->Vp * c->tsize);
  c->cur = 0;
}

// Replica exchange step on buffer 0 (device-side decisions, then field swaps).
void launch_swaps(ghx_ctx* c) {
  if (!c->temper || c->npairs == 0) return;
  launch_action(c, 0);
  hipLaunchKernelGGL(swap_decide_kernel, dim3(1), dim3(1), 0, c->stream, c->dreps, c->N, c->sums, c->nrep, c->dpairs, c->npairs,
                     c->neven, c->dcfg, c->dq0, c->dq1, c->dswap, c->dattempts, c->daccepts, c->dflags);
  const unsigned nb = (c->maxvol + kBlock - 1) / kBlock;
  const int nphi = has_phi(c) ? int(c->nphi()) : 0;
  for (int pass = 0; pass < 2; ++pass) {
    const int k0 = pass ? c->neven : 0, k1 = pass ? c->npairs : c->neven;
    if (k1 <= k0) continue;
    if (c->tsize == 8)
      hipLaunchKernelGGL(swap_apply_kernel<double>, dim3(nb, k1 - k0), dim3(kBlock), 0, c->stream, c->dreps, c->dpairs, c->dflags, k0,
                         c->Vp, (double*)c->U[0], int(c->nU()), (double*)c->phi[0], nphi);
    else
      hipLaunchKernelGGL(swap_apply_kernel<float>, dim3(nb, k1 - k0), dim3(kBlock), 0, c->stream, c->dreps, c->dpairs, c->dflags, k0,
                         c->Vp, (float*)c->U[0], int(c->nU()), (float*)c->phi[0], nphi);
  }
}

// One trajectory of every replica, all on the stream.
void launch_trajectory(ghx_ctx* c) {
  const LaunchCtx l = c->lc();
  const bool hp = has_phi(c);
  c->ops.refresh(l, c->P, c->pi, c->mc, hp ? kTagP : kTagP_pure, kTagPi);
  if (c->frozen) c->ops.frozen_site(l, c->phi[c->cur], c->pi, 0);  // tangent momenta (gauge_higgs_hmc.hpp:172)
  stream_copy(c, c->Usave, c->U[0], c->nU() * c->Vp * c->tsize);
  if (hp) stream_copy(c, c->phisave, c->phi[0], c->nphi() * c->Vp * c->tsize);
  launch_hamiltonian(c, 0, c->Hi);
  launch_md(c);
  launch_settle(c, c->reunit);
  launch_hamiltonian(c, 0, c->Hf);
  hipLaunchKernelGGL(metropolis_kernel, dim3((c->nrep + 63) / 64), dim3(64), 0, c->stream, c->dreps, c->nrep, c->Hi, c->Hf, c->mc,
                     hp ? kTagMC : kTagMC_pure);
  if (c->tsize == 8)
    hipLaunchKernelGGL(restore_kernel<double>, dim3(c->nblk), dim3(kBlock), 0, c->stream, c->dreps, c->dblk_rep, c->mc, c->Vp,
                       (double*)c->U[0], (const double*)c->Usave, int(c->nU()), (double*)c->phi[0], (const double*)c->phisave,
                       hp ? int(c->nphi()) : 0);
  else
    hipLaunchKernelGGL(restore_kernel<float>, dim3(c->nblk), dim3(kBlock), 0, c->stream, c->dreps, c->dblk_rep, c->mc, c->Vp,
                       (float*)c->U[0], (const float*)c->Usave, int(c->nU()), (float*)c->phi[0], (const float*)c->phisave,
                       hp ? int(c->nphi()) : 0);
}

template <typename T>
void soa_to_aos(const std::vector<T>& soa, size_t Vp, unsigned off, unsigned vol, int nplane_per_site, double* out) {
  for (unsigned s = 0; s < vol; ++s)
    for (int c = 0; c < nplane_per_site; ++c) out[size_t(s) * nplane_per_site + c] = double(soa[size_t(c) * Vp + off + s]);
}
template <typename T>
void aos_to_soa(std::vector<T>& soa, size_t Vp, unsigned off, unsigned vol, int nplane_per_site, const double* in) {
  for (unsigned s = 0; s < vol; ++s)
    for (int c = 0; c < nplane_per_site; ++c) soa[size_t(c) * Vp + off + s] = T(in[size_t(s) * nplane_per_site + c]);
}

// The AoS site record [mu][...] maps to planes (mu*per_mu + c): identical ordering.
int get_planes(ghx_ctx* ctx, const void* dev, int nplanes, int replica, double* out) {
  const RepDesc& rd = ctx->hreps[replica];
  if (ctx->tsize == 8) {
    std::vector<double> h(size_t(nplanes) * ctx->Vp);
    HIPCHK(hipMemcpyAsync(h.data(), dev, h.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    soa_to_aos(h, ctx->Vp, rd.offset, rd.vol, nplanes, out);
  } else {
    std::vector<float> h(size_t(nplanes) * ctx->Vp);
    HIPCHK(hipMemcpyAsync(h.data(), dev, h.size() * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    soa_to_aos(h, ctx->Vp, rd.offset, rd.vol, nplanes, out);
  }
  return GHX_OK;
}
int set_planes(ghx_ctx* ctx, void* dev, int nplanes, int replica, const double* in) {
  const RepDesc& rd = ctx->hreps[replica];
  if (ctx->tsize == 8) {
    std::vector<double> h(size_t(nplanes) * ctx->Vp);
    HIPCHK(hipMemcpyAsync(h.data(), dev, h.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    aos_to_soa(h, ctx->Vp, rd.offset, rd.vol, nplanes, in);
    HIPCHK(hipMemcpyAsync(dev, h.data(), h.size() * 8, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  } else {
    std::vector<float> h(size_t(nplanes) * ctx->Vp);
    HIPCHK(hipMemcpyAsync(h.data(), dev, h.size() * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    aos_to_soa(h, ctx->Vp, rd.offset, rd.vol, nplanes, in);
    HIPCHK(hipMemcpyAsync(dev, h.data(), h.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  }
  return GHX_OK;
}

// Links in the reference AoS layout (N x N complex per link) from / to the
// stored rows; in FP32 MD the last row is rebuilt exactly as on the device.
void rebuild_last_row_host(int N, std::complex<double>* m) {
  if (N == 2) {
    m[2] = -std::conj(m[1]);
    m[3] = std::conj(m[0]);
  } else {
    for (int k = 0; k < 3; ++k) {
      const int i = (k + 1) % 3, j = (k + 2) % 3;
      m[6 + k] = std::conj(m[i] * m[3 + j] - m[j] * m[3 + i]);
    }
  }
}
int get_links(ghx_ctx* ctx, int replica, double* out) {
  const int nr = ctx->link_rows(), N = ctx->N;
  if (nr == N) return get_planes(ctx, ctx->U[ctx->cur], int(ctx->nU()), replica, out);
  const size_t nl = size_t(ctx->hreps[replica].vol) * ctx->D, per = size_t(2) * N * nr;
  std::vector<double> c(nl * per);
  int rc = get_planes(ctx, ctx->U[ctx->cur], int(ctx->nU()), replica, c.data());
  if (rc) return rc;
  for (size_t l = 0; l < nl; ++l) {
    std::complex<double>* m = reinterpret_cast<std::complex<double>*>(out + l * 2 * N * N);
    for (size_t q = 0; q < per / 2; ++q) m[q] = {c[l * per + 2 * q], c[l * per + 2 * q + 1]};
    rebuild_last_row_host(N, m);
  }
  return GHX_OK;
}
int set_links(ghx_ctx* ctx, int replica, const double* in) {
  const int nr = ctx->link_rows(), N = ctx->N;
  if (nr == N) return set_planes(ctx, ctx->U[ctx->cur], int(ctx->nU()), replica, in);
  const size_t nl = size_t(ctx->hreps[replica].vol) * ctx->D, per = size_t(2) * N * nr;
  std::vector<double> c(nl * per);
  for (size_t l = 0; l < nl; ++l)
    for (size_t q = 0; q < per; ++q) c[l * per + q] = in[l * 2 * N * N + q];
  return set_planes(ctx, ctx->U[ctx->cur], int(ctx->nU()), replica, c.data());
}

void drop_graph(ghx_ctx* c) {
  if (c->graph) { (void)hipGraphExecDestroy(c->graph); c->graph = nullptr; }
}

}  // namespace

namespace ghx {
int ctx_fail(ghx_ctx* c, int code, const std::string& msg) { return fail(c, code, msg); }
int ctx_check_ready(ghx_ctx* ctx) { return check_ready(ctx); }
void ctx_launch_forces(ghx_ctx* c, int b, int parts) {
  const LaunchCtx l = c->lc();
  if (c->big)
    c->bops.matter(l, c->tb, c->U[b], c->phi[b], nullptr, c->pi, c->G, c->Fphi, nullptr, 0.0, 0.0, F_ABS_EPS | parts, BIG_FORCE);
  c->ops.md_step(l, c->U[b], c->U[b ^ 1], c->P, c->phi[b], c->phi[b ^ 1], c->pi, c->F, c->Fphi, 0.0, 0.0, F_FORCE | F_ABS_EPS | parts);
}
}  // namespace ghx

extern "C" {

int ghx_create(ghx_ctx** out, int D, int N, int n_replicas, const int* L, int precision) {
  if (!out) return GHX_E_INVALID;
  *out = nullptr;
  if ((D != 3 && D != 4) || (N != 2 && N != 3) || n_replicas < 1 || !L || (precision != GHX_FP64 && precision != GHX_FP32MD))
    return GHX_E_INVALID;
  std::unique_ptr<ghx_ctx> c(new ghx_ctx);
  ghx_ctx* ctx = c.get();
  ctx->D = D; ctx->N = N; ctx->nrep = n_replicas; ctx->prec = precision;
  ctx->tsize = precision == GHX_FP64 ? 8 : 4;
  unsigned off = 0, red = 0;
  for (int r = 0; r < n_replicas; ++r) {
    RepDesc rd{};
    unsigned long long vol = 1;
    for (int m = 0; m < 4; ++m) rd.L[m] = 1;
    for (int m = 0; m < D; ++m) {
      const int e = L[r * D + m];
      if (e < 2) return GHX_E_INVALID;
      rd.L[m] = e;
      rd.stride[m] = unsigned(vol);
      vol *= unsigned(e);
    }
    if (vol > (1ull << 30)) return GHX_E_INVALID;
    rd.vol = unsigned(vol);
    rd.offset = off;
    rd.beta = 2.3; rd.kappa = 0.1; rd.lambda = 0.5; rd.tau = 1.0; rd.eps = 1.0 / 20;
    rd.seed = 1;
    const unsigned nb = (rd.vol + kBlock - 1) / kBlock;
    for (unsigned b = 0; b < nb; ++b) ctx->hblk_rep.push_back(r);
    off += nb * kBlock;
    const unsigned nr = (rd.vol + kRedBlock - 1) / kRedBlock;
    rd.blk_begin = red;
    for (unsigned b = 0; b < nr; ++b) { ctx->hred_rep.push_back(r); ctx->hred_first.push_back(rd.offset + b * kRedBlock); }
    red += nr;
    rd.blk_end = red;
    ctx->hreps.push_back(rd);
  }
  ctx->Vp = off;
  ctx->nblk = unsigned(ctx->hblk_rep.size());
  ctx->nred = red;
  HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  HIPCHK(hipMalloc((void**)&ctx->dreps, sizeof(RepDesc) * n_replicas));
  HIPCHK(hipMalloc((void**)&ctx->dblk_rep, sizeof(int) * ctx->nblk));
  HIPCHK(hipMalloc((void**)&ctx->dred_rep, sizeof(int) * ctx->nred));
  HIPCHK(hipMalloc((void**)&ctx->dred_first, sizeof(unsigned) * ctx->nred));
  HIPCHK(hipMalloc((void**)&ctx->partials, sizeof(double) * kNQ * ctx->nred));
  HIPCHK(hipMalloc((void**)&ctx->sums, sizeof(double) * kNQ * n_replicas));
  HIPCHK(hipMalloc((void**)&ctx->Hi, sizeof(double) * 5 * n_replicas));
  HIPCHK(hipMalloc((void**)&ctx->Hf, sizeof(double) * 5 * n_replicas));
  HIPCHK(hipMalloc((void**)&ctx->Hwork, sizeof(double) * 5 * n_replicas));
  HIPCHK(hipMalloc((void**)&ctx->mc, sizeof(McState) * n_replicas));
  HIPCHK(hipMemsetAsync(ctx->mc, 0, sizeof(McState) * n_replicas, ctx->stream));
  HIPCHK(hipMemcpyAsync(ctx->dblk_rep, ctx->hblk_rep.data(), sizeof(int) * ctx->nblk, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(ctx->dred_rep, ctx->hred_rep.data(), sizeof(int) * ctx->nred, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(ctx->dred_first, ctx->hred_first.data(), sizeof(unsigned) * ctx->nred, hipMemcpyHostToDevice, ctx->stream));
  int rc = upload_reps(ctx);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  *out = c.release();
  return GHX_OK;
}

int ghx_destroy(ghx_ctx* ctx) {
  if (!ctx) return GHX_E_INVALID;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  drop_graph(ctx);
  surrogate_free(ctx);
  free_multi(ctx);
  free_fields(ctx);
  free_tables(ctx);
  for (void* p : {(void*)ctx->dreps, (void*)ctx->dblk_rep, (void*)ctx->dred_rep, (void*)ctx->dred_first, (void*)ctx->partials,
                  (void*)ctx->sums, (void*)ctx->Hi, (void*)ctx->Hf, (void*)ctx->Hwork, (void*)ctx->mc, (void*)ctx->dpairs,
                  (void*)ctx->dflags, (void*)ctx->dcfg, (void*)ctx->dq0, (void*)ctx->dq1, (void*)ctx->dattempts,
                  (void*)ctx->daccepts, (void*)ctx->dswap})
    if (p) (void)hipFree(p);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return GHX_OK;
}

const char* ghx_last_error(const ghx_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int ghx_set_rep(ghx_ctx* ctx, const char* spec, int* d_out) {
  if (!ctx) return GHX_E_INVALID;
  try {
    ctx->rep = build_rep(ctx->N, spec ? spec : "");
  } catch (const std::exception& e) {
    ctx->ops_ok = false;
    return fail(ctx, GHX_E_REP, e.what());
  }
  drop_graph(ctx);
  // State sized or meant for the previous irrep goes with it: the plugged-in
  // potential (d^2 x d^2 superoperator, on-site planes), the surrogate and the
  // frozen-length flag. The caller sets them again for the new irrep.
  free_multi(ctx);
  surrogate_free(ctx);
  ctx->frozen = false;
  int rc = select_ops(ctx);
  if (rc) return rc;
  if ((rc = alloc_fields(ctx))) return rc;
  if (d_out) *d_out = ctx->rep.d;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return GHX_OK;
}

int ghx_rep_info(const ghx_ctx* ctx, int* d, int* nality, int* real_rep, long* dimT) {
  if (!ctx) return GHX_E_INVALID;
  if (d) *d = ctx->rep.d;
  if (nality) *nality = ctx->rep.nality;
  if (real_rep) *real_rep = ctx->rep.real ? 1 : 0;
  if (dimT) *dimT = ctx->rep.dimT;
  return GHX_OK;
}

int ghx_rep_generators(const ghx_ctx* ctx, double* T) {
  if (!ctx || !T) return GHX_E_INVALID;
  for (size_t i = 0; i < ctx->rep.T.size(); ++i) { T[2 * i] = ctx->rep.T[i].real(); T[2 * i + 1] = ctx->rep.T[i].imag(); }
  return GHX_OK;
}

int ghx_rep_basis(const ghx_ctx* ctx, double* W) {
  if (!ctx || !W) return GHX_E_INVALID;
  std::copy(ctx->rep.W.begin(), ctx->rep.W.end(), W);
  return GHX_OK;
}

int ghx_set_couplings(ghx_ctx* ctx, int replica, double beta, double kappa, double lambda, double tau) {
  if (!ctx || replica < 0 || replica >= ctx->nrep) return GHX_E_INVALID;
  RepDesc& rd = ctx->hreps[replica];
  rd.beta = beta; rd.kappa = kappa; rd.lambda = lambda; rd.tau = tau;
  rd.eps = tau / (ctx->n_outer > 0 ? ctx->n_outer : ctx->nmd);
  return upload_reps(ctx);
}

int ghx_set_integrator(ghx_ctx* ctx, int nmd, int integrator, int reunit_each_traj) {
  if (!ctx || nmd < 1 || (integrator != GHX_LEAPFROG && integrator != GHX_OMELYAN)) return GHX_E_INVALID;
  ctx->nmd = nmd; ctx->integ = integrator; ctx->reunit = reunit_each_traj != 0;
  for (auto& rd : ctx->hreps) rd.eps = rd.tau / (ctx->n_outer > 0 ? ctx->n_outer : nmd);
  drop_graph(ctx);
  return upload_reps(ctx);
}

int ghx_set_nested(ghx_ctx* ctx, int n_outer, int m_inner) {
  if (!ctx || n_outer < 0 || m_inner < 1) return GHX_E_INVALID;
  ctx->n_outer = n_outer;
  ctx->m_inner = m_inner;
  for (auto& rd : ctx->hreps) rd.eps = rd.tau / (n_outer > 0 ? n_outer : ctx->nmd);
  drop_graph(ctx);
  return upload_reps(ctx);
}

int ghx_set_nested_inner(ghx_ctx* ctx, int integrator) {
  if (!ctx || (integrator != GHX_LEAPFROG && integrator != GHX_OMELYAN)) return GHX_E_INVALID;
  ctx->inner_integ = integrator;
  drop_graph(ctx);
  return GHX_OK;
}

int ghx_set_seed(ghx_ctx* ctx, int replica, uint64_t seed) {
  if (!ctx || replica < 0 || replica >= ctx->nrep) return GHX_E_INVALID;
  ctx->hreps[replica].seed = seed;
  return upload_reps(ctx);
}

int ghx_init_fields(ghx_ctx* ctx, int replica, int link_mode, double sigma_U, int phi_mode, uint64_t phi_stream, double phi_sigma) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (replica < 0 || replica >= ctx->nrep) return fail(ctx, GHX_E_INVALID, "replica out of range");
  const LaunchCtx l = ctx->lc();
  if (ctx->cur != 0) launch_settle(ctx, false);
  ctx->ops.init_links(l, ctx->U[0], replica, link_mode == GHX_HOT ? 1 : 0, sigma_U);
  if (has_phi(ctx)) {
    if (ctx->tsize == 8)
      hipLaunchKernelGGL(init_phi_kernel<double>, dim3(ctx->nblk), dim3(kBlock), 0, ctx->stream, ctx->dreps, ctx->dblk_rep, ctx->Vp,
                         (double*)ctx->phi[0], ctx->rep.d, ctx->rep.real ? 1 : 0, replica, phi_mode ? 1 : 0,
                         (unsigned long long)phi_stream, phi_sigma);
    else
      hipLaunchKernelGGL(init_phi_kernel<float>, dim3(ctx->nblk), dim3(kBlock), 0, ctx->stream, ctx->dreps, ctx->dblk_rep, ctx->Vp,
                         (float*)ctx->phi[0], ctx->rep.d, ctx->rep.real ? 1 : 0, replica, phi_mode ? 1 : 0,
                         (unsigned long long)phi_stream, phi_sigma);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return GHX_OK;
}

int ghx_set_counters(ghx_ctx* ctx, int replica, uint64_t traj_count, uint64_t accept_count) {
  if (!ctx || replica < 0 || replica >= ctx->nrep) return GHX_E_INVALID;
  McState m{};
  HIPCHK(hipMemcpyAsync(&m, ctx->mc + replica, sizeof m, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  m.traj_count = traj_count; m.accept_count = accept_count;
  HIPCHK(hipMemcpyAsync(ctx->mc + replica, &m, sizeof m, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return GHX_OK;
}

int ghx_get_counters(const ghx_ctx* cctx, int replica, uint64_t* traj_count, uint64_t* accept_count, double* last_dH) {
  ghx_ctx* ctx = const_cast<ghx_ctx*>(cctx);
  if (!ctx || replica < 0 || replica >= ctx->nrep) return GHX_E_INVALID;
  McState m{};
  HIPCHK(hipMemcpyAsync(&m, ctx->mc + replica, sizeof m, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (traj_count) *traj_count = m.traj_count;
  if (accept_count) *accept_count = m.accept_count;
  if (last_dH) *last_dH = m.last_dH;
  return GHX_OK;
}

int ghx_refresh_momenta(ghx_ctx* ctx) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  ctx->ops.refresh(ctx->lc(), ctx->P, ctx->pi, ctx->mc, has_phi(ctx) ? kTagP : kTagP_pure, kTagPi);
  if (ctx->frozen) ctx->ops.frozen_site(ctx->lc(), ctx->phi[ctx->cur], ctx->pi, 0);
  HIPCHK(hipGetLastError());
  return GHX_OK;
}

// Frozen-length scalar |phi_x| = 1 (GaugeHiggsHMC::frozen_phi, gauge_higgs_hmc.hpp:38, 61-103).
int ghx_set_frozen_phi(ghx_ctx* ctx, int on) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (on && !has_phi(ctx)) return fail(ctx, GHX_E_INVALID, "frozen-length scalar needs a scalar field");
  ctx->frozen = on != 0;
  drop_graph(ctx);
  return GHX_OK;
}
// normalize_phi (gauge_higgs_hmc.hpp:74-83): phi_x /= |phi_x| on every replica.
int ghx_normalize_phi(ghx_ctx* ctx) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (!has_phi(ctx)) return fail(ctx, GHX_E_INVALID, "no scalar field");
  ctx->ops.frozen_site(ctx->lc(), ctx->phi[ctx->cur], ctx->pi, 1);
  HIPCHK(hipGetLastError());
  return GHX_OK;
}

int ghx_kick(ghx_ctx* ctx, double eps) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  const LaunchCtx l = ctx->lc();
  const int b = ctx->cur;
  if (ctx->big)
    ctx->bops.matter(l, ctx->tb, ctx->U[b], ctx->phi[b], ctx->phi[b ^ 1], ctx->pi, ctx->G, nullptr, nullptr, eps, 0.0, F_KICK | F_ABS_EPS,
                     BIG_STEP);
  ctx->ops.md_step(l, ctx->U[b], ctx->U[b ^ 1], ctx->P, ctx->phi[b], ctx->phi[b ^ 1], ctx->pi, ctx->F, ctx->Fphi, eps, 0.0,
                   F_KICK | F_ABS_EPS);
  HIPCHK(hipGetLastError());
  return GHX_OK;
}

int ghx_drift(ghx_ctx* ctx, double eps) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  const LaunchCtx l = ctx->lc();
  const int b = ctx->cur;
  ctx->ops.drift(l, ctx->U[b], ctx->U[b ^ 1], ctx->P, ctx->phi[b], ctx->phi[b ^ 1], ctx->pi, eps, F_DRIFT | F_ABS_EPS);
  if (ctx->big) ctx->bops.phi_drift(l, ctx->phi[b], ctx->phi[b ^ 1], ctx->pi, int(ctx->nphi()), eps, 1);
  ctx->cur ^= 1;
  HIPCHK(hipGetLastError());
  return GHX_OK;
}

int ghx_md_evolve(ghx_ctx* ctx) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  launch_md(ctx);
  HIPCHK(hipGetLastError());
  return GHX_OK;
}

int ghx_hamiltonian(ghx_ctx* ctx, double* H_parts) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  launch_hamiltonian(ctx, ctx->cur, ctx->Hwork);
  HIPCHK(hipGetLastError());
  if (H_parts) {
    HIPCHK(hipMemcpyAsync(H_parts, ctx->Hwork, sizeof(double) * 5 * ctx->nrep, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  }
  return GHX_OK;
}

int ghx_trajectory(ghx_ctx* ctx, int* accepted, double* dH) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (ctx->cur != 0) launch_settle(ctx, false);
  launch_trajectory(ctx);
  HIPCHK(hipGetLastError());
  if (accepted || dH) {
    std::vector<McState> m(ctx->nrep);
    HIPCHK(hipMemcpyAsync(m.data(), ctx->mc, sizeof(McState) * ctx->nrep, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (int r = 0; r < ctx->nrep; ++r) {
      if (accepted) accepted[r] = m[r].accepted;
      if (dH) dH[r] = m[r].last_dH;
    }
    for (int r = 0; r < ctx->nrep; ++r)
      if (!std::isfinite(m[r].last_dH)) return fail(ctx, GHX_E_NAN, "non-finite dH in replica " + std::to_string(r));
  }
  return GHX_OK;
}

int ghx_trajectories(ghx_ctx* ctx, int n, int use_graph) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if ((rc = check_ladders(ctx))) return rc;
  if (ctx->cur != 0) launch_settle(ctx, false);
  if (!use_graph) {
    for (int i = 0; i < n; ++i) { launch_trajectory(ctx); launch_swaps(ctx); }
    HIPCHK(hipGetLastError());
    return GHX_OK;
  }
  if (!ctx->graph) {
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    launch_trajectory(ctx);
    launch_swaps(ctx);
    HIPCHK(hipStreamEndCapture(ctx->stream, &g));
    HIPCHK(hipGraphInstantiate(&ctx->graph, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
  }
  for (int i = 0; i < n; ++i) HIPCHK(hipGraphLaunch(ctx->graph, ctx->stream));
  return GHX_OK;
}

int ghx_observables(ghx_ctx* ctx, double* out) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (!out) return GHX_E_INVALID;
  const int b = ctx->cur;
  launch_action(ctx, b);
  std::vector<double> s(size_t(kNQ) * ctx->nrep), pl(ctx->nrep);
  HIPCHK(hipMemcpyAsync(s.data(), ctx->sums, sizeof(double) * s.size(), hipMemcpyDeviceToHost, ctx->stream));
  const LaunchCtx l = ctx->lc();
  ctx->ops.polyakov(l, ctx->U[b]);
  hipLaunchKernelGGL(finalize_kernel, dim3(ctx->nrep), dim3(kRedBlock), 0, ctx->stream, ctx->dreps, ctx->partials, 1, ctx->Hwork);
  HIPCHK(hipMemcpyAsync(pl.data(), ctx->Hwork, sizeof(double) * ctx->nrep, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (int r = 0; r < ctx->nrep; ++r) {
    const RepDesc& rd = ctx->hreps[r];
    const double* q = &s[size_t(r) * kNQ];
    const double V = rd.vol, nplaq = V * (ctx->D * (ctx->D - 1) / 2);
    double* o = out + size_t(r) * 6;
    o[0] = q[0] / (ctx->N * nplaq);
    o[1] = (rd.beta / ctx->N) * (ctx->N * nplaq - q[0]);
    o[2] = has_phi(ctx) ? q[2] - 2.0 * rd.kappa * q[1] : 0.0;
    o[3] = has_phi(ctx) ? q[1] / (V * ctx->D) : 0.0;
    o[4] = has_phi(ctx) ? q[3] / V : 0.0;
    o[5] = pl[r] / ((V / rd.L[ctx->D - 1]) * ctx->N);
  }
  return GHX_OK;
}

namespace {
// Probe irrep of a measurement from a rep spec; the Young-row rules and
// messages are those of the scalar's representation. Returns d_R' or 0 (with
// the context's error set).
long parse_probe(ghx_ctx* ctx, const char* spec, Probe* p) {
  const std::string s = spec ? spec : "";
  *p = Probe{0, 0, {0, 0}};
  if (s == "fund") { p->kind = 1; return ctx->N; }
  if (s == "adj") { p->kind = 2; return long(ctx->N) * ctx->N - 1; }
  try {
    const std::vector<int> rows = parse_young_rows(ctx->N, s, nullptr);
    if (rows[0] + 1 > kProbeMaxH) throw std::runtime_error("probe rep: first row longer than " + std::to_string(kProbeMaxH - 1));
    p->kind = 3;
    p->nrows = int(rows.size());
    for (size_t i = 0; i < rows.size(); ++i) p->rows[i] = rows[i];
    return young_dimension(ctx->N, rows);
  } catch (const std::exception& ex) {
    fail(ctx, GHX_E_REP, ex.what());
    return 0;
  }
}

// Sum the per-block partials of every replica (fixed order) and copy them out.
int replica_sums(ghx_ctx* ctx, std::vector<double>& out) {
  hipLaunchKernelGGL(finalize_kernel, dim3(ctx->nrep), dim3(kRedBlock), 0, ctx->stream, ctx->dreps, ctx->partials, 1, ctx->Hwork);
  out.assign(ctx->nrep, 0.0);
  HIPCHK(hipMemcpyAsync(out.data(), ctx->Hwork, sizeof(double) * ctx->nrep, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return GHX_OK;
}
}  // namespace

int ghx_wilson_loop(ghx_ctx* ctx, const char* probe, int Ra, int Rb, double* W) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (!W || Ra < 0 || Rb < 0) return GHX_E_INVALID;
  Probe p;
  const long d = parse_probe(ctx, probe, &p);
  if (d == 0) return GHX_E_REP;
  ctx->mops.wilson(ctx->lc(), ctx->U[ctx->cur], p, Ra, Rb);
  HIPCHK(hipGetLastError());
  std::vector<double> s;
  if ((rc = replica_sums(ctx, s))) return rc;
  for (int r = 0; r < ctx->nrep; ++r) {
    const double nloop = double(ctx->hreps[r].vol) * (ctx->D * (ctx->D - 1) / 2);
    W[r] = s[r] / (nloop * double(d));
  }
  return GHX_OK;
}

int ghx_polyakov_loop_rep(ghx_ctx* ctx, const char* probe, int tdir, double* P) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (!P || tdir < 0 || tdir >= ctx->D) return GHX_E_INVALID;
  Probe p;
  const long d = parse_probe(ctx, probe, &p);
  if (d == 0) return GHX_E_REP;
  ctx->mops.polyakov(ctx->lc(), ctx->U[ctx->cur], p, tdir);
  HIPCHK(hipGetLastError());
  std::vector<double> s;
  if ((rc = replica_sums(ctx, s))) return rc;
  for (int r = 0; r < ctx->nrep; ++r) {
    const RepDesc& rd = ctx->hreps[r];
    P[r] = s[r] / (double(rd.vol / unsigned(rd.L[tdir])) * double(d));
  }
  return GHX_OK;
}

int ghx_max_unitarity(ghx_ctx* ctx, int replica, double* out) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (replica < 0 || replica >= ctx->nrep || !out) return GHX_E_INVALID;
  ctx->ops.unitarity(ctx->lc(), ctx->U[ctx->cur]);
  const RepDesc& rd = ctx->hreps[replica];
  std::vector<double> p(rd.blk_end - rd.blk_begin);
  HIPCHK(hipMemcpyAsync(p.data(), ctx->partials + rd.blk_begin, sizeof(double) * p.size(), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  *out = p.empty() ? 0.0 : *std::max_element(p.begin(), p.end());
  return GHX_OK;
}

int ghx_reunitarize(ghx_ctx* ctx) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  launch_settle(ctx, true);
  HIPCHK(hipGetLastError());
  return GHX_OK;
}

int ghx_get_fields(ghx_ctx* ctx, int replica, double* U, double* P, double* phi, double* pi) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (replica < 0 || replica >= ctx->nrep) return GHX_E_INVALID;
  if (U && (rc = get_links(ctx, replica, U))) return rc;
  if (P && (rc = get_planes(ctx, ctx->P, int(ctx->nP()), replica, P))) return rc;
  if (has_phi(ctx)) {
    if (phi && (rc = get_planes(ctx, ctx->phi[ctx->cur], int(ctx->nphi()), replica, phi))) return rc;
    if (pi && (rc = get_planes(ctx, ctx->pi, int(ctx->nphi()), replica, pi))) return rc;
  }
  return GHX_OK;
}

int ghx_set_fields(ghx_ctx* ctx, int replica, const double* U, const double* P, const double* phi, const double* pi) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (replica < 0 || replica >= ctx->nrep) return GHX_E_INVALID;
  if (U && (rc = set_links(ctx, replica, U))) return rc;
  if (P && (rc = set_planes(ctx, ctx->P, int(ctx->nP()), replica, P))) return rc;
  if (has_phi(ctx)) {
    if (phi && (rc = set_planes(ctx, ctx->phi[ctx->cur], int(ctx->nphi()), replica, phi))) return rc;
    if (pi && (rc = set_planes(ctx, ctx->pi, int(ctx->nphi()), replica, pi))) return rc;
  }
  return GHX_OK;
}

int ghx_forces(ghx_ctx* ctx, int replica, double* F_link, double* F_phi) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (replica < 0 || replica >= ctx->nrep) return GHX_E_INVALID;
  const LaunchCtx l = ctx->lc();
  const int b = ctx->cur;
  if (ctx->big)
    ctx->bops.matter(l, ctx->tb, ctx->U[b], ctx->phi[b], nullptr, ctx->pi, ctx->G, ctx->Fphi, nullptr, 0.0, 0.0, F_ABS_EPS, BIG_FORCE);
  ctx->ops.md_step(l, ctx->U[b], ctx->U[b ^ 1], ctx->P, ctx->phi[b], ctx->phi[b ^ 1], ctx->pi, ctx->F, ctx->Fphi, 0.0, 0.0,
                   F_FORCE | F_ABS_EPS);
  HIPCHK(hipGetLastError());
  if (F_link && (rc = get_planes(ctx, ctx->F, int(ctx->nP()), replica, F_link))) return rc;
  if (F_phi && has_phi(ctx) && (rc = get_planes(ctx, ctx->Fphi, int(ctx->nphi()), replica, F_phi))) return rc;
  return GHX_OK;
}

int ghx_replica_swap(ghx_ctx* ctx, int i, int j, double u, int* accepted, double* delta) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (i < 0 || j < 0 || i >= ctx->nrep || j >= ctx->nrep || i == j) return GHX_E_INVALID;
  const RepDesc &ri = ctx->hreps[i], &rj = ctx->hreps[j];
  for (int m = 0; m < ctx->D; ++m)
    if (ri.L[m] != rj.L[m]) return fail(ctx, GHX_E_INVALID, "replica swap needs equal extents");
  if (ctx->cur != 0) launch_settle(ctx, false);
  launch_action(ctx, 0);
  std::vector<double> s(size_t(kNQ) * ctx->nrep);
  HIPCHK(hipMemcpyAsync(s.data(), ctx->sums, sizeof(double) * s.size(), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  // S_r(C) = (beta_r/N)(N nplaq - sumReTr(C)) + sum V(C) - 2 kappa_r B(C); only the
  // coupling-dependent terms survive in S_i(C_j)+S_j(C_i)-S_i(C_i)-S_j(C_j).
  const double *qi = &s[size_t(i) * kNQ], *qj = &s[size_t(j) * kNQ];
  if (has_phi(ctx) && ri.lambda != rj.lambda) return fail(ctx, GHX_E_INVALID, "replica swap needs equal lambda");
  const double dgauge = (ri.beta - rj.beta) / ctx->N * (qi[0] - qj[0]);
  const double dkappa = has_phi(ctx) ? (ri.kappa - rj.kappa) * 2.0 * (qi[1] - qj[1]) : 0.0;
  const double Delta = dgauge + dkappa;
  const bool acc = (Delta <= 0.0) || (u < std::exp(-Delta));
  if (delta) *delta = Delta;
  if (accepted) *accepted = acc ? 1 : 0;
  if (!acc) return GHX_OK;
  // exchange U and phi site ranges (same volume): planes are strided by Vp
  const size_t ts = ctx->tsize;
  void* tmp = nullptr;
  HIPCHK(hipMalloc(&tmp, size_t(ri.vol) * ts));
  auto swap_planes = [&](void* base, size_t nplanes) -> int {
    for (size_t c = 0; c < nplanes; ++c) {
      char* pi_ = (char*)base + (c * ctx->Vp + ri.offset) * ts;
      char* pj_ = (char*)base + (c * ctx->Vp + rj.offset) * ts;
      HIPCHK(hipMemcpyAsync(tmp, pi_, size_t(ri.vol) * ts, hipMemcpyDeviceToDevice, ctx->stream));
      HIPCHK(hipMemcpyAsync(pi_, pj_, size_t(ri.vol) * ts, hipMemcpyDeviceToDevice, ctx->stream));
      HIPCHK(hipMemcpyAsync(pj_, tmp, size_t(ri.vol) * ts, hipMemcpyDeviceToDevice, ctx->stream));
    }
    return GHX_OK;
  };
  if ((rc = swap_planes(ctx->U[0], ctx->nU()))) return rc;
  if (has_phi(ctx) && (rc = swap_planes(ctx->phi[0], ctx->nphi()))) return rc;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipFree(tmp));
  return GHX_OK;
}

int ghx_set_ladders(ghx_ctx* ctx, int n_ladders, const int* lengths, const int* members, uint64_t swap_seed) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (n_ladders < 0 || (n_ladders > 0 && (!lengths || !members))) return fail(ctx, GHX_E_INVALID, "bad ladder arguments");
  std::vector<std::vector<int>> lad;
  std::vector<int> seen(ctx->nrep, 0);
  int off = 0;
  for (int l = 0; l < n_ladders; ++l) {
    std::vector<int> m(members + off, members + off + lengths[l]);
    off += lengths[l];
    for (int r : m) {
      if (r < 0 || r >= ctx->nrep || seen[r]++) return fail(ctx, GHX_E_INVALID, "ladder members must be distinct replicas");
    }
    for (size_t k = 1; k < m.size(); ++k) {
      const RepDesc &a = ctx->hreps[m[k - 1]], &b = ctx->hreps[m[k]];
      for (int d = 0; d < ctx->D; ++d)
        if (a.L[d] != b.L[d]) return fail(ctx, GHX_E_INVALID, "replicas of one ladder need equal extents");
      if (has_phi(ctx) && a.lambda != b.lambda) return fail(ctx, GHX_E_INVALID, "replicas of one ladder need equal lambda");
    }
    lad.push_back(m);
  }
  ctx->hpairs.clear();
  for (int pass = 0; pass < 2; ++pass) {
    for (const auto& m : lad)
      for (size_t p = size_t(pass); p + 1 < m.size(); p += 2) { ctx->hpairs.push_back(m[p]); ctx->hpairs.push_back(m[p + 1]); }
    if (pass == 0) ctx->neven = int(ctx->hpairs.size() / 2);
  }
  ctx->npairs = int(ctx->hpairs.size() / 2);
  for (void** p : {(void**)&ctx->dpairs, (void**)&ctx->dflags, (void**)&ctx->dcfg, (void**)&ctx->dq0, (void**)&ctx->dq1,
                   (void**)&ctx->dattempts, (void**)&ctx->daccepts, (void**)&ctx->dswap})
    if (*p) { (void)hipFree(*p); *p = nullptr; }
  const int np = std::max(ctx->npairs, 1);
  HIPCHK(hipMalloc((void**)&ctx->dpairs, sizeof(int) * 2 * np));
  HIPCHK(hipMalloc((void**)&ctx->dflags, sizeof(int) * np));
  HIPCHK(hipMalloc((void**)&ctx->dcfg, sizeof(int) * ctx->nrep));
  HIPCHK(hipMalloc((void**)&ctx->dq0, sizeof(double) * ctx->nrep));
  HIPCHK(hipMalloc((void**)&ctx->dq1, sizeof(double) * ctx->nrep));
  HIPCHK(hipMalloc((void**)&ctx->dattempts, sizeof(unsigned long long) * np));
  HIPCHK(hipMalloc((void**)&ctx->daccepts, sizeof(unsigned long long) * np));
  HIPCHK(hipMalloc((void**)&ctx->dswap, sizeof(SwapState)));
  if (ctx->npairs) HIPCHK(hipMemcpyAsync(ctx->dpairs, ctx->hpairs.data(), sizeof(int) * ctx->hpairs.size(), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemsetAsync(ctx->dflags, 0, sizeof(int) * np, ctx->stream));
  HIPCHK(hipMemsetAsync(ctx->dattempts, 0, sizeof(unsigned long long) * np, ctx->stream));
  HIPCHK(hipMemsetAsync(ctx->daccepts, 0, sizeof(unsigned long long) * np, ctx->stream));
  const SwapState st{swap_seed, 0};
  HIPCHK(hipMemcpyAsync(ctx->dswap, &st, sizeof st, hipMemcpyHostToDevice, ctx->stream));
  ctx->maxvol = 0;
  for (const auto& rd : ctx->hreps) ctx->maxvol = std::max(ctx->maxvol, rd.vol);
  ctx->temper = ctx->npairs > 0;
  drop_graph(ctx);
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return GHX_OK;
}

int ghx_set_tempering(ghx_ctx* ctx, int enabled) {
  if (!ctx) return GHX_E_INVALID;
  ctx->temper = enabled && ctx->npairs > 0;
  drop_graph(ctx);
  return GHX_OK;
}

int ghx_swap_step(ghx_ctx* ctx) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if ((rc = check_ladders(ctx))) return rc;
  if (ctx->cur != 0) launch_settle(ctx, false);
  launch_swaps(ctx);
  HIPCHK(hipGetLastError());
  return GHX_OK;
}

int ghx_swap_stats(ghx_ctx* ctx, int* npairs, int* pairs, uint64_t* attempts, uint64_t* accepts, int* last_flags) {
  if (!ctx) return GHX_E_INVALID;
  if (npairs) *npairs = ctx->npairs;
  if (ctx->npairs == 0) return GHX_OK;
  if (pairs) std::copy(ctx->hpairs.begin(), ctx->hpairs.end(), pairs);
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (attempts) HIPCHK(hipMemcpy(attempts, ctx->dattempts, sizeof(uint64_t) * ctx->npairs, hipMemcpyDeviceToHost));
  if (accepts) HIPCHK(hipMemcpy(accepts, ctx->daccepts, sizeof(uint64_t) * ctx->npairs, hipMemcpyDeviceToHost));
  if (last_flags) HIPCHK(hipMemcpy(last_flags, ctx->dflags, sizeof(int) * ctx->npairs, hipMemcpyDeviceToHost));
  return GHX_OK;
}

int ghx_set_tuning(ghx_ctx* ctx, int key, int value) {
  if (!ctx) return GHX_E_INVALID;
  if (key == GHX_TUNE_STREAM_SEG) {
    if (value < 0) return fail(ctx, GHX_E_INVALID, "stream segment must be >= 0");
    ctx->stream_seg = value;
    drop_graph(ctx);
    return GHX_OK;
  }
  if (key == GHX_TUNE_BLOCK_ORDER) {
    if (value < 0 || value > 6) return fail(ctx, GHX_E_INVALID, "block order must be 0 to 6");
    ctx->block_order = value;
    drop_graph(ctx);
    return GHX_OK;
  }
  if (key == GHX_TUNE_LINK_LANES) {
    if (value < 0 || value > 1) return fail(ctx, GHX_E_INVALID, "link lanes must be 0 or 1");
    ctx->link_lanes = value;
    drop_graph(ctx);
    return GHX_OK;
  }
  if (key == GHX_TUNE_SYM_ENGINE) {
    if (value < 0 || value > 1) return fail(ctx, GHX_E_INVALID, "symmetric-power engine must be 0 or 1");
    const bool was = ctx->sym_enabled;
    ctx->sym_enabled = value != 0;
    ctx->tb.sym.on = (ctx->sym_valid && ctx->sym_enabled) ? 1 : 0;
    if (int rc = check_big_lds(ctx)) {  // the tensor apply does not fit: keep the engine
      ctx->sym_enabled = was;
      ctx->tb.sym.on = (ctx->sym_valid && ctx->sym_enabled) ? 1 : 0;
      return rc;
    }
    drop_graph(ctx);
    return GHX_OK;
  }
  if (key == GHX_TUNE_SYM_PAIRS) {
    if (value < 0 || value > 2) return fail(ctx, GHX_E_INVALID, "the symmetric-power link kernel must be 0, 1 or 2");
    const int was = ctx->sym_pair_mode;
    ctx->sym_pair_mode = value;
    if (int rc = check_big_lds(ctx)) {
      ctx->sym_pair_mode = was;
      check_big_lds(ctx);
      return rc;
    }
    drop_graph(ctx);
    return GHX_OK;
  }
  if (key == GHX_TUNE_X_LANES) {
    if (value < 0 || value > 1) return fail(ctx, GHX_E_INVALID, "x lanes must be 0 or 1");
    ctx->xlane_enabled = value != 0;
    drop_graph(ctx);
    return GHX_OK;
  }
  if (key == GHX_TUNE_FUSE_INNER) {
    if (value < 0 || value > 1) return fail(ctx, GHX_E_INVALID, "fused inner step must be 0 or 1");
    ctx->fuse_inner = value != 0;
    drop_graph(ctx);
    return GHX_OK;
  }
  if (key == GHX_TUNE_LINK_SPLIT) {
    if (value < 0 || value > 1) return fail(ctx, GHX_E_INVALID, "link split must be 0 or 1");
    ctx->link_split = value != 0;
    drop_graph(ctx);
    return GHX_OK;
  }
  if (key == GHX_TUNE_STORE_POLICY) {
    if (value < 0 || value > 1) return fail(ctx, GHX_E_INVALID, "store policy must be 0 (streamed) or 1 (cached)");
    ctx->store_cached = value;
    drop_graph(ctx);
    return GHX_OK;
  }
  if (key == GHX_TUNE_SPLIT_INNER) {
    if (value < 0 || value > 1) return fail(ctx, GHX_E_INVALID, "split inner scale must be 0 or 1");
    ctx->split_inner = value != 0;
    drop_graph(ctx);
    return GHX_OK;
  }
  return fail(ctx, GHX_E_INVALID, "unknown tuning key");
}

int ghx_get_tuning(const ghx_ctx* ctx, int key, int* value) {
  if (!ctx || !value) return GHX_E_INVALID;
  if (key == GHX_TUNE_STREAM_SEG) { *value = ctx->stream_seg; return GHX_OK; }
  if (key == GHX_TUNE_BLOCK_ORDER) { *value = ctx->block_walk().order; return GHX_OK; }
  if (key == GHX_TUNE_LINK_LANES) { *value = ctx->link_lanes; return GHX_OK; }
  if (key == GHX_TUNE_SYM_ENGINE) { *value = ctx->tb.sym.on; return GHX_OK; }
  if (key == GHX_TUNE_SYM_PAIRS) { *value = ctx->tb.sym.on ? ctx->tb.sym.pair : 0; return GHX_OK; }
  if (key == GHX_TUNE_FUSE_INNER) {  // whether a large irrep's inner-scale step runs as one launch
    *value = ctx->fuse_inner && ctx->big && ctx->D == 4 && ctx->link_lanes && !(ctx->nrep == 1 && ctx->stream_seg > 0);
    return GHX_OK;
  }
  if (key == GHX_TUNE_SPLIT_INNER) { *value = inner_split_applies(ctx) ? 1 : 0; return GHX_OK; }
  if (key == GHX_TUNE_STORE_POLICY) { *value = ctx->store_cached; return GHX_OK; }
  if (key == GHX_TUNE_LINK_SPLIT) {  // whether the link-lane step splits the staples: link lanes, no scalar in the kernel, D >= 3
    *value = ctx->link_split && ctx->link_lanes && (ctx->big || ctx->rep.d == 0) && ctx->D >= 3 && !(ctx->nrep == 1 && ctx->stream_seg > 0);
    return GHX_OK;
  }
  if (key == GHX_TUNE_X_LANES) {  // whether the MD step uses it (not with link lanes or the 4D-tile walks)
    *value = ctx->xlane_enabled && ctx->xlane_ok() && !ctx->link_lanes && ctx->block_walk().order < 3;
    return GHX_OK;
  }
  return GHX_E_INVALID;
}

// Time each candidate launch shape on the live fields (state saved and
// restored around the trial steps) and keep the fastest.
int ghx_autotune(ghx_ctx* ctx, int steps, int* best_seg, double* best_ms) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (steps < 1) return GHX_E_INVALID;
  if (ctx->cur != 0) launch_settle(ctx, false);
  const size_t nUb = ctx->nU() * ctx->Vp * ctx->tsize, nPb = ctx->nP() * ctx->Vp * ctx->tsize;
  const size_t nphb = has_phi(ctx) ? ctx->nphi() * ctx->Vp * ctx->tsize : 0;
  void *Ps = nullptr, *pis = nullptr;
  HIPCHK(hipMalloc(&Ps, nPb));
  if (nphb) HIPCHK(hipMalloc(&pis, nphb));
  HIPCHK(hipMemcpyAsync(ctx->Usave, ctx->U[0], nUb, hipMemcpyDeviceToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(Ps, ctx->P, nPb, hipMemcpyDeviceToDevice, ctx->stream));
  if (nphb) {
    HIPCHK(hipMemcpyAsync(ctx->phisave, ctx->phi[0], nphb, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(pis, ctx->pi, nphb, hipMemcpyDeviceToDevice, ctx->stream));
  }
  // candidates (stream segment, block order or -1 for the link-parallel step),
  // each with streamed and with cached stores; the walks only for one site per
  // thread
  std::vector<std::pair<int, int>> shapes = {{0, 0}};
  for (int order : {1, 2}) {
    ctx->block_order = order;
    if (ctx->block_walk().order == order) shapes.push_back({0, order});
  }
  shapes.push_back({0, -1});
  if (ctx->nrep == 1)
    for (int sg : {1, 2, 4, 8, 16, 32, 64})
      if (sg <= ctx->hreps[0].L[ctx->D - 1]) shapes.push_back({sg, 0});
  struct Cand {
    int first, second, store, split;
  };
  // the link-lane shape also with and without its staples split over planes
  // (where the kernel carries no scalar)
  const bool can_split = (ctx->big || ctx->rep.d == 0) && ctx->D >= 3;
  std::vector<Cand> cand;
  for (int st : {0, 1})
    for (const auto& sh : shapes) {
      cand.push_back({sh.first, sh.second, st, 0});
      if (sh.second < 0 && can_split) cand.push_back({sh.first, sh.second, st, 1});
    }
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  auto use = [&](const Cand& c) {
    ctx->stream_seg = c.first;
    ctx->block_order = c.second < 0 ? 0 : c.second;
    ctx->link_lanes = c.second < 0 ? 1 : 0;
    ctx->store_cached = c.store;
    ctx->link_split = c.split != 0;
  };
  auto time_it = [&](int n, double& out) -> int {  // GHX_OK or an error code
    launch_step(ctx, 1.0, 1.0, true);  // the shape's first launch stays out of the timing
    HIPCHK(hipEventRecord(e0, ctx->stream));
    for (int i = 0; i < n; ++i) launch_step(ctx, 1.0, 1.0, true);
    HIPCHK(hipEventRecord(e1, ctx->stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    out = ms;
    return GHX_OK;
  };
  // The GPU runs a newly started kernel slower for its first few tens of
  // milliseconds, so the first candidate runs untimed first, the two fastest
  // are timed again, and the winner runs last: the shape is chosen on warm
  // timings and the caller's next steps start warm.
  use(cand[0]);
  for (int i = 0; i < steps; ++i) launch_step(ctx, 1.0, 1.0, true);
  std::vector<double> ms(cand.size());
  for (size_t k = 0; k < cand.size() && rc == GHX_OK; ++k) {
    use(cand[k]);
    rc = time_it(steps, ms[k]);
  }
  size_t b0 = 0, b1 = cand.size() > 1 ? 1 : 0;
  for (size_t k = 0; k < cand.size(); ++k)
    if (ms[k] < ms[b0]) b0 = k;
  for (size_t k = 0; k < cand.size(); ++k)
    if (k != b0 && (b1 == b0 || ms[k] < ms[b1])) b1 = k;
  if (rc == GHX_OK && b1 != b0) {  // runner-up again, then the leader again; keep the better second timing
    double m0 = 0.0, m1 = 0.0;
    use(cand[b1]);
    rc = time_it(steps, m1);
    use(cand[b0]);
    if (rc == GHX_OK) rc = time_it(steps, m0);
    if (rc == GHX_OK && m1 < m0) std::swap(b0, b1);
  }
  use(cand[b0]);
  double bms = 0.0;
  if (rc == GHX_OK) rc = time_it(steps, bms);
  // A large irrep's link-lane step is a few percent of its MD step, below the
  // spread of the timings above, while a nested chain runs it alone 2m+1
  // times per outer half step: the staple split is decided on gauge-only
  // launches of that kernel (two rounds, the better of each kept).
  if (rc == GHX_OK && ctx->big && ctx->link_lanes && can_split) {
    auto time_gauge = [&](bool split, double& out) -> int {
      ctx->link_split = split;
      const LaunchCtx l = ctx->lc();
      const int n = 8 * steps;
      for (int i = 0; i <= n; ++i) {
        if (i == 1) HIPCHK(hipEventRecord(e0, ctx->stream));  // the first launch stays out of the timing
        const int a = ctx->cur, b = ctx->cur ^ 1;
        ctx->bops.gauge_step(l, ctx->U[a], ctx->U[b], ctx->P, ctx->phi[ctx->cur], ctx->phi[ctx->cur], ctx->pi, 1.0, 1.0,
                             F_KICK | F_DRIFT | kInParts);
        ctx->cur = b;
      }
      HIPCHK(hipEventRecord(e1, ctx->stream));
      HIPCHK(hipEventSynchronize(e1));
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, e0, e1));
      out = ms;
      return GHX_OK;
    };
    double g[2] = {1e30, 1e30};
    for (int r = 0; r < 2 && rc == GHX_OK; ++r)
      for (int s : {0, 1}) {
        double t = 0.0;
        if (rc == GHX_OK) rc = time_gauge(s != 0, t);
        g[s] = std::min(g[s], t);
      }
    ctx->link_split = g[1] < g[0];
    cand[b0].split = ctx->link_split ? 1 : 0;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != GHX_OK) {  // the fields stay as the failed launch left them
    (void)hipFree(Ps);
    if (pis) (void)hipFree(pis);
    return rc;
  }
  const int best = cand[b0].first;
  ctx->cur = 0;
  HIPCHK(hipMemcpyAsync(ctx->U[0], ctx->Usave, nUb, hipMemcpyDeviceToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(ctx->P, Ps, nPb, hipMemcpyDeviceToDevice, ctx->stream));
  if (nphb) {
    HIPCHK(hipMemcpyAsync(ctx->phi[0], ctx->phisave, nphb, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->pi, pis, nphb, hipMemcpyDeviceToDevice, ctx->stream));
  }
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipFree(Ps));
  if (pis) HIPCHK(hipFree(pis));
  drop_graph(ctx);
  if (best_seg) *best_seg = best;
  if (best_ms) *best_ms = bms / steps;
  return GHX_OK;
}

int ghx_sync(ghx_ctx* ctx) {
  if (!ctx) return GHX_E_INVALID;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return GHX_OK;
}

int ghx_time_md_steps(ghx_ctx* ctx, int steps, double* ms_total) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (steps < 1 || !ms_total) return GHX_E_INVALID;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, ctx->stream));
  for (int i = 0; i < steps; ++i) launch_step(ctx, 1.0, 1.0, true);
  HIPCHK(hipEventRecord(e1, ctx->stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms_total = ms;
  return GHX_OK;
}

int ghx_rng_gauss(uint64_t seed, uint64_t key0, int64_t n, double* out) {
  if (n < 0 || (n > 0 && !out)) return GHX_E_INVALID;
  if (n == 0) return GHX_OK;
  ghx_ctx* ctx = nullptr;  // HIPCHK's error slot
  constexpr int64_t kChunk = int64_t(1) << 24;
  double* d = nullptr;
  const int64_t cap = n < kChunk ? n : kChunk;
  HIPCHK(hipMalloc(&d, size_t(cap) * sizeof(double)));
  int rc = GHX_OK;
  for (int64_t o = 0; o < n && rc == GHX_OK; o += cap) {
    const int64_t m = n - o < cap ? n - o : cap;
    hipLaunchKernelGGL(rng_gauss_kernel, dim3(unsigned((m + kRngBlock - 1) / kRngBlock)), dim3(kRngBlock), 0, nullptr,
                       (unsigned long long)seed,
                       (unsigned long long)(key0 + uint64_t(o)), (long long)m, d);
    if (hipGetLastError() != hipSuccess || hipMemcpy(out + o, d, size_t(m) * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      rc = GHX_E_HIP;
  }
  (void)hipFree(d);
  return rc;
}

int ghx_rep_query(int N, const char* spec, int* d, int* nality, int* real_rep, long* dimT, double* W, double* T) {
  if (N != 2 && N != 3) return GHX_E_INVALID;
  RepInfo r;
  try {
    r = build_rep(N, spec ? spec : "");
  } catch (const std::exception&) {
    return GHX_E_REP;
  }
  if (d) *d = r.d;
  if (nality) *nality = r.nality;
  if (real_rep) *real_rep = r.real ? 1 : 0;
  if (dimT) *dimT = r.dimT;
  if (W) std::copy(r.W.begin(), r.W.end(), W);
  if (T)
    for (size_t i = 0; i < r.T.size(); ++i) { T[2 * i] = r.T[i].real(); T[2 * i + 1] = r.T[i].imag(); }
  return GHX_OK;
}

// Channels of the multi-invariant potential (scalar_invariants.hpp:20-52): the
// distinct eigenvalues of M -> sum_a [T^a_R, [T^a_R, M]]; host only.
int ghx_casimir_channels(int N, const char* spec, double* C2, int* n) {
  if ((N != 2 && N != 3) || !n) return GHX_E_INVALID;
  RepInfo r;
  try {
    r = build_rep(N, spec ? spec : "");
  } catch (const std::exception&) {
    return GHX_E_REP;
  }
  if (r.d == 0) return GHX_E_INVALID;
  const std::vector<double> lam = casimir_channels(r);
  *n = int(lam.size());
  if (C2) std::copy(lam.begin(), lam.end(), C2);
  return GHX_OK;
}

// Plug in the multi-invariant potential V = -mu2 phi^dag phi + sum_c f_c V_c
// (MultiInvariantPotential, scalar_invariants.hpp:182-208, as GaugeHiggsHMC::potential):
// nf must equal the number of channels; f == NULL restores the quartic potential.
int ghx_set_multi_potential(ghx_ctx* ctx, const double* f, int nf, double mu2) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  drop_graph(ctx);
  HIPCHK(hipStreamSynchronize(ctx->stream));
  free_multi(ctx);
  if (!f) return GHX_OK;
  if (!has_phi(ctx)) return fail(ctx, GHX_E_INVALID, "multi-invariant potential needs a scalar field");
  if (ctx->big && ctx->rep.d > kMultiMaxD)
    return fail(ctx, GHX_E_UNSUPPORTED, "multi-invariant potential is built for irreps of dimension <= " + std::to_string(kMultiMaxD) +
                                            " (this one has " + std::to_string(ctx->rep.d) + ")");
  const std::vector<double> lam = casimir_channels(ctx->rep);
  if (nf != int(lam.size()))
    return fail(ctx, GHX_E_INVALID, "need " + std::to_string(lam.size()) + " channel couplings, got " + std::to_string(nf));
  const std::vector<std::complex<double>> A = multi_superop(ctx->rep, lam, std::vector<double>(f, f + nf));
  const std::vector<double> Ad = hermitian_superop(ctx->rep.d, A);
  if (ctx->tsize == 8) {
    HIPCHK(hipMalloc(&ctx->dAq, Ad.size() * 8));
    HIPCHK(hipMemcpy(ctx->dAq, Ad.data(), Ad.size() * 8, hipMemcpyHostToDevice));
  } else {
    const std::vector<float> Af(Ad.begin(), Ad.end());
    HIPCHK(hipMalloc(&ctx->dAq, Af.size() * 4));
    HIPCHK(hipMemcpy(ctx->dAq, Af.data(), Af.size() * 4, hipMemcpyHostToDevice));
  }
  HIPCHK(hipMalloc(&ctx->dFon, ctx->nphi() * ctx->Vp * ctx->tsize));
  HIPCHK(hipMemset(ctx->dFon, 0, ctx->nphi() * ctx->Vp * ctx->tsize));
  HIPCHK(hipMalloc((void**)&ctx->dVon, size_t(ctx->Vp) * 8));
  HIPCHK(hipMemset(ctx->dVon, 0, size_t(ctx->Vp) * 8));
  ctx->mu2 = mu2;
  ctx->multi = true;
  return GHX_OK;
}

int64_t ghx_volume(const ghx_ctx* ctx) {
  if (!ctx) return 0;
  int64_t v = 0;
  for (const auto& rd : ctx->hreps) v += rd.vol;
  return v;
}


// ---- checkpoint / resume (roadmap item of the reference, README.md:126-127) ----
// Binary file, little-endian: magic "GHXCKPT1", then int32 {version = 3, D, N,
// n_replicas, precision, d, frozen, nmd, integrator, reunit, n_outer, m_inner,
// inner integrator} (version 2 files end the header after m_inner and load as
// a leapfrog inner scale),
// int32 extents[n_replicas*D], per replica {u64 seed, traj_count,
// accept_count; f64 last_dH, beta, kappa, lambda, tau}, u64 {swap seed, swap
// counter}, int32 n_pairs, u64 attempts[n_pairs], u64 accepts[n_pairs], then
// per replica the links (reference AoS [(s*D+mu)*N*N + i*N + j] complex, FP64)
// and phi ([s*d+k] complex, FP64). Momenta are not stored: every trajectory
// refreshes them from its counter. Loading requires a context with the same
// geometry, irrep, frozen flag, integrator settings and replica-exchange pairs.
namespace {
constexpr char kCkptMagic[8] = {'G', 'H', 'X', 'C', 'K', 'P', 'T', '1'};
constexpr int kCkptVersion = 3, kCkptHdr = 13;
struct CkptRep {
  uint64_t seed, traj, acc;
  double last_dH, beta, kappa, lambda, tau;
};
}  // namespace

int ghx_save_checkpoint(ghx_ctx* ctx, const char* path) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (!path) return GHX_E_INVALID;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(ctx, GHX_E_INVALID, std::string("cannot open ") + path);
  bool ok = std::fwrite(kCkptMagic, 1, 8, f) == 8;
  const int32_t hdr[kCkptHdr] = {kCkptVersion, ctx->D, ctx->N, ctx->nrep, ctx->prec, ctx->rep.d, ctx->frozen ? 1 : 0,
                                 ctx->nmd, ctx->integ, ctx->reunit ? 1 : 0, ctx->n_outer, ctx->m_inner, ctx->inner_integ};
  ok = ok && std::fwrite(hdr, sizeof hdr, 1, f) == 1;
  for (int r = 0; r < ctx->nrep && ok; ++r) {
    int32_t L[4] = {0, 0, 0, 0};
    for (int m = 0; m < ctx->D; ++m) L[m] = ctx->hreps[r].L[m];
    ok = std::fwrite(L, sizeof(int32_t), ctx->D, f) == size_t(ctx->D);
  }
  for (int r = 0; r < ctx->nrep && ok; ++r) {
    McState m{};
    HIPCHK(hipMemcpy(&m, ctx->mc + r, sizeof m, hipMemcpyDeviceToHost));
    const RepDesc& rd = ctx->hreps[r];
    const CkptRep c{rd.seed, m.traj_count, m.accept_count, m.last_dH, rd.beta, rd.kappa, rd.lambda, rd.tau};
    ok = std::fwrite(&c, sizeof c, 1, f) == 1;
  }
  SwapState sw{0, 0};
  if (ctx->dswap) HIPCHK(hipMemcpy(&sw, ctx->dswap, sizeof sw, hipMemcpyDeviceToHost));
  ok = ok && std::fwrite(&sw, sizeof sw, 1, f) == 1;
  // swap statistics per pair, after the pair count they belong to
  const int32_t np = ctx->npairs;
  std::vector<unsigned long long> att(size_t(std::max(np, 0)), 0), acc(att.size(), 0);
  if (np > 0) {
    HIPCHK(hipMemcpy(att.data(), ctx->dattempts, sizeof(unsigned long long) * np, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(acc.data(), ctx->daccepts, sizeof(unsigned long long) * np, hipMemcpyDeviceToHost));
  }
  ok = ok && std::fwrite(&np, sizeof np, 1, f) == 1;
  if (np > 0) ok = ok && std::fwrite(att.data(), sizeof(unsigned long long), np, f) == size_t(np) &&
                   std::fwrite(acc.data(), sizeof(unsigned long long), np, f) == size_t(np);
  for (int r = 0; r < ctx->nrep && ok; ++r) {
    const size_t nl = size_t(ctx->hreps[r].vol) * ctx->D * ctx->N * ctx->N * 2;
    std::vector<double> U(nl);
    if ((rc = get_links(ctx, r, U.data()))) { std::fclose(f); return rc; }
    ok = std::fwrite(U.data(), 8, nl, f) == nl;
    if (ok && has_phi(ctx)) {
      const size_t np = size_t(ctx->hreps[r].vol) * ctx->nphi();
      std::vector<double> ph(np);
      if ((rc = get_planes(ctx, ctx->phi[ctx->cur], int(ctx->nphi()), r, ph.data()))) { std::fclose(f); return rc; }
      ok = std::fwrite(ph.data(), 8, np, f) == np;
    }
  }
  ok = (std::fclose(f) == 0) && ok;
  return ok ? GHX_OK : fail(ctx, GHX_E_INVALID, std::string("write failed: ") + path);
}

int ghx_load_checkpoint(ghx_ctx* ctx, const char* path) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (!path) return GHX_E_INVALID;
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(ctx, GHX_E_INVALID, std::string("cannot open ") + path);
  auto bad = [&](const std::string& why) { std::fclose(f); return fail(ctx, GHX_E_INVALID, "checkpoint " + std::string(path) + ": " + why); };
  char magic[8];
  int32_t hdr[kCkptHdr];
  if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, kCkptMagic, 8) != 0) return bad("not a ghx checkpoint");
  if (std::fread(hdr, sizeof(int32_t), kCkptHdr - 1, f) != size_t(kCkptHdr - 1) || (hdr[0] != 2 && hdr[0] != kCkptVersion))
    return bad("unknown version");
  hdr[kCkptHdr - 1] = GHX_LEAPFROG;
  if (hdr[0] == kCkptVersion && std::fread(hdr + kCkptHdr - 1, sizeof(int32_t), 1, f) != 1) return bad("truncated");
  if (hdr[1] != ctx->D || hdr[2] != ctx->N || hdr[3] != ctx->nrep || hdr[5] != ctx->rep.d)
    return bad("lattice dimension, colours, replica count or scalar dimension differ from this context");
  if (hdr[6] != (ctx->frozen ? 1 : 0)) return bad("frozen-length flag differs from this context (set it before loading)");
  if (hdr[7] != ctx->nmd || hdr[8] != ctx->integ || hdr[9] != (ctx->reunit ? 1 : 0) || hdr[10] != ctx->n_outer ||
      hdr[11] != ctx->m_inner || (ctx->n_outer > 0 && hdr[12] != ctx->inner_integ))
    return bad("integrator settings (nmd, integrator, reunitarization, nested scales) differ from this context");
  for (int r = 0; r < ctx->nrep; ++r) {
    int32_t L[4] = {0, 0, 0, 0};
    if (std::fread(L, sizeof(int32_t), ctx->D, f) != size_t(ctx->D)) return bad("truncated");
    for (int m = 0; m < ctx->D; ++m)
      if (L[m] != ctx->hreps[r].L[m]) return bad("extents differ from this context");
  }
  std::vector<CkptRep> reps(ctx->nrep);
  if (std::fread(reps.data(), sizeof(CkptRep), ctx->nrep, f) != size_t(ctx->nrep)) return bad("truncated");
  SwapState sw{0, 0};
  if (std::fread(&sw, sizeof sw, 1, f) != 1) return bad("truncated");
  int32_t np = 0;
  if (std::fread(&np, sizeof np, 1, f) != 1) return bad("truncated");
  if (np != ctx->npairs) return bad("replica-exchange pairs differ from this context (set the ladders before loading)");
  std::vector<unsigned long long> att(size_t(std::max(np, 0))), acc(att.size());
  if (np > 0 && (std::fread(att.data(), sizeof(unsigned long long), np, f) != size_t(np) ||
                 std::fread(acc.data(), sizeof(unsigned long long), np, f) != size_t(np)))
    return bad("truncated");
  std::vector<std::vector<double>> Us(ctx->nrep), phs(ctx->nrep);
  for (int r = 0; r < ctx->nrep; ++r) {
    Us[r].resize(size_t(ctx->hreps[r].vol) * ctx->D * ctx->N * ctx->N * 2);
    if (std::fread(Us[r].data(), 8, Us[r].size(), f) != Us[r].size()) return bad("truncated");
    if (has_phi(ctx)) {
      phs[r].resize(size_t(ctx->hreps[r].vol) * ctx->nphi());
      if (std::fread(phs[r].data(), 8, phs[r].size(), f) != phs[r].size()) return bad("truncated");
    }
  }
  std::fclose(f);
  if (ctx->cur != 0) launch_settle(ctx, false);
  for (int r = 0; r < ctx->nrep; ++r) {
    const CkptRep& c = reps[r];
    if ((rc = ghx_set_couplings(ctx, r, c.beta, c.kappa, c.lambda, c.tau))) return rc;
    if ((rc = ghx_set_seed(ctx, r, c.seed))) return rc;
    McState m{};
    m.traj_count = c.traj;
    m.accept_count = c.acc;
    m.last_dH = c.last_dH;
    HIPCHK(hipMemcpy(ctx->mc + r, &m, sizeof m, hipMemcpyHostToDevice));
    if ((rc = set_links(ctx, r, Us[r].data()))) return rc;
    if (has_phi(ctx) && (rc = set_planes(ctx, ctx->phi[ctx->cur], int(ctx->nphi()), r, phs[r].data()))) return rc;
  }
  if (ctx->dswap) HIPCHK(hipMemcpy(ctx->dswap, &sw, sizeof sw, hipMemcpyHostToDevice));
  if (np > 0) {
    HIPCHK(hipMemcpy(ctx->dattempts, att.data(), sizeof(unsigned long long) * np, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->daccepts, acc.data(), sizeof(unsigned long long) * np, hipMemcpyHostToDevice));
  }
  drop_graph(ctx);
  return GHX_OK;
}

}  // extern "C"
