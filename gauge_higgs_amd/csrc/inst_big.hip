// Large tensor irreps (matrix-free apply in LDS, one workgroup per site).
#include "bigrep.hpp"
#include "launch.hpp"
#include "symrep_dev.hpp"
namespace ghx {
template <typename T, int N, int D>
struct BigLaunch {
  // SU(3), first Young row 2..kSymPairMaxK: two links per wave (sym_link_pair_kernel)
  template <int K>
  static void pair_launch(const LaunchCtx& c, const BigRepTables& tb, const BigArgs<T>& a, T* hops, int both) {
    hipLaunchKernelGGL((sym_link_pair_kernel<T, D, K>), dim3(c.Vp, (D + 1) / 2), dim3(kSymWave),
                       2 * size_t(tb.sym.pair_total) * sizeof(cx<T>), c.stream, a, tb, hops, both);
  }
  // FP32, second Young row <= 1: the columns of Sym^k(U) in registers (sym_link_col_kernel)
  template <int K>
  static void col_launch(const LaunchCtx& c, const BigRepTables& tb, const BigArgs<T>& a, T* hops, int both) {
    if constexpr (sizeof(T) == 4) {
      if (tb.sym.k2 == 0)
        hipLaunchKernelGGL((sym_link_col_kernel<D, K, 0>), dim3(c.Vp, (D + 1) / 2), dim3(kSymWave), 0, c.stream, a, tb, hops, both);
      else
        hipLaunchKernelGGL((sym_link_col_kernel<D, K, 1>), dim3(c.Vp, (D + 1) / 2), dim3(kSymWave), 0, c.stream, a, tb, hops, both);
    }
  }
  static bool link_pair(const LaunchCtx& c, const BigRepTables& tb, const BigArgs<T>& a, T* hops, int both) {
    if constexpr (N != 3) {
      return false;
    } else {
      if (!tb.sym.pair) return false;
      if (sizeof(T) == 4 && tb.sym.pair == 2 && tb.sym.k2 <= 1) {
        switch (tb.sym.k1) {
          case 2: col_launch<2>(c, tb, a, hops, both); return true;
          case 3: col_launch<3>(c, tb, a, hops, both); return true;
          case 4: col_launch<4>(c, tb, a, hops, both); return true;
          case 5: col_launch<5>(c, tb, a, hops, both); return true;
          case 6: col_launch<6>(c, tb, a, hops, both); return true;
          default: return false;
        }
      }
      switch (tb.sym.k1) {
        case 2: pair_launch<2>(c, tb, a, hops, both); return true;
        case 3: pair_launch<3>(c, tb, a, hops, both); return true;
        case 4: pair_launch<4>(c, tb, a, hops, both); return true;
        case 5: pair_launch<5>(c, tb, a, hops, both); return true;
        case 6: pair_launch<6>(c, tb, a, hops, both); return true;
        default: return false;
      }
    }
  }
  static void matter(const LaunchCtx& c, const BigRepTables& tb, const void* U, const void* phi_in, void* phi_out, void* pi, void* G,
                     void* Fphi, double* q1, double ck, double cd, int flags, int mode) {
    if (c.frozen) flags |= F_FROZEN;
    BigArgs<T> a{c.reps, nullptr, c.blk_rep, c.Vp, (const T*)U, (const T*)phi_in, (T*)phi_out, (T*)pi, (T*)G, (T*)Fphi, q1, ck, cd, flags, mode};
    if (c.multi && mode != BIG_ACTION && !(flags & F_NO_ONSITE) && (mode == BIG_FORCE || (flags & F_KICK))) {
      multi(c, tb.d, phi_in, c.Fon, nullptr);
      a.Fon = (const T*)c.Fon;
    }
    if (tb.sym.on) {  // symmetric-power engine: lift every site's scalar, one wave per link, then one per site
      T* hops = static_cast<T*>(tb.sym.hops);
      if (!(flags & F_NO_HOP)) {
        if constexpr (sizeof(T) == 4)
          hipLaunchKernelGGL(sym_lift_mfma_kernel, dim3(c.Vp / (kLiftMfmaWaves * kLiftMfmaSites)), dim3(kSymWave * kLiftMfmaWaves), 0,
                             c.stream, c.Vp, (const float*)phi_in, tb.d, tb.sym);
        else
          hipLaunchKernelGGL(sym_lift_kernel<T>, dim3(c.Vp / kSymWave, (tb.sym.dimS + kLiftChunk - 1) / kLiftChunk), dim3(kSymWave), 0,
                             c.stream, c.Vp, (const T*)phi_in, tb.d, tb.sym);
        const int both = mode == BIG_ACTION ? 0 : 1;  // the action needs the forward hops only
        if (!link_pair(c, tb, a, hops, both))
          hipLaunchKernelGGL((sym_link_kernel<T, N, D>), dim3(c.Vp, D), dim3(kSymWave), size_t(tb.sym.lds_total) * sizeof(cx<T>),
                             c.stream, a, tb, hops, both);
      }
      hipLaunchKernelGGL((sym_finish_kernel<T, N, D>), dim3(c.Vp / kFinishSites), dim3(kSymWave, kFinishSites), 0, c.stream, a, tb,
                         (const T*)hops);
      return;
    }
    const size_t lds = size_t(tb.dimT) * sizeof(cx<T>);
    hipLaunchKernelGGL((big_matter_kernel<T, N, D>), dim3(c.Vp), dim3(kBigBlock), lds, c.stream, a, tb);
  }
  static void action(const LaunchCtx& c, const void* U, const void* P, const void* phi, const void* pi, const double* q1, int d) {
    ActArgs a{c.reps, c.red_rep, c.red_first, c.Vp, c.partials};
    if (c.multi) multi(c, d, phi, nullptr, c.Von);
    hipLaunchKernelGGL((big_action_kernel<T, N, D>), dim3(c.nred), dim3(kRedBlock), 0, c.stream, a, (const T*)U, (const T*)P,
                       (const T*)phi, (const T*)pi, q1, d, c.multi ? (const double*)c.Von : nullptr);
  }
  // Multi-invariant on-site force planes and/or per-site potential (d <= kMultiMaxD).
  // Sites per block: the largest of 64, 32, 16 whose LDS (d^2 + 4d reals per site) fits 64 KB.
  static void multi(const LaunchCtx& c, int d, const void* phi, void* Fon, double* Von) {
    const size_t per = size_t(d * d + 4 * d) * sizeof(T);
    const unsigned nsites = c.nblk * unsigned(kBlock);
    if (per * 64 <= 65536)
      hipLaunchKernelGGL((multi_onsite_lds_kernel<T, 64>), dim3(nsites / 64), dim3(64), per * 64, c.stream, c.reps, c.blk_rep, c.Vp,
                         d, (const T*)phi, (const T*)c.Aq, c.mu2, (T*)Fon, Von);
    else if (per * 32 <= 65536)
      hipLaunchKernelGGL((multi_onsite_lds_kernel<T, 32>), dim3(nsites / 32), dim3(32), per * 32, c.stream, c.reps, c.blk_rep, c.Vp,
                         d, (const T*)phi, (const T*)c.Aq, c.mu2, (T*)Fon, Von);
    else
      hipLaunchKernelGGL((multi_onsite_lds_kernel<T, 16>), dim3(nsites / 16), dim3(16), per * 16, c.stream, c.reps, c.blk_rep, c.Vp,
                         d, (const T*)phi, (const T*)c.Aq, c.mu2, (T*)Fon, Von);
  }
  static void phi_drift(const LaunchCtx& c, const void* phi_in, void* phi_out, const void* pi, int nplanes, double cd, int abs_eps) {
    hipLaunchKernelGGL(phi_drift_kernel<T>, dim3(c.nblk), dim3(kBlock), 0, c.stream, c.reps, c.blk_rep, c.Vp, (const T*)phi_in,
                       (T*)phi_out, (T*)const_cast<void*>(pi), nplanes, cd, abs_eps, c.frozen);
  }
  static OnsiteArgs<T> onsite_args(const LaunchCtx& c, const void* phi_in, void* phi_out, void* pi, int d, double ck, double cd,
                                   int flags) {
    if (c.frozen) flags |= F_FROZEN;
    const T* Fon = nullptr;
    if (c.multi) {
      if (!(flags & F_NO_ONSITE) && (flags & F_KICK)) multi(c, d, phi_in, c.Fon, nullptr);
      Fon = (const T*)c.Fon;
    }
    return OnsiteArgs<T>{c.reps, c.blk_rep, c.Vp, (const T*)phi_in, (T*)phi_out, (T*)pi, d, ck, cd, flags, Fon};
  }
  static void onsite(const LaunchCtx& c, const void* phi_in, void* phi_out, void* pi, int d, double ck, double cd, int flags) {
    const OnsiteArgs<T> o = onsite_args(c, phi_in, phi_out, pi, d, ck, cd, flags);
    hipLaunchKernelGGL(onsite_step_kernel<T>, dim3(c.Vp / 64), dim3(64 * kOnsiteGroups), 0, c.stream, o);
  }
  // The inner-scale step (F_NO_HOP): the on-site update of the scalar and the
  // gauge update of the links, in one launch where the links run
  // link-parallel (D = 4), else as the two launches it replaces.
  static void onsite_gauge(const LaunchCtx& c, const void* U_in, void* U_out, void* P, const void* phi_in, void* phi_out, void* pi,
                           int d, double ck, double cd, int flags) {
    if constexpr (kLinkSites * D == 64 * kOnsiteGroups) {
      if (c.fuse_inner && c.link_lanes && c.stream_seg == 0) {
        const OnsiteArgs<T> o = onsite_args(c, phi_in, phi_out, pi, d, ck, cd, flags);
        const int gf = c.frozen ? flags | F_FROZEN : flags;
        const StepArgs<T> a{c.reps, c.blk_rep, c.Vp, (const T*)U_in, (T*)U_out, (T*)P, (const T*)phi_in, (T*)phi_out, (T*)pi, nullptr,
                            nullptr, ck, cd, gf, (const T*)c.G, (const T*)c.S, c.surr_sign, (const T*)c.Fon, c.walk};
        const unsigned nlink = c.nblk * unsigned(kBlock / kLinkSites);
        hipLaunchKernelGGL((md_link_onsite_kernel<T, N, D>), dim3(nlink + c.Vp / 64), dim3(kLinkSites * D), 0, c.stream, a, nlink, o);
        return;
      }
    }
    onsite(c, phi_in, phi_out, pi, d, ck, cd, flags);
    Launch<T, N, D, RepExt>::md_step(c, U_in, U_out, P, phi_in, phi_out, pi, nullptr, nullptr, ck, cd, flags);
  }
  static void onsite_seq(const LaunchCtx& c, const void* phi_in, void* phi_out, void* pi, int d, const void* seq) {
    OnsiteSeq q = *static_cast<const OnsiteSeq*>(seq);
    if (c.frozen)
      for (int s = 0; s < q.n; ++s) q.flags[s] |= F_FROZEN;
    const OnsiteArgs<T> o{c.reps, c.blk_rep, c.Vp, (const T*)phi_in, (T*)phi_out, (T*)pi, d, 0.0, 0.0, 0, nullptr};
    hipLaunchKernelGGL(onsite_seq_kernel<T>, dim3(c.Vp / 64), dim3(64 * kOnsiteGroups), 0, c.stream, o, q);
  }
  // the gauge update of one inner sub-step (the external link force of RepExt
  // has no scalar components, so the scalar buffers are not touched)
  static void gauge_step(const LaunchCtx& c, const void* U_in, void* U_out, void* P, const void* phi_in, void* phi_out, void* pi,
                         double ck, double cd, int flags) {
    Launch<T, N, D, RepExt>::md_step(c, U_in, U_out, P, phi_in, phi_out, pi, nullptr, nullptr, ck, cd, flags);
  }
};
template <typename T, int N, int D>
size_t matter_static_lds() {
  hipFuncAttributes fa{};
  if (hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&big_matter_kernel<T, N, D>)) != hipSuccess) return size_t(-1);
  return fa.sharedSizeBytes;
}
template <typename T, int N, int D>
void fill(Ops* o, BigOps* b) {
  *o = Launch<T, N, D, RepExt>::ops();
  b->matter = &BigLaunch<T, N, D>::matter;
  b->action = &BigLaunch<T, N, D>::action;
  b->phi_drift = &BigLaunch<T, N, D>::phi_drift;
  b->onsite = &BigLaunch<T, N, D>::onsite;
  b->onsite_gauge = &BigLaunch<T, N, D>::onsite_gauge;
  b->matter_static_lds = &matter_static_lds<T, N, D>;
  b->onsite_seq = &BigLaunch<T, N, D>::onsite_seq;
  b->gauge_step = &BigLaunch<T, N, D>::gauge_step;
}
bool ops_big(int prec, int N, int D, Ops* o, BigOps* b) {
  if (N == 3 && D == 4) { prec ? fill<float, 3, 4>(o, b) : fill<double, 3, 4>(o, b); return true; }
  if (N == 3 && D == 3) { prec ? fill<float, 3, 3>(o, b) : fill<double, 3, 3>(o, b); return true; }
  if (N == 2 && D == 4) { prec ? fill<float, 2, 4>(o, b) : fill<double, 2, 4>(o, b); return true; }
  if (N == 2 && D == 3) { prec ? fill<float, 2, 3>(o, b) : fill<double, 2, 3>(o, b); return true; }
  return false;
}
}  // namespace ghx
