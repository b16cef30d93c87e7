// Templated launchers that fill an Ops table for one (T, N, D, rep) combination.
#pragma once
#include "kernels.hpp"
#include "ops.hpp"

namespace ghx {

template <class REP, typename P>
struct ParamFill {
  static typename REP::template Params<P> make(const LaunchCtx&) { return {}; }
};
template <typename R, int N, int NB, int DR, typename P>
struct ParamFill<RepTensor<R, N, NB, DR>, P> {
  using Rep = RepTensor<R, N, NB, DR>;
  static typename Rep::template Params<P> make(const LaunchCtx& c) {
    typename Rep::template Params<P> p;
    for (int l = 0; l < DR; ++l)
      for (int t = 0; t < Rep::DIMT; ++t) p.W[l][t] = P(c.W[l * Rep::DIMT + t]);
    return p;
  }
};

template <typename T, int N, int D, template <typename, int> class REPT>
struct Launch {
  using REP = REPT<T, N>;
  using REPD = REPT<double, N>;
  static void md_step(const LaunchCtx& c, const void* U_in, void* U_out, void* P, const void* phi_in, void* phi_out, void* pi,
                      void* F, void* Fphi, double ck, double cd, int flags) {
    if (c.frozen) flags |= F_FROZEN;
    if constexpr (REP::DR > 0) {
      if (c.multi && !(flags & F_NO_ONSITE) && (flags & (F_KICK | F_FORCE))) {
        multi_onsite(c, phi_in, c.Fon, nullptr);
        flags |= F_EXT_ONSITE;
      }
    }
    StepArgs<T> a{c.reps, c.blk_rep, c.Vp, (const T*)U_in, (T*)U_out, (T*)P, (const T*)phi_in, (T*)phi_out, (T*)pi, (T*)F, (T*)Fphi,
                  ck, cd, flags, (const T*)c.G, (const T*)c.S, c.surr_sign, (const T*)c.Fon, c.walk};
    // x-lines in lanes (link_update): SU(3) with two stored rows, and only where
    // the context found every wave to hold whole x-lines
    constexpr bool kXLable = N == 3 && link_rows<T, N>() == 2;
    const bool xl = kXLable && c.xlane && !c.link_lanes && c.walk.order < 3;
    if (c.stream_seg > 0) {
      const unsigned V3 = c.V3, Lt = c.Lt;
      unsigned ntile = (V3 + kBlock - 1) / kBlock;
      ntile = (ntile + 7) / 8 * 8;
      const unsigned nseg = (Lt + unsigned(c.stream_seg) - 1) / unsigned(c.stream_seg);
      if constexpr (kXLable)
        if (xl) {
          hipLaunchKernelGGL((md_stream_kernel<T, N, D, REP, true>), dim3(ntile * nseg), dim3(kBlock), 0, c.stream, a,
                             ParamFill<REP, T>::make(c), ntile, c.stream_seg);
          return;
        }
      // the store policy (cached or streamed stores) is a kernel template
      // constant; the x-lane kernels always stream
      if (c.store_cached)
        hipLaunchKernelGGL((md_stream_kernel<T, N, D, REP, false, true>), dim3(ntile * nseg), dim3(kBlock), 0, c.stream, a,
                           ParamFill<REP, T>::make(c), ntile, c.stream_seg);
      else
        hipLaunchKernelGGL((md_stream_kernel<T, N, D, REP>), dim3(ntile * nseg), dim3(kBlock), 0, c.stream, a, ParamFill<REP, T>::make(c),
                           ntile, c.stream_seg);
    } else if (c.link_lanes) {
      if constexpr (REP::DR == 0 && D >= 3) {
        if (c.link_split && !(flags & F_NO_GAUGE)) {  // the staples split over planes: (D - 1) times the waves
          const dim3 grid(c.nblk * (kBlock / kLinkSites)), block(kLinkSites * D * (D - 1));
          if (c.store_cached)
            hipLaunchKernelGGL((md_link3_kernel<T, N, D, REP, true>), grid, block, 0, c.stream, a, ParamFill<REP, T>::make(c));
          else
            hipLaunchKernelGGL((md_link3_kernel<T, N, D, REP>), grid, block, 0, c.stream, a, ParamFill<REP, T>::make(c));
          return;
        }
      }
      if (c.store_cached)
        hipLaunchKernelGGL((md_link_kernel<T, N, D, REP, true>), dim3(c.nblk * (kBlock / kLinkSites)), dim3(kLinkSites * D), 0, c.stream,
                           a, ParamFill<REP, T>::make(c));
      else
        hipLaunchKernelGGL((md_link_kernel<T, N, D, REP>), dim3(c.nblk * (kBlock / kLinkSites)), dim3(kLinkSites * D), 0, c.stream, a,
                           ParamFill<REP, T>::make(c));
    } else {
      if constexpr (kXLable)
        if (xl) {
          hipLaunchKernelGGL((md_step_kernel<T, N, D, REP, true>), dim3(c.nblk), dim3(kBlock), 0, c.stream, a, ParamFill<REP, T>::make(c));
          return;
        }
      if (c.store_cached)
        hipLaunchKernelGGL((md_step_kernel<T, N, D, REP, false, true>), dim3(c.nblk), dim3(kBlock), 0, c.stream, a,
                           ParamFill<REP, T>::make(c));
      else
        hipLaunchKernelGGL((md_step_kernel<T, N, D, REP>), dim3(c.nblk), dim3(kBlock), 0, c.stream, a, ParamFill<REP, T>::make(c));
    }
  }
  static void drift(const LaunchCtx& c, const void* U_in, void* U_out, const void* P, const void* phi_in, void* phi_out,
                    const void* pi, double cd, int flags) {
    if (c.frozen) flags |= F_FROZEN;
    StepArgs<T> a{c.reps, c.blk_rep, c.Vp, (const T*)U_in, (T*)U_out, (T*)P, (const T*)phi_in, (T*)phi_out, (T*)pi, nullptr, nullptr,
                  0.0, cd, flags, nullptr, nullptr, 0.0, nullptr};
    hipLaunchKernelGGL((drift_kernel<T, N, D, REP::DR>), dim3(c.nblk), dim3(kBlock), 0, c.stream, a);
  }
  // Multi-invariant on-site force planes and/or per-site potential from phi.
  static void multi_onsite(const LaunchCtx& c, const void* phi, void* Fon, double* Von) {
    if constexpr (REP::DR > 0)
      hipLaunchKernelGGL((multi_onsite_kernel<T, REP::DR>), dim3(c.nblk), dim3(kBlock), 0, c.stream, c.reps, c.blk_rep, c.Vp,
                         (const T*)phi, (const T*)c.Aq, c.mu2, (T*)Fon, Von);
  }
  static void action(const LaunchCtx& c, const void* U, const void* P, const void* phi, const void* pi) {
    ActArgs a{c.reps, c.red_rep, c.red_first, c.Vp, c.partials};
    if constexpr (REP::DR > 0) {
      if (c.multi) {
        multi_onsite(c, phi, nullptr, c.Von);
        a.Von = c.Von;
      }
    }
    hipLaunchKernelGGL((action_kernel<T, N, D, REPD>), dim3(c.nred), dim3(kRedBlock), 0, c.stream, a, (const T*)U, (const T*)P,
                       (const T*)phi, (const T*)pi, ParamFill<REPD, double>::make(c));
  }
  static void refresh(const LaunchCtx& c, void* P, void* pi, const McState* st, unsigned long long tagP, unsigned long long tagPi) {
    hipLaunchKernelGGL((refresh_kernel<T, N, D>), dim3(c.nblk), dim3(kBlock), 0, c.stream, c.reps, c.blk_rep, c.Vp, (T*)P, (T*)pi,
                       c.d, c.real_rep, st, tagP, tagPi);
  }
  static void frozen_site(const LaunchCtx& c, void* phi, void* pi, int mode) {
    hipLaunchKernelGGL(frozen_site_kernel<T>, dim3(c.nblk), dim3(kBlock), 0, c.stream, c.reps, c.blk_rep, c.Vp, (T*)phi, (T*)pi, c.d,
                       mode);
  }
  static void init_links(const LaunchCtx& c, void* U, int replica, int mode, double sigma) {
    hipLaunchKernelGGL((init_links_kernel<T, N, D>), dim3(c.nblk), dim3(kBlock), 0, c.stream, c.reps, c.blk_rep, c.Vp, (T*)U, replica,
                       mode, sigma);
  }
  static void reunit(const LaunchCtx& c, const void* U_in, void* U_out) {
    hipLaunchKernelGGL((reunit_kernel<T, N, D>), dim3(c.nblk), dim3(kBlock), 0, c.stream, c.reps, c.blk_rep, c.Vp, (const T*)U_in,
                       (T*)U_out);
  }
  static void unitarity(const LaunchCtx& c, const void* U) {
    ActArgs a{c.reps, c.red_rep, c.red_first, c.Vp, c.partials};
    hipLaunchKernelGGL((unitarity_kernel<T, N, D>), dim3(c.nred), dim3(kRedBlock), 0, c.stream, a, (const T*)U);
  }
  static void polyakov(const LaunchCtx& c, const void* U) {
    ActArgs a{c.reps, c.red_rep, c.red_first, c.Vp, c.partials};
    hipLaunchKernelGGL((polyakov_kernel<T, N, D>), dim3(c.nred), dim3(kRedBlock), 0, c.stream, a, (const T*)U);
  }
  static Ops ops() { return Ops{&md_step, &drift, &action, &refresh, &init_links, &reunit, &unitarity, &polyakov, &frozen_site}; }
};

// Dispatch helper over precision and D in {3, 4}.
template <int N, template <typename, int> class REPT>
bool pick(int prec, int D, Ops* o) {
  if (D == 4) { *o = prec ? Launch<float, N, 4, REPT>::ops() : Launch<double, N, 4, REPT>::ops(); return true; }
  if (D == 3) { *o = prec ? Launch<float, N, 3, REPT>::ops() : Launch<double, N, 3, REPT>::ops(); return true; }
  return false;
}

}  // namespace ghx
