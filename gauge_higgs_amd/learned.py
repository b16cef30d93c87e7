"""Exact-force HMC against multi-timescale HMC with the learned surrogate force.

Workload: 4D SU(3) + complex scalar in the Young-rows {6,1} irrep (d = 48),
beta = 5.7, kappa = 0.05, lambda = 0.5, tau = 1, FP32 MD with the FP64
Hamiltonian and Metropolis test. The exact chain is the reference's HMC
(gauge_higgs_hmc.hpp:149-165, Omelyan, nmd = 16). The learned chain uses the
Sexton-Weingarten split (theory_notes.md:129-134, u1.hpp:278-320 for U(1)):
the expensive hopping forces (matter link force and the D(U) phi hops of the
scalar force, 12 tensor applies per site in the reference formulation) on
n_outer coarse steps, the gauge force, the on-site potential force and the
MLP surrogate S of the matter link force on m_inner fine steps each, with S
subtracted again on the coarse scale. Whatever S is, the integrated force is
the exact one, so the chain samples exactly (reversible, area preserving,
Metropolis on the exact FP64 H); S only moves error from the coarse scale to
the fine one.

The surrogate is trained on (configuration, exact matter link force) pairs
drawn from the exact chain, by the MFMA forward/backward kernels on the GPU.
Costs are GPU seconds per accepted trajectory (wall time of the trajectory
launches with the device synchronised; all work runs on one stream).
"""
import time

import numpy as np

from . import hmc as _h


def _chain(g, n, record=True, use_graph=True):
    dhs, acc, t = [], 0, 0.0
    obs = []
    for _ in range(n):
        a0 = g.accept_count
        g.lat.sync()
        t0 = time.perf_counter()
        g.lat.trajectories(1, use_graph=use_graph)
        g.lat.sync()
        t += time.perf_counter() - t0
        acc += g.accept_count - a0
        dhs.append(g.last_dH)
        if record:
            o = g.observables()
            obs.append((o["plaq"], o["Llink"]))
    dhs = np.array(dhs)
    a = acc / n
    out = {"trajectories": n, "gpu_s_per_traj": t / n, "acceptance": a, "acceptance_err": float(np.sqrt(max(a * (1 - a), 1.0 / n) / n)),
           "exp_minus_dH": float(np.mean(np.exp(-np.clip(dhs, -50, 50)))),
           "exp_minus_dH_err": float(np.std(np.exp(-np.clip(dhs, -50, 50))) / np.sqrt(n)),
           "mean_abs_dH": float(np.mean(np.abs(dhs)))}
    if record:
        ob = np.array(obs)
        out.update(plaq=float(ob[:, 0].mean()), plaq_err=float(ob[:, 0].std() / np.sqrt(n)),
                   Llink=float(ob[:, 1].mean()), Llink_err=float(ob[:, 1].std() / np.sqrt(n)))
    out["gpu_s_per_accepted"] = out["gpu_s_per_traj"] / max(out["acceptance"], 1e-9)
    out["gpu_s_per_accepted_err"] = out["gpu_s_per_accepted"] * out["acceptance_err"] / max(out["acceptance"], 1e-9)
    return out


def _ratio(a, b):
    """a / b for two chain summaries' GPU-s per accepted trajectory, with the
    error from both (independent chains, relative errors in quadrature)."""
    x, y = a["gpu_s_per_accepted"], b["gpu_s_per_accepted"]
    r = x / y
    return {"ratio": r, "err": r * float(np.hypot(a["gpu_s_per_accepted_err"] / x, b["gpu_s_per_accepted_err"] / y))}


def basis_bound(B, F, subsets=((0,), (0, 1), (0, 1, 2), (0, 1, 2, 3))):
    """Fraction of |F|^2 left unexplained when every link's force F is replaced
    by its least-squares projection onto the span of its own basis vectors
    B[link, k, :] (k in the subset): the residual of a network that picked the
    best coefficients for every link, a lower bound for any surrogate over these
    bases. Also the residual of one global coefficient per basis."""
    F = np.asarray(F, dtype=np.float64)
    out = {}
    fn = float(np.sum(F ** 2))
    for sub in subsets:
        Bs = np.asarray(B, dtype=np.float64)[:, list(sub), :]          # [links, k, a]
        A = np.transpose(Bs, (0, 2, 1))                                # [links, a, k]
        G = np.einsum("lak,lam->lkm", A, A) + 1e-30 * np.eye(len(sub))
        c = np.linalg.solve(G, np.einsum("lak,la->lk", A, F)[..., None])[..., 0]
        r_link = float(np.sum((F - np.einsum("lak,lk->la", A, c)) ** 2)) / fn
        Ag = A.reshape(-1, len(sub))
        cg, *_ = np.linalg.lstsq(Ag, F.reshape(-1), rcond=None)
        r_glob = float(np.sum((F.reshape(-1) - Ag @ cg) ** 2)) / fn
        out["B" + "".join(str(k) for k in sub)] = {"per_link_optimum": r_link, "global_coefficients": r_glob}
    return out


def split_key(split):
    """Result key of a split (n_outer, m_inner[, inner integrator])."""
    inner = split[2] if len(split) > 2 else _h.LEAPFROG
    return "omelyan_outer%d_%sinner%d" % (split[0], "omelyan_" if inner == _h.OMELYAN else "", split[1])


RESIDUAL_GATE = 0.9  # use S only if it explains > 10 % of |F|^2 on a held-out configuration


def compare(L=12, rep="6,1", beta=5.7, kappa=0.05, lam=0.5, nmd=16, exact_nmds=(16, 20, 24, 28, 32, 40), n_therm=40, n_meas=200,
            splits=((5, 4, _h.OMELYAN), (5, 10)), train_configs=8, train_iters=200, lr=2e-3, seed=1, precision=_h.FP32MD,
            therm_split=(12, 6), warm=(0.4, 0.27), use_graph=True):
    """Cost per accepted trajectory of exact-force HMC and of the multi-timescale
    chain with and without the learned surrogate, from one thermalised
    configuration.

    Start: links exp(i sigma g.T) with sigma = warm[0] and Gaussian phi with
    sigma = warm[1] (|phi|^2 near its equilibrium value), thermalised by the
    split integrator (any exact chain samples the same distribution; the stiff
    on-site potential of a d = 48 scalar stalls coarse single-scale chains far
    from equilibrium). The exact chains run Omelyan at each nmd in exact_nmds
    (the reference setting is nmd = 16); the surrogate is trained on
    (configuration, exact matter link force) pairs from the thermalised chain;
    each split (n_outer, m_inner[, inner integrator], leapfrog inner by default)
    runs with the Omelyan outer scale, with the surrogate forced on
    ("surrogate_on") and without it ("split_no_surrogate"). The learned chain is
    the first split; the default pairs 5 outer steps of 4 inner Omelyan steps
    (9 inner launches per half step) with the same outer scale over 10 inner
    leapfrog steps (11 launches), the cheapest leapfrog-inner split measured.

    The learned chain keeps the exact force as fallback and arbiter: it uses the
    surrogate only when its held-out residual |S - F|^2 / |F|^2 is below
    RESIDUAL_GATE, otherwise the coarse scale carries the exact hopping force
    alone; "learned" reports whichever of the two measured chains that is."""
    g = _h.GaugeHiggsHMC(L, rep, seed=seed, D=4, N=3, precision=precision)
    g.configure(beta=beta, kappa=kappa, lam=lam, nmd=nmd, integ=_h.OMELYAN)
    g.hot(warm[0], 12345, warm[1])
    g.lat.autotune(steps=10)  # launch shape for this lattice (fields restored); every chain below uses it
    t0 = time.perf_counter()
    g.lat.set_nested(*therm_split)
    therm = _chain(g, n_therm, record=True, use_graph=use_graph)
    therm_s = time.perf_counter() - t0

    # training pairs: (configuration, exact matter link force) along the chain
    g.lat.surrogate_init(seed=7)
    losses = []
    t0 = time.perf_counter()
    for _ in range(train_configs):
        g.lat.trajectories(1, use_graph=use_graph)
        l0, l1 = g.lat.surrogate_train(iters=train_iters, lr=lr)
        losses.append((l0, l1))
    g.lat.sync()
    train_s = time.perf_counter() - t0
    g.lat.trajectories(1, use_graph=use_graph)
    # held-out residual: fraction of |F|^2 the surrogate leaves unexplained on a new configuration
    _, Bh, S, F = g.lat.surrogate_snapshot()
    resid = float(np.sum((S - F) ** 2) / np.sum(F ** 2))
    bound = basis_bound(Bh.reshape(F.shape[0], 4, -1), F)
    U0, phi0 = g.get("U"), g.get("phi")

    def restart():
        g.set("U", U0)
        g.set("phi", phi0)

    out = {"workload": f"su3_{rep.replace(',', '')}_L{L}", "L": L, "sites": L ** 4, "beta": beta, "kappa": kappa, "lambda": lam,
           "start": f"hot sigma_U={warm[0]}, Gaussian phi sigma={warm[1]}; {n_therm} thermalisation trajectories "
                    f"(split {therm_split[0]}x{therm_split[1]}, acceptance {therm['acceptance']:.2f}, plaq {therm['plaq']:.4f}, {therm_s:.1f} s)",
           "launch": {"link_lanes": g.lat.link_lanes(), "block_order": g.lat.block_order(), "store_cached": g.lat.store_policy(), "link_split": g.lat.link_split()},
           "exact": {}, "surrogate_on": {}, "split_no_surrogate": {}}
    g.lat.set_nested(0, 1)
    g.lat.surrogate_enable(False)
    for n in exact_nmds:
        restart()
        g.configure(nmd=n)
        out["exact"][f"omelyan_nmd{n}"] = _chain(g, n_meas, use_graph=use_graph)
    g.configure(nmd=nmd)
    for sp in splits:
        key = split_key(sp)
        for name, surr in (("surrogate_on", True), ("split_no_surrogate", False)):
            restart()
            g.lat.set_nested(*sp)
            g.lat.surrogate_enable(surr)
            out[name][key] = _chain(g, n_meas, use_graph=use_graph)
    g.lat.surrogate_enable(False)
    g.lat.set_nested(0, 1)
    g.lat.close()
    out["training"] = {"configs": train_configs, "pairs": train_configs * 4 * L ** 4, "adam_steps": train_configs * train_iters,
                       "seconds": train_s, "loss_first": losses[0][0], "loss_last": losses[-1][1],
                       "heldout_residual_fraction": resid, "basis_bound": bound}
    use_s = resid < RESIDUAL_GATE
    key = split_key(splits[0])
    out["learned"] = {"surrogate_used": use_s, "gate": f"held-out residual {resid:.3f} vs {RESIDUAL_GATE}",
                      "chain": ("surrogate_on" if use_s else "split_no_surrogate") + "/" + key}
    learned = out["surrogate_on" if use_s else "split_no_surrogate"][key]
    out["learned"]["gpu_s_per_accepted"] = learned["gpu_s_per_accepted"]
    out["learned"]["gpu_s_per_accepted_err"] = learned["gpu_s_per_accepted_err"]
    # every exact chain against the learned one; no selection of the best chain
    out["exact_over_learned"] = {k: _ratio(v, learned) for k, v in out["exact"].items()}
    out["surrogate_gain"] = _ratio(out["split_no_surrogate"][key], out["surrogate_on"][key])
    return out
