"""Replica sets on one GPU: many lattices (uneven extents, own couplings and
seeds) advanced by the same launches, with replica exchange along beta or
kappa ladders decided on the device (replica_tempering.hpp:78-218).

`ReplicaSet` mirrors ReplicaTempering's interface (step, pair_acceptance,
hmc_acceptance, enabled); `kappa_scan` builds the 32-chain SU(2)+fundamental
configuration with kappa evenly spaced over [0.20, 0.32] and L in {6,8,10,12}.
"""
import ctypes as C

import numpy as np

from .hmc import FP32MD, HOT, OMELYAN, Lattices

SWAP_SEED = 0x5A7AB1EC     # replica_tempering.hpp:103
SEED_STRIDE = 1000003      # replica_tempering.hpp:102


class ReplicaSet:
    def __init__(self, D, N, extents, rep, betas, kappas, lambdas, seed0=1, nmd=20, tau=1.0, integ=OMELYAN,
                 precision=FP32MD, phi_sigma=0.3, ladders=None, swap_seed=SWAP_SEED):
        self.lat = Lattices(D, N, extents, rep=rep, precision=precision)
        n = self.lat.nrep
        self.n = n
        self.betas, self.kappas, self.lambdas = list(betas), list(kappas), list(lambdas)
        self.lat.set_integrator(nmd, integ, True)
        for r in range(n):
            self.lat.set_couplings(r, self.betas[r], self.kappas[r], self.lambdas[r], tau)
            self.lat.set_seed(r, seed0 + r * SEED_STRIDE)
            self.lat.init_fields(r, HOT, 0.8, 1, 12345, phi_sigma)
        self.ladders = ladders or []
        if self.ladders:
            lengths = (C.c_int * len(self.ladders))(*[len(l) for l in self.ladders])
            flat = [r for l in self.ladders for r in l]
            members = (C.c_int * len(flat))(*flat)
            self.lat._check(self.lat._lib.ghx_set_ladders(self.lat.h, len(self.ladders), lengths, members, swap_seed))
        self._enabled = bool(self.ladders)

    @property
    def enabled(self):
        return self._enabled

    @enabled.setter
    def enabled(self, v):
        self._enabled = bool(v)
        self.lat._check(self.lat._lib.ghx_set_tempering(self.lat.h, int(self._enabled)))

    def step(self, n=1, use_graph=True):
        """n cycles of (one trajectory of every replica, then the swap passes)."""
        self.lat.trajectories(n, use_graph=use_graph)

    def swap_stats(self):
        npairs = C.c_int()
        self.lat._check(self.lat._lib.ghx_swap_stats(self.lat.h, C.byref(npairs), None, None, None, None))
        k = npairs.value
        if k == 0:
            return [], np.zeros(0), np.zeros(0), np.zeros(0)
        pairs = (C.c_int * (2 * k))()
        att = (C.c_uint64 * k)()
        acc = (C.c_uint64 * k)()
        flags = (C.c_int * k)()
        self.lat._check(self.lat._lib.ghx_swap_stats(self.lat.h, C.byref(npairs), pairs, att, acc, flags))
        pl = [(pairs[2 * i], pairs[2 * i + 1]) for i in range(k)]
        return pl, np.array(list(att)), np.array(list(acc)), np.array(list(flags))

    def pair_acceptance(self):
        pl, att, acc, _ = self.swap_stats()
        return {p: (a / t if t else 0.0) for p, t, a in zip(pl, att, acc)}

    def hmc_acceptance(self, k):
        t, a, _ = self.lat.counters(k)
        return a / t if t else 0.0

    def observables(self):
        return self.lat.observables_all()


def kappa_scan(seed0=1, nmd=20, precision=FP32MD, n_chains=32, Ls=(6, 8, 10, 12), kmin=0.20, kmax=0.32, beta=2.3, lam=0.5):
    """32 SU(2)+fundamental chains: chain i has kappa = kmin + (kmax-kmin) i/(n-1) and
    L = Ls[i % 4]; swaps run between kappa-neighbours of equal L."""
    kappas = [kmin + (kmax - kmin) * i / (n_chains - 1) for i in range(n_chains)]
    Lc = [Ls[i % len(Ls)] for i in range(n_chains)]
    ladders = [[i for i in range(n_chains) if Lc[i] == L] for L in Ls]
    return ReplicaSet(4, 2, [[L] * 4 for L in Lc], "fund", [beta] * n_chains, kappas, [lam] * n_chains, seed0=seed0, nmd=nmd,
                      precision=precision, ladders=ladders)


def measure_scan(rs, n_therm=200, n_meas=1000, use_graph=True):
    """Run a replica set: n_therm cycles, then n_meas cycles recording every
    chain's L_link (scalar_higgs.hpp:130-144). Returns per-chain kappa, L,
    <L_link>, chi_link = V Var(L_link) (observables.hpp:104-145 Stats), the HMC
    and swap acceptances, and the time per cycle (one trajectory of every chain
    plus the swap passes, graph-replayed)."""
    import time
    rs.step(n_therm, use_graph=use_graph)
    rs.lat.sync()
    series = np.zeros((n_meas, rs.n))
    t = 0.0
    for k in range(n_meas):
        t0 = time.perf_counter()
        rs.step(1, use_graph=use_graph)
        rs.lat.sync()
        t += time.perf_counter() - t0
        series[k] = rs.observables()[:, 3]
    vols = np.array(rs.lat.vols, dtype=float)
    return {
        "kappa": np.array(rs.kappas), "L": np.array([e[0] for e in rs.lat.extents]), "Llink": series.mean(axis=0),
        "chi_link": vols * series.var(axis=0), "hmc_acceptance": np.array([rs.hmc_acceptance(k) for k in range(rs.n)]),
        "swap_acceptance": rs.pair_acceptance(), "s_per_cycle": t / n_meas, "series": series,
    }


def chi_peaks(m):
    """kappa of the largest chi_link for each L, with a parabolic refinement
    through the maximum and its two kappa neighbours."""
    out = {}
    for L in sorted(set(m["L"].tolist())):
        sel = np.where(m["L"] == L)[0]
        k, c = m["kappa"][sel], m["chi_link"][sel]
        i = int(np.argmax(c))
        kp = float(k[i])
        if 0 < i < len(k) - 1:
            den = c[i - 1] - 2 * c[i] + c[i + 1]
            if den < 0:
                kp = float(k[i] + 0.5 * (k[i + 1] - k[i]) * (c[i - 1] - c[i + 1]) / den)
        out[int(L)] = {"kappa_peak": kp, "chi_max": float(c[i])}
    return out
