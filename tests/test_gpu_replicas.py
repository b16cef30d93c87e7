"""Batched replicas on one GPU (uneven extents in one launch per MD step) and
replica exchange along kappa / beta ladders, against independent oracle
chains plus a host model of the reference swap rule
(replica_tempering.hpp:157-194)."""
import numpy as np
import pytest

import gauge_higgs_amd as gha
import oracle_lib as O

pytestmark = pytest.mark.gpu

SWAP_SEED = 0x5A7AB1EC


def test_uneven_replicas_match_independent_chains():
    ext = [[4, 4, 4, 4], [4, 4, 4, 6], [6, 4, 4, 4], [5, 3, 4, 4]]
    kap = [0.20, 0.25, 0.28, 0.31]
    bet = [2.3, 2.2, 2.4, 2.3]
    rs = gha.ReplicaSet(4, 2, ext, "fund", bet, kap, [0.5] * 4, seed0=11, nmd=8, precision=gha.FP64)
    rs.enabled = False
    ors = []
    for r in range(4):
        o = O.Oracle(4, 2, ext[r], "fund", seed=11 + r * 1000003)
        o.couplings(beta=bet[r], kappa=kap[r], lam=0.5, nmd=8)
        o.hot(0.8)
        o.phi_gaussian(12345, 0.3)
        ors.append(o)
    for _ in range(2):
        acc, dH = rs.lat.trajectory_all()
        for r, o in enumerate(ors):
            a, d, Hi, _ = o.trajectory()
            assert abs(dH[r] - d) < 1e-9 and acc[r] == a
    for r, o in enumerate(ors):
        assert np.max(np.abs(rs.lat.get("U", r) - o.get("U"))) < 1e-10
        assert np.max(np.abs(rs.lat.get("phi", r) - o.get("phi"))) < 1e-10


def _host_swap_pass(ors, pairs, betas, kappas, N, D, counter):
    """Reference rule on oracle chains; returns flags and the updated counter."""
    flags = []
    for i, j in pairs:
        oi, oj = ors[i].observables(), ors[j].observables()
        V = ors[i].V
        nplaq = V * D * (D - 1) // 2
        q0i, q0j = oi["plaq"] * N * nplaq, oj["plaq"] * N * nplaq
        q1i, q1j = oi["Llink"] * V * D, oj["Llink"] * V * D
        delta = (betas[i] - betas[j]) / N * (q0i - q0j) + 2 * (kappas[i] - kappas[j]) * (q1i - q1j)
        counter += 1
        acc = delta <= 0 or O.uniform(SWAP_SEED, O.key(0x57A9, counter)) < np.exp(-delta)
        flags.append(int(acc))
        if acc:
            Ui, Uj, pi_, pj = ors[i].get("U"), ors[j].get("U"), ors[i].get("phi"), ors[j].get("phi")
            ors[i].set("U", Uj); ors[j].set("U", Ui); ors[i].set("phi", pj); ors[j].set("phi", pi_)
    return flags, counter


@pytest.mark.parametrize("ladder_kind", ["kappa", "beta"])
def test_replica_exchange_matches_host_model(ladder_kind):
    n = 5
    ext = [[4, 4, 4, 4]] * 4 + [[4, 4, 4, 6]]
    if ladder_kind == "kappa":
        bet, kap = [2.3] * n, [0.22, 0.25, 0.28, 0.31, 0.27]
    else:
        bet, kap = [2.0, 2.2, 2.4, 2.6, 2.3], [0.25] * n
    ladders = [[0, 1, 2, 3], [4]]
    rs = gha.ReplicaSet(4, 2, ext, "fund", bet, kap, [0.5] * n, seed0=3, nmd=8, precision=gha.FP64, ladders=ladders)
    ors = []
    for r in range(n):
        o = O.Oracle(4, 2, ext[r], "fund", seed=3 + r * 1000003)
        o.couplings(beta=bet[r], kappa=kap[r], lam=0.5, nmd=8)
        o.hot(0.8)
        o.phi_gaussian(12345, 0.3)
        ors.append(o)
    counter = 0
    even, odd = [(0, 1), (2, 3)], [(1, 2)]
    for cycle in range(4):
        rs.step(1, use_graph=(cycle % 2 == 1))
        for o in ors:
            o.trajectory()
        fe, counter = _host_swap_pass(ors, even, bet, kap, 2, 4, counter)
        fo, counter = _host_swap_pass(ors, odd, bet, kap, 2, 4, counter)
        pl, att, acc, flags = rs.swap_stats()
        assert pl == even + odd
        assert list(flags) == fe + fo, (cycle, flags, fe, fo)
        for r in range(n):
            assert np.max(np.abs(rs.lat.get("U", r) - ors[r].get("U"))) < 1e-9
            assert np.max(np.abs(rs.lat.get("phi", r) - ors[r].get("phi"))) < 1e-9
    assert list(att) == [4, 4, 4]


def test_kappa_scan_32_chains_runs_in_one_launch_per_step():
    rs = gha.kappa_scan(seed0=1, nmd=20)
    assert rs.n == 32 and sorted(set(len(l) for l in rs.ladders)) == [8]
    assert rs.lat.D == 4 and sum(rs.lat.vols) == 8 * (6 ** 4 + 8 ** 4 + 10 ** 4 + 12 ** 4)
    rs.step(30)
    obs = rs.observables()
    assert np.all(np.isfinite(obs))
    # L_link grows with kappa inside every ladder (8 kappas from 0.20 to 0.32)
    for lad in rs.ladders:
        ll = obs[lad, 3]
        assert ll[-1] > ll[0]
    acc = [rs.hmc_acceptance(k) for k in range(32)]
    assert np.mean(acc) >= 0.5  # includes the thermalisation from the hot start
    pa = rs.pair_acceptance()
    assert len(pa) == 4 * 7 and all(0.0 <= v <= 1.0 for v in pa.values())


def test_tempering_preserves_each_rung_marginal():
    """Canonical gate (test_replica_tempering_canonical.cpp): with swaps on, each
    kappa rung samples the same L_link distribution as an independent chain.
    Both rungs sit below the 4^4 transition (kappa ~ 0.26), where L_link is
    unimodal and the 20-block error estimate is reliable; at the transition the
    two-phase tunnelling time exceeds the block length."""
    kap = [0.21, 0.23]
    res = {}
    for enabled in (False, True):
        rs = gha.ReplicaSet(4, 2, [[4, 4, 4, 4]] * 2, "fund", [2.3] * 2, kap, [0.5] * 2, seed0=77, nmd=10,
                            precision=gha.FP64, ladders=[[0, 1]])
        rs.enabled = enabled
        rs.step(100)
        series = []
        for _ in range(600):
            rs.step(1)
            series.append(rs.observables()[:, 3])
        s = np.array(series)
        nb = 20
        bm = s[: len(s) // nb * nb].reshape(nb, -1, 2).mean(axis=1)
        res[enabled] = (s.mean(axis=0), bm.std(axis=0, ddof=1) / np.sqrt(nb))
        if enabled:
            pa = rs.pair_acceptance()[(0, 1)]
            assert 0.0 < pa <= 1.0
    (m0, e0), (m1, e1) = res[False], res[True]
    z = np.abs(m0 - m1) / np.sqrt(e0 ** 2 + e1 ** 2)
    assert np.all(z < 4.0), (m0, m1, e0, e1)
