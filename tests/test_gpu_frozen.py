"""Frozen-length scalar |phi_x| = 1 on the GPU (GaugeHiggsHMC::frozen_phi,
gauge_higgs_hmc.hpp:61-103, 133, 144, 172) against the reference's own dumps:
normalisation, tangent-projected momenta, kick + projection, geodesic drift,
and whole trajectories of both integrators."""
import numpy as np
import pytest

import gauge_higgs_amd as gha
import oracle_lib as O

pytestmark = pytest.mark.gpu

# (tag, D, N, rep, beta, kappa, nmd, L); the last two run on the large-irrep kernels
CASES = [("su2_fund_frozen", 4, 2, "fund", 2.3, 0.6, 20, 4), ("su2_adj_frozen", 3, 2, "adj", 2.3, 0.5, 12, 4),
         ("su2_4_frozen", 3, 2, "4", 2.3, 0.5, 10, 3), ("su3_21_frozen", 3, 3, "2,1", 5.7, 0.3, 10, 3)]


def rel(a, b):
    a = np.asarray(a)
    a = a.view(np.float64).ravel() if np.iscomplexobj(a) else a.ravel()
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def make(D, rep, kappa, nmd, integ, prec=gha.FP64, N=2, beta=2.3, L=4):
    g = gha.GaugeHiggsHMC(L, rep, seed=1, D=D, N=N, precision=prec)
    g.configure(beta=beta, kappa=kappa, lam=0.5, tau=1.0, nmd=nmd, integ=integ)
    g.hot(0.8, 12345, 0.3)
    g.frozen_phi = True
    g.normalize_phi()
    return g


@pytest.mark.parametrize("tag,D,N,rep,beta,kappa,nmd,L", CASES, ids=[c[0] for c in CASES])
def test_frozen_against_reference(goldens, tag, D, N, rep, beta, kappa, nmd, L):
    arr, sc = goldens
    g = make(D, rep, kappa, nmd, gha.LEAPFROG, N=N, beta=beta, L=L)
    assert rel(g.get("phi"), arr[tag + "_phi0"]) < 1e-15
    g.refresh_momenta()
    assert rel(g.get("pi"), arr[tag + "_pi0"]) < 1e-15
    g.kick(0.05)
    assert rel(g.get("pi"), arr[tag + "_pi_kick"]) < 1e-12
    g.drift(0.05)
    assert rel(g.get("phi"), arr[tag + "_phi_drift"]) < 1e-12
    assert rel(g.get("pi"), arr[tag + "_pi_drift"]) < 1e-12
    for integ, name in ((gha.LEAPFROG, "lf"), (gha.OMELYAN, "omel")):
        g = make(D, rep, kappa, nmd, integ, N=N, beta=beta, L=L)
        dh, acc = [], []
        for _ in range(2):
            acc.append(float(g.trajectory()))
            dh.append(g.last_dH)
        assert np.max(np.abs(np.array(dh) - arr[f"{tag}_{name}_dH"])) < 1e-9
        assert np.array_equal(np.array(acc), arr[f"{tag}_{name}_acc"])
        assert rel(g.get("U"), arr[f"{tag}_{name}_U2"]) < 1e-10
        assert rel(g.get("phi"), arr[f"{tag}_{name}_phi2"]) < 1e-10


@pytest.mark.parametrize("prec,tol", [(gha.FP64, 1e-13), (gha.FP32MD, 2e-6)])
def test_frozen_length_preserved(prec, tol):
    g = make(4, "fund", 0.6, 10, gha.OMELYAN, prec)
    g.lat.trajectories(5, use_graph=True)
    assert np.max(np.abs(np.linalg.norm(g.get("phi"), axis=1) - 1.0)) < 5 * tol
    g.refresh_momenta()
    phi, pi = g.get("phi"), g.get("pi")
    assert np.max(np.abs(np.einsum("sk,sk->s", phi.conj(), pi).real)) < 10 * tol


def test_frozen_reversible_and_nested():
    for nested in (False, True):
        g = make(4, "fund", 0.6, 8, gha.LEAPFROG)
        g.configure(reunit_each_traj=False)
        if nested:
            g.lat.set_nested(4, 3)
        g.refresh_momenta()
        U0, phi0 = g.get("U"), g.get("phi")
        g.md_evolve()
        g.set("P", -g.get("P"))
        g.set("pi", -g.get("pi"))
        g.md_evolve()
        assert np.max(np.abs(g.get("U") - U0)) < 1e-10
        assert np.max(np.abs(g.get("phi") - phi0)) < 1e-10


@pytest.mark.parametrize("prec,tol", [(gha.FP64, 1e-13), (gha.FP32MD, 2e-6)])
def test_frozen_large_irrep_invariants(prec, tol):
    """SU(3) {6,1} (d = 48) on the large-irrep kernels: |phi| = 1 and
    phi.pi = 0 hold through graph-captured trajectories, the multi-timescale
    split is reversible, and the kick/drift API keeps the constraint."""
    g = gha.GaugeHiggsHMC(3, "6,1", seed=1, D=4, N=3, precision=prec)
    g.configure(beta=5.7, kappa=0.05, lam=0.5, tau=1.0, nmd=6, integ=gha.OMELYAN)
    g.hot(0.8, 12345, 0.3)
    g.frozen_phi = True
    g.normalize_phi()
    g.lat.trajectories(2, use_graph=True)
    assert np.max(np.abs(np.linalg.norm(g.get("phi"), axis=1) - 1.0)) < 5 * tol
    g.refresh_momenta()
    g.kick(0.05)
    g.drift(0.05)
    phi, pi = g.get("phi"), g.get("pi")
    assert np.max(np.abs(np.linalg.norm(phi, axis=1) - 1.0)) < 5 * tol
    assert np.max(np.abs(np.einsum("sk,sk->s", phi.conj(), pi).real)) < 10 * tol * max(1.0, np.abs(pi).max())
    if prec == gha.FP64:
        g.configure(reunit_each_traj=False)
        g.lat.set_nested(2, 3)
        g.refresh_momenta()
        U0, phi0 = g.get("U"), g.get("phi")
        g.md_evolve()
        g.set("P", -g.get("P"))
        g.set("pi", -g.get("pi"))
        g.md_evolve()
        assert np.max(np.abs(g.get("U") - U0)) < 1e-10 and np.max(np.abs(g.get("phi") - phi0)) < 1e-10


def test_frozen_multi_large_irrep():
    """The discrete-subgroup set-up of the reference's multi driver on a large
    irrep: frozen |phi| = 1 with the multi-invariant potential, SU(2) spin 6
    (d = 13), against the FP64 oracle over one trajectory."""
    f = 0.1 + 0.05 * np.arange(13)
    o = O.Oracle(3, 2, 3, "12", seed=1)
    o.multi_potential(f, 0.4)
    o.couplings(beta=2.3, kappa=0.3, lam=0.5, tau=1.0, nmd=8, omelyan=True)
    o.hot(0.8)
    o.phi_gaussian(12345, 0.3)
    o.frozen(True)
    o.normalize_phi()
    g = gha.GaugeHiggsHMC(3, "12", seed=1, D=3, N=2)
    g.configure(beta=2.3, kappa=0.3, lam=0.5, tau=1.0, nmd=8, integ=gha.OMELYAN)
    g.lat.set_multi_potential(f, 0.4)
    g.hot(0.8, 12345, 0.3)
    g.frozen_phi = True
    g.normalize_phi()
    assert np.max(np.abs(g.get("phi") - o.get("phi"))) < 1e-14
    acc = g.trajectory()
    a_o, dH_o, _, _ = o.trajectory()
    assert acc == a_o and abs(g.last_dH - dH_o) < 1e-7
    assert np.max(np.abs(g.get("phi") - o.get("phi"))) < 1e-8
