"""Large irrep SU(3) {6,1} (d = 48, N^n = 2187) on the GPU: the LDS
matrix-free apply path against the reference's own dump (3^4 lattice) and the
FP64 oracle; reversibility of the MD with this path."""
import numpy as np
import pytest

import gauge_higgs_amd as gha
import oracle_lib as O

pytestmark = pytest.mark.gpu
TAG = "su3_61"


def relmax(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def make(prec=gha.FP64, nmd=16, L=3):
    g = gha.GaugeHiggsHMC(L, "6,1", seed=1, D=4, N=3, precision=prec).configure(beta=5.7, kappa=0.05, lam=0.5, nmd=nmd)
    g.hot(0.8, 12345, 0.3)
    return g


def test_rep_and_path():
    g = make()
    assert g.d == 48 and g.lat.dimT == 2187 and g.lat.nality == 1


def test_against_reference_goldens(goldens):
    arr, sc = goldens
    g = make()
    assert np.max(np.abs(g.get("U").ravel().view(np.float64) - arr[TAG + "_U0"])) < 1e-14
    assert np.max(np.abs(g.get("phi").ravel().view(np.float64) - arr[TAG + "_phi0"])) < 1e-15
    F, Fp = g.forces()
    assert relmax(F.ravel(), arr[TAG + "_Flink"]) < 1e-12
    assert relmax(Fp.ravel().view(np.float64), arr[TAG + "_Fphi"]) < 1e-12
    ob = g.observables()
    for k, key in [("Sg", "_Sg"), ("SH", "_SH"), ("plaq", "_plaq"), ("Llink", "_Llink"), ("Lphi", "_Lphi"), ("poly", "_poly")]:
        assert abs(ob[k] - sc[TAG + key]) <= 1e-12 * max(1.0, abs(sc[TAG + key])), k
    g.refresh_momenta()
    assert relmax(g.get("P").ravel(), arr[TAG + "_P0"]) < 5e-16
    assert relmax(g.get("pi").ravel().view(np.float64), arr[TAG + "_pi0"]) < 5e-16
    assert abs(g.hamiltonian() - sc[TAG + "_H0"]) < 1e-12 * abs(sc[TAG + "_H0"])
    g.kick(0.05)
    assert relmax(g.get("P").ravel(), arr[TAG + "_P_kick"]) < 1e-13
    assert relmax(g.get("pi").ravel().view(np.float64), arr[TAG + "_pi_kick"]) < 1e-13
    g.drift(0.05)
    assert np.max(np.abs(g.get("U").ravel().view(np.float64) - arr[TAG + "_U_drift"])) < 1e-13
    assert np.max(np.abs(g.get("phi").ravel().view(np.float64) - arr[TAG + "_phi_drift"])) < 1e-13


def test_trajectory_against_oracle():
    g = make(nmd=4)
    o = O.Oracle(4, 3, 3, "6,1", seed=1)
    o.couplings(beta=5.7, kappa=0.05, lam=0.5, nmd=4, omelyan=True)
    o.hot(0.8)
    o.phi_gaussian(12345, 0.3)
    acc = g.trajectory()
    a, dH, Hi, Hf = o.trajectory()
    assert abs(g.last_dH - dH) < 1e-9 and acc == a
    assert np.max(np.abs(g.get("U") - o.get("U"))) < 1e-10
    assert np.max(np.abs(g.get("phi") - o.get("phi"))) < 1e-10


def test_fp32_forces():
    g = make(prec=gha.FP32MD)
    o = O.Oracle(4, 3, 3, "6,1", seed=1)
    o.couplings(beta=5.7, kappa=0.05, lam=0.5)
    o.set("U", g.get("U"))
    o.set("phi", g.get("phi"))
    F, Fp = g.forces()
    Fo, Fpo = o.forces()
    assert relmax(F, Fo) < 2e-5 and relmax(Fp, Fpo) < 2e-5


def test_reversibility():
    g = make(nmd=8, L=4)
    g.configure(reunit_each_traj=False)
    g.refresh_momenta()
    U0, phi0 = g.get("U"), g.get("phi")
    g.md_evolve()
    g.set("P", -g.get("P"))
    g.set("pi", -g.get("pi"))
    g.md_evolve()
    assert np.max(np.abs(g.get("U") - U0)) < 1e-9
    assert np.max(np.abs(g.get("phi") - phi0)) < 1e-9
