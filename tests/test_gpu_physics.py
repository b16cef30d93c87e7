"""The reference's MD-exactness and physics checks rerun on the GPU:
finite-difference forces (test_gauge.cpp:13-49, test_scalar.cpp:96-120),
reversibility to 1e-9 (test_gauge.cpp:66-85, test_scalar.cpp:123-141),
<exp(-dH)> ~ 1 (test_gauge.cpp:111-128), the plaquette gate of the gh_hmc
defaults (README: 0.603 at beta=2.3), hipGraph replay equivalence, and the
drop-in CLIs."""
import os
import re
import subprocess

import numpy as np
import pytest

import gauge_higgs_amd as gha
import oracle_lib as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("D,N,rep", [(3, 2, "fund"), (3, 3, "2"), (4, 2, ""), (3, 2, "adj")])
def test_gpu_link_force_finite_difference(D, N, rep):
    g = gha.GaugeHiggsHMC(4, rep, seed=301, D=D, N=N).configure(beta=2.0, kappa=0.35, lam=0.4)
    g.hot(0.5, 77, 0.7)
    F, _ = g.forces()
    T = O.generators(N)
    U0 = g.get("U")
    eps, worst = 1e-5, 0.0
    V = U0.shape[0] // D

    def S():
        ob = g.observables()
        return ob["Sg"] + ob["SH"]
    for s in (0, 5, V - 1):
        for mu in range(D):
            for a in range(N * N - 1):
                U = U0.copy()
                U[s * D + mu] = O.expi(T[a] * eps) @ U0[s * D + mu]
                g.set("U", U)
                Sp = S()
                U[s * D + mu] = O.expi(-T[a] * eps) @ U0[s * D + mu]
                g.set("U", U)
                Sm = S()
                worst = max(worst, abs((Sp - Sm) / (2 * eps) - F[s * D + mu, a]))
    g.set("U", U0)
    assert worst < 1e-5


@pytest.mark.parametrize("D,N,rep,prec,tol", [(3, 2, "fund", gha.FP64, 1e-9), (3, 3, "", gha.FP64, 1e-9), (4, 3, "2", gha.FP64, 1e-9),
                                              (3, 2, "adj", gha.FP64, 1e-9), (4, 3, "2", gha.FP32MD, 2e-4)])
def test_gpu_reversibility(D, N, rep, prec, tol):
    g = gha.GaugeHiggsHMC(4, rep, seed=401, D=D, N=N, precision=prec).configure(beta=2.3, kappa=0.3, lam=0.5, nmd=12,
                                                                                 reunit_each_traj=False)
    g.hot(0.5, 1, 0.6)
    g.refresh_momenta()
    U0 = g.get("U")
    phi0 = g.get("phi") if rep else None
    g.md_evolve()
    g.set("P", -g.get("P"))
    if rep:
        g.set("pi", -g.get("pi"))
    g.md_evolve()
    assert np.max(np.abs(g.get("U") - U0)) < tol
    if rep:
        assert np.max(np.abs(g.get("phi") - phi0)) < tol


@pytest.mark.parametrize("D,N,rep,kappa", [(3, 2, "", 0.0), (3, 3, "", 0.0), (3, 2, "fund", 0.2), (3, 3, "2", 0.1)])
def test_gpu_exp_minus_dH(D, N, rep, kappa):
    g = gha.GaugeHiggsHMC(4, rep, seed=2024, D=D, N=N).configure(beta=2.3, kappa=kappa, lam=0.5, nmd=16)
    g.hot(0.3, 12345, 0.3, phi_mode=0 if rep else 1)
    n = 300
    s = 0.0
    for _ in range(n):
        g.trajectory()
        s += np.exp(-g.last_dH)
    assert abs(s / n - 1.0) < 0.15
    assert g.acceptance() > 0.5


def test_gpu_plaquette_gate_gh_hmc_defaults():
    """gh_hmc defaults: L=8, beta=2.3, 50 therm + 100 meas, nmd=20, Omelyan, seed 1."""
    g = gha.GaugeHMC(8, seed=1, D=4, N=2).configure(beta=2.3, nmd=20, tau=1.0)
    g.hot(0.8)
    g.lat.trajectories(50, use_graph=True)
    g.lat.set_counters(0, 0, 0)
    plaq, sexp = [], 0.0
    for _ in range(100):
        g.lat.trajectories(1, use_graph=True)
        plaq.append(g.observables()["plaq"])
        sexp += np.exp(-g.last_dH)
    m = np.mean(plaq)
    assert abs(m - 0.603) < 0.004, m
    assert abs(sexp / 100 - 1) < 0.1
    assert g.lat.max_unitarity() < 1e-12


@pytest.mark.parametrize("prec", [gha.FP64, gha.FP32MD])
def test_graph_replay_matches_stream(prec):
    """Trajectories replayed from one captured hipGraph equal eagerly launched ones bit for bit."""
    res = []
    for use_graph in (False, True):
        g = gha.GaugeHiggsHMC(6, "fund", seed=7, D=4, N=2, precision=prec).configure(beta=2.3, kappa=0.25, lam=0.5, nmd=8)
        g.hot(0.8, 12345, 0.3)
        per = []
        for _ in range(4):
            g.lat.trajectories(1, use_graph=use_graph)
            g.lat.sync()
            per.append((g.last_dH, g.traj_count, g.accept_count))
        res.append((per, g.get("U"), g.get("phi")))
    assert res[0][0] == res[1][0]
    assert np.array_equal(res[0][1], res[1][1]) and np.array_equal(res[0][2], res[1][2])


def test_cli_gh_hmc_and_hmc_higgs():
    out = subprocess.run([os.path.join(ROOT, "gauge_higgs_amd", "bin", "gh_hmc"), "4", "2.3", "5", "10", "10"], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert re.search(r"^plaquette\s+=\s+0\.\d+ \+/- ", out.stdout, re.M)
    assert re.search(r"^acceptance\s+=", out.stdout, re.M)
    out = subprocess.run([os.path.join(ROOT, "gauge_higgs_amd", "bin", "hmc_higgs"), "fund", "4", "2.3", "0.265", "0.5", "5", "10", "10"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    for key in ("plaquette", "L_phi", "L_link", "Polyakov", "chi_link", "Binder_link", "chi_phi", "acceptance", "<exp(-dH)>"):
        assert re.search(r"^" + re.escape(key) + r"\s+=", out.stdout, re.M), key
    bad = subprocess.run([os.path.join(ROOT, "gauge_higgs_amd", "bin", "hmc_higgs"), "1,2", "4", "2.3", "0.2", "0.5"],
                         capture_output=True, text=True, timeout=60)
    assert bad.returncode == 1 and "rep error:" in bad.stderr
