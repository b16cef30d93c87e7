"""Multi-invariant quartic potential (MultiInvariantPotential,
scalar_invariants.hpp:20-208, plugged in as GaugeHiggsHMC::potential) on the
GPU against the reference's own dumps and the FP64 oracle: the scalar force and
action with the potential, trajectories of both integrators, FP32 forces,
the frozen-length combination and the drop-in CLI."""
import os
import subprocess

import numpy as np
import pytest

import gauge_higgs_amd as gha
import oracle_lib as O

pytestmark = pytest.mark.gpu

# (tag, D, N, rep, beta, kappa, mu2, nmd, L, s): s scales the tolerances (the
# degree-12 Lagrange projectors of d = 13 lose ~1e-10 relative to roundoff).
# The last two run on the large-irrep kernels (LDS multi kernel, runtime d).
MULTI = [("multi_su2_fund", 4, 2, "fund", 2.3, 0.2, 0.4, 12, 4, 1.0), ("multi_su2_adj", 3, 2, "adj", 2.3, 0.2, 0.4, 12, 4, 1.0),
         ("multi_su3_2", 4, 3, "2", 5.7, 0.1, 0.5, 12, 4, 1.0), ("multi_su3_21", 3, 3, "2,1", 5.7, 0.1, 0.5, 10, 3, 1.0),
         ("multi_su2_12", 3, 2, "12", 2.3, 0.1, 0.4, 10, 3, 1e3)]
IDS = [c[0] for c in MULTI]


def rel(a, b):
    a = np.asarray(a)
    a = a.view(np.float64).ravel() if np.iscomplexobj(a) else a.ravel()
    b = np.asarray(b)
    b = b.view(np.float64).ravel() if np.iscomplexobj(b) else b.ravel()
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def make(case, integ=gha.LEAPFROG, prec=gha.FP64, f=None):
    tag, D, N, rep, beta, kappa, mu2, nmd, L, _ = case
    g = gha.GaugeHiggsHMC(L, rep, seed=1, D=D, N=N, precision=prec)
    g.configure(beta=beta, kappa=kappa, lam=0.5, tau=1.0, nmd=nmd, integ=integ)
    g.lat.set_multi_potential(f, mu2)
    g.hot(0.8, 12345, 0.3)
    return g


@pytest.mark.parametrize("case", MULTI, ids=IDS)
def test_against_reference(goldens, case):
    arr, sc = goldens
    tag, s = case[0], case[-1]
    f = arr[tag + "_f"]
    g = make(case, f=f)
    _, Fp = g.forces()
    assert rel(Fp, arr[tag + "_Fphi"]) < 1e-11 * s
    assert abs(g.observables()["SH"] - sc[tag + "_SH"]) < 1e-11 * s * max(1.0, abs(sc[tag + "_SH"]))
    g.refresh_momenta()
    assert abs(g.hamiltonian() - sc[tag + "_H0"]) < 1e-11 * s * abs(sc[tag + "_H0"])
    for integ, name in ((gha.LEAPFROG, "lf"), (gha.OMELYAN, "omel")):
        g = make(case, integ, f=f)
        dh, acc = [], []
        for _ in range(2):
            acc.append(float(g.trajectory()))
            dh.append(g.last_dH)
        assert np.max(np.abs(np.array(dh) - arr[f"{tag}_{name}_dH"])) < 1e-8 * max(1.0, s / 100)
        assert np.array_equal(np.array(acc), arr[f"{tag}_{name}_acc"])
        assert rel(g.get("phi"), arr[f"{tag}_{name}_phi2"]) < 1e-10 * s


@pytest.mark.parametrize("case", MULTI, ids=IDS)
def test_against_oracle_and_fp32(case):
    tag, D, N, rep, beta, kappa, mu2, nmd, L, s = case
    f = 0.1 + 0.25 * np.arange(len(gha.casimir_channels(N, rep)))
    o = O.Oracle(D, N, L, rep, seed=1)
    o.multi_potential(f, mu2)
    o.couplings(beta=beta, kappa=kappa, lam=0.5, tau=1.0, nmd=6, omelyan=True)
    o.hot(0.8)
    o.phi_gaussian(12345, 0.3)
    g = make(case, gha.OMELYAN, f=f)
    g.configure(nmd=6)
    F, Fp = g.forces()
    Fo, Fpo = o.forces()
    assert rel(F, Fo) < 1e-12 * s and rel(Fp, Fpo) < 1e-12 * s
    acc = g.trajectory()
    a_o, dH_o, _, _ = o.trajectory()
    assert abs(g.last_dH - dH_o) < 1e-9 * s and acc == a_o
    h = make(case, gha.OMELYAN, prec=gha.FP32MD, f=f)
    h.set("U", g.get("U"))
    h.set("phi", g.get("phi"))
    o.set("U", h.get("U"))
    o.set("phi", h.get("phi"))
    F32, Fp32 = h.forces()
    Fo, Fpo = o.forces()
    assert rel(F32, Fo) < 2e-5 * max(1.0, s / 100) and rel(Fp32, Fpo) < 2e-5 * max(1.0, s / 100)


def test_frozen_with_multi_reversible():
    case = MULTI[0]
    g = make(case, f=[0.3, 0.5])
    g.configure(reunit_each_traj=False)
    g.frozen_phi = True
    g.normalize_phi()
    g.refresh_momenta()
    U0, phi0 = g.get("U"), g.get("phi")
    g.md_evolve()
    g.set("P", -g.get("P"))
    g.set("pi", -g.get("pi"))
    g.md_evolve()
    assert np.max(np.abs(g.get("U") - U0)) < 1e-10 and np.max(np.abs(g.get("phi") - phi0)) < 1e-10
    assert np.max(np.abs(np.linalg.norm(g.get("phi"), axis=1) - 1)) < 1e-12


def test_large_irrep_multi_nested_and_limits():
    """The multi-timescale split on a large irrep moves the multi-invariant
    on-site force to the inner scale (onsite_step_kernel reads the same
    planes): one inner step reproduces the single-scale leapfrog to rounding,
    and the split is reversible. Irreps above the LDS kernel's bound are
    refused."""
    case = MULTI[3]
    f = [0.2, 0.35, 0.5, 0.65]
    a = make(case, gha.LEAPFROG, f=f)
    b = make(case, gha.LEAPFROG, f=f)
    a.refresh_momenta()
    b.set("P", a.get("P"))
    b.set("pi", a.get("pi"))
    b.lat.set_nested(case[7], 1)
    a.md_evolve()
    b.md_evolve()
    for k in ("U", "P", "phi", "pi"):
        assert np.max(np.abs(a.get(k) - b.get(k))) < 1e-11, k
    g = make(case, gha.LEAPFROG, f=f)
    g.configure(reunit_each_traj=False)
    g.lat.set_nested(2, 3)
    g.refresh_momenta()
    U0, phi0 = g.get("U"), g.get("phi")
    g.md_evolve()
    g.set("P", -g.get("P"))
    g.set("pi", -g.get("pi"))
    g.md_evolve()
    assert np.max(np.abs(g.get("U") - U0)) < 1e-10 and np.max(np.abs(g.get("phi") - phi0)) < 1e-10
    g = gha.GaugeHiggsHMC(3, "6,1", seed=1, D=3, N=3)
    with pytest.raises(gha.GhxError, match="dimension <= 16"):
        g.lat.set_multi_potential(np.ones(4), 0.5)


def test_restore_quartic():
    case = MULTI[0]
    a = make(case, f=None)
    b = make(case, f=[0.3, 0.5])
    b.lat.set_multi_potential(None, 0.0)
    assert np.array_equal(a.forces()[1], b.forces()[1])


def test_cli_hmc_higgs_multi(tmp_path):
    exe = os.path.join(os.path.dirname(gha.__file__), "bin", "hmc_higgs_multi")
    out = subprocess.run([exe, "2", "--N=3"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and "3 quartic channels" in out.stdout
    ts = str(tmp_path / "ts.txt")
    out = subprocess.run([exe, "2", "4", "5.7", "0.1", "0.5", "0.2,0.35,0.5", "5", "10", "8", "1.0", "1", ts],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "tau_int(L_link)" in out.stdout and "acceptance" in out.stdout
    assert len(open(ts).read().strip().splitlines()) == 11
    bad = subprocess.run([exe, "2", "4", "5.7", "0.1", "0.5", "0.2,0.35"], capture_output=True, text=True, timeout=60)
    assert bad.returncode == 1 and "need 3 couplings" in bad.stderr
    # a large irrep (SU(2) spin 6, 13 channels) with the frozen scalar, as in discrete-subgroup runs
    out = subprocess.run([exe, "12", "--N=2"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and "13 quartic channels" in out.stdout
    out = subprocess.run([exe, "12", "3", "2.3", "0.3", "0.4", "auto", "2", "4", "6", "1.0", "1", "--N=2", "--D=3"],
                         capture_output=True, text=True, timeout=120, env=dict(os.environ, GH_FROZEN="1"))
    assert out.returncode == 0, out.stderr
    assert "FROZEN" in out.stdout and "acceptance" in out.stdout
