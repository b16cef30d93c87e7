"""Measurement layer on the GPU (observables.hpp:36-101): Wilson loops and
Polyakov loops traced in probe irreps, through the C ABI (ghx_wilson_loop,
ghx_polyakov_loop_rep), against the reference's own values on its
thermalised fields (fixture keys in loop_cases.py)."""
import numpy as np
import pytest

import gauge_higgs_amd as gha
import loop_cases as LC

pytestmark = pytest.mark.gpu


def lattices(goldens, tag, D, N, ext, prec=gha.FP64, nrep=1):
    arr, _ = goldens
    lat = gha.Lattices(D, N, [list(ext)] * nrep, rep="", precision=prec)
    U = arr[tag + "_U"].view(np.complex128).reshape(-1, N, N)
    for r in range(nrep):
        lat.set("U", U, replica=r)
    return lat


@pytest.mark.parametrize("prec,tol", [(gha.FP64, 1e-12), (gha.FP32MD, 1e-6)], ids=["fp64", "fp32"])
@pytest.mark.parametrize("case", LC.CASES, ids=[c[0] for c in LC.CASES])
def test_loops_match_reference(goldens, case, prec, tol):
    """Every probe, loop size and Polyakov direction of the fixtures. FP64
    agrees to rounding; FP32 link storage moves the averaged loops by O(1e-8)."""
    _, sc = goldens
    tag, D, N, ext, labels = case
    lat = lattices(goldens, tag, D, N, ext, prec)
    for Ra, Rb in LC.SIZES:
        w = lat.wilson_loops("fund", Ra, Rb)[0]
        assert abs(w - sc[f"{tag}_Wfund_{Ra}x{Rb}"]) < tol, (Ra, Rb)
    worst = 0.0
    for lab in labels:
        spec = LC.SPEC[lab]
        for Ra, Rb in LC.sizes_for(lab):
            w = lat.wilson_loops(spec, Ra, Rb)[0]
            err = abs(w - sc[f"{tag}_W_{lab}_{Ra}x{Rb}"])
            worst = max(worst, err)
            assert err < tol, (lab, Ra, Rb, w, sc[f"{tag}_W_{lab}_{Ra}x{Rb}"])
        for t in range(D):
            p = lat.polyakov_loops_rep(spec, t)[0]
            assert abs(p - sc[f"{tag}_P_{lab}_t{t}"]) < tol, (lab, t)
    print(f"{tag} prec={prec} worst |W - W_ref| = {worst:.2e}")


def test_loops_per_replica_and_cold(goldens):
    """Two replicas measured in one launch: the fixture field and a cold field
    (every loop 1 in every probe)."""
    tag, D, N, ext, _ = LC.CASES[0]
    _, sc = goldens
    lat = lattices(goldens, tag, D, N, ext, nrep=2)
    V = int(np.prod(ext))
    lat.set("U", np.tile(np.eye(N, dtype=np.complex128), (V * D, 1, 1)), replica=1)
    for lab in ("fund", "adj", "21", "61"):
        w = lat.wilson_loops(LC.SPEC[lab], 2, 2)
        assert abs(w[0] - sc[f"{tag}_W_{lab}_2x2"]) < 1e-12
        assert abs(w[1] - 1.0) < 1e-13
        p = lat.polyakov_loops_rep(LC.SPEC[lab], 1)
        assert abs(p[0] - sc[f"{tag}_P_{lab}_t1"]) < 1e-12
        assert abs(p[1] - 1.0) < 1e-13


def test_loops_on_hmc_fields():
    """On an evolved gauge-Higgs field the 1x1 fundamental loop is the
    plaquette, the default Polyakov direction is the observable's, and the
    adjoint probe equals Young rows {2} for SU(2)."""
    g = gha.GaugeHiggsHMC(4, "fund", seed=3, D=4, N=2)
    g.configure(beta=2.3, kappa=0.3, lam=0.5, tau=1.0, nmd=10, integ=gha.OMELYAN)
    g.hot(0.8, 12345, 0.3)
    g.lat.trajectories(3, use_graph=False)
    o = g.observables()
    assert abs(g.wilson_loop_fund(1, 1) - o["plaq"]) < 1e-13
    assert abs(g.polyakov_loop_rep("fund") - o["poly"]) < 1e-13
    assert abs(g.wilson_loop_rep("adj", 2, 1) - g.wilson_loop_rep("2", 2, 1)) < 1e-13
    g.cold()
    assert abs(g.wilson_loop_rep("3", 2, 3) - 1.0) < 1e-13


@pytest.mark.parametrize("probe", ["2,3", "1,1,1", "", "0"])
def test_invalid_probe_is_an_error(probe):
    lat = gha.Lattices(4, 3, [[4, 4, 4, 4]], rep="")
    lat.init_fields(0, gha.COLD, 0.0, 0, 0, 1.0)
    with pytest.raises(gha.GhxError):
        lat.wilson_loops(probe, 1, 1)
    with pytest.raises(gha.GhxError):
        lat.polyakov_loops_rep(probe, 0)


def test_invalid_loop_arguments():
    lat = gha.Lattices(3, 2, [[4, 4, 4]], rep="")
    lat.init_fields(0, gha.COLD, 0.0, 0, 0, 1.0)
    with pytest.raises(gha.GhxError):
        lat.wilson_loops("fund", -1, 1)
    with pytest.raises(gha.GhxError):
        lat.polyakov_loops_rep("fund", 3)
