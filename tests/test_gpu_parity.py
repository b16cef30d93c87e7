"""GPU (HIP, gfx950) vs FP64 oracle parity through the C ABI.

Same seeds, same fields: the oracle restates the reference (pinned by
test_oracle_golden.py), so agreement here is agreement with the reference.
Tolerances (FP64 mode): per-site forces, actions and single kick/drift steps
to 1e-12 relative; fields after one MD trajectory to 1e-10 absolute; dH to
1e-9 absolute (|H| ~ 1e3-1e4, so this is ~1e-13 of H). Integer RNG keys are
exact; Gaussian draws (phi, P, pi) are equal except where glibc misrounds log
or cos, where the device holds the correctly rounded draw (gauss_policy.py).
"""
import numpy as np
import pytest

import gauge_higgs_amd as gha
import oracle_lib as O
from gauss_policy import TAG_P, TAG_P_PURE, TAG_PI, assert_equal_but_glibc_misrounds, momentum_keys, scalar_keys

pytestmark = pytest.mark.gpu

# (id, D, N, rep, L, beta, kappa)
CASES = [
    ("su2_pure", 4, 2, "", 4, 2.3, 0.0),
    ("su3_pure", 4, 3, "", 4, 5.7, 0.0),
    ("su2_fund", 4, 2, "fund", 4, 2.3, 0.265),
    ("su2_fund_aniso", 4, 2, "fund", (4, 3, 5, 6), 2.3, 0.265),
    ("su2_adj_d3", 3, 2, "adj", 4, 2.3, 0.3),
    ("su3_fund_d3", 3, 3, "fund", 4, 5.7, 0.2),
    ("su3_sym2", 4, 3, "2", 4, 5.7, 0.1),
    ("su3_adj_d3", 3, 3, "adj", 3, 5.7, 0.15),
    ("su3_asym_11", 4, 3, "1,1", 3, 5.7, 0.1),
    ("su2_spin1_real", 3, 2, "2:real", 4, 2.3, 0.2),
    ("su2_spin32", 3, 2, "3", 4, 2.3, 0.2),
]
IDS = [c[0] for c in CASES]


def make_pair(case, nmd=8, omelyan=True, precision=gha.FP64, seed=1):
    _, D, N, rep, L, beta, kappa = case
    ext = list(L) if not np.isscalar(L) else [L] * D
    g = gha.GaugeHiggsHMC(ext, rep, seed=seed, D=D, N=N, precision=precision)
    g.configure(beta=beta, kappa=kappa, lam=0.5, tau=1.0, nmd=nmd, integ=gha.OMELYAN if omelyan else gha.LEAPFROG)
    o = O.Oracle(D, N, ext, rep, seed=seed)
    o.couplings(beta=beta, kappa=kappa, lam=0.5, tau=1.0, nmd=nmd, omelyan=omelyan)
    g.hot(0.8, 12345, 0.3)
    o.hot(0.8)
    if rep:
        o.phi_gaussian(12345, 0.3)
    return g, o


def relmax(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_start_fields(case):
    g, o = make_pair(case)
    assert np.max(np.abs(g.get("U") - o.get("U"))) < 1e-14
    if case[3]:
        d = g.d
        assert_equal_but_glibc_misrounds(g.get("phi").view(np.float64), o.get("phi").view(np.float64),
                                         lambda i: scalar_keys(i, d, 12345), 1, 0.3, case[0] + " phi")


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_momentum_refresh(case):
    g, o = make_pair(case)
    g.refresh_momenta()
    o.refresh()
    _, D, N, rep = case[:4]
    sP = O.key(TAG_P if rep else TAG_P_PURE, 0)
    assert_equal_but_glibc_misrounds(g.get("P"), o.get("P"), lambda i: momentum_keys(i, D, N * N - 1, sP), 1,
                                     what=case[0] + " P")
    if rep:
        d, sPi = g.d, O.key(TAG_PI, 0)
        assert_equal_but_glibc_misrounds(g.get("pi").view(np.float64), o.get("pi").view(np.float64),
                                         lambda i: scalar_keys(i, d, sPi), 1, what=case[0] + " pi")


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_forces(case):
    g, o = make_pair(case)
    F, Fp = g.forces()
    Fo, Fpo = o.forces()
    assert relmax(F, Fo) < 1e-12
    if case[3]:
        assert relmax(Fp, Fpo) < 1e-12


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_observables_and_hamiltonian(case):
    g, o = make_pair(case)
    g.refresh_momenta()
    o.refresh()
    ob, obo = g.observables(), o.observables()
    for k in ("plaq", "Sg", "SH", "Llink", "Lphi", "poly"):
        assert abs(ob[k] - obo[k]) <= 1e-12 * max(1.0, abs(obo[k])), k
    hp, hpo = g.hamiltonian_parts(), o.ham_parts()
    assert np.all(np.abs(hp - hpo) <= 1e-12 * np.maximum(1.0, np.abs(hpo)))


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_kick_then_drift(case):
    g, o = make_pair(case)
    g.refresh_momenta()
    o.refresh()
    g.kick(0.05)
    o.kick(0.05)
    assert relmax(g.get("P"), o.get("P")) < 1e-13
    if case[3]:
        assert relmax(g.get("pi"), o.get("pi")) < 1e-13
    g.drift(0.05)
    o.drift(0.05)
    assert np.max(np.abs(g.get("U") - o.get("U"))) < 1e-13
    if case[3]:
        assert np.max(np.abs(g.get("phi") - o.get("phi"))) < 1e-13


@pytest.mark.parametrize("omelyan", [False, True])
@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_md_evolve(case, omelyan):
    g, o = make_pair(case, nmd=6, omelyan=omelyan)
    g.refresh_momenta()
    o.refresh()
    g.md_evolve()
    o.md_evolve()
    assert np.max(np.abs(g.get("U") - o.get("U"))) < 1e-10
    assert relmax(g.get("P"), o.get("P")) < 1e-10
    if case[3]:
        assert np.max(np.abs(g.get("phi") - o.get("phi"))) < 1e-10


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_trajectories(case):
    g, o = make_pair(case, nmd=10)
    for _ in range(2):
        acc = g.trajectory()
        acc_o, dH_o, Hi, _ = o.trajectory()
        assert abs(g.last_dH - dH_o) < 1e-9 * max(1.0, abs(Hi) / 1e3)
        assert acc == acc_o
    assert g.traj_count == o.traj_count and g.accept_count == o.accept_count
    assert np.max(np.abs(g.get("U") - o.get("U"))) < 1e-10
    if case[3]:
        assert np.max(np.abs(g.get("phi") - o.get("phi"))) < 1e-10


GOLDEN_CASES = [("su2_pure", 4, 2, "", 4, 2.3, 0.0, 20), ("su3_pure", 4, 3, "", 4, 5.7, 0.0, 12),
                ("su2_fund", 4, 2, "fund", 4, 2.3, 0.265, 20), ("su2_adj", 3, 2, "adj", 4, 2.3, 0.3, 12),
                ("su3_fund", 3, 3, "fund", 4, 5.7, 0.2, 12), ("su3_sym2", 4, 3, "2", 4, 5.7, 0.1, 20)]


@pytest.mark.parametrize("gc", GOLDEN_CASES, ids=[c[0] for c in GOLDEN_CASES])
def test_against_reference_goldens(goldens, gc):
    """The GPU against the reference's own dump (no oracle in between)."""
    arr, sc = goldens
    tag, D, N, rep, L, beta, kappa, nmd = gc
    for integ in ("lf", "omel"):
        g = gha.GaugeHiggsHMC(L, rep, seed=1, D=D, N=N)
        g.configure(beta=beta, kappa=kappa, lam=0.5, tau=1.0, nmd=nmd, integ=gha.OMELYAN if integ == "omel" else gha.LEAPFROG)
        g.hot(0.8, 12345, 0.3)
        if integ == "lf":
            F, Fp = g.forces()
            ref = arr[tag + ("_Flink" if rep else "_F")].reshape(F.shape)
            assert relmax(F, ref) < 1e-12
            if rep:
                assert relmax(Fp.view(np.float64).ravel(), arr[tag + "_Fphi"]) < 1e-12
            ob = g.observables()
            assert abs(ob["Sg"] - sc[tag + "_Sg"]) < 1e-12 * abs(sc[tag + "_Sg"])
            assert abs(ob["plaq"] - sc[tag + "_plaq"]) < 1e-13
        ntr = 2 if rep else 3
        dH = []
        acc = []
        for _ in range(ntr):
            acc.append(float(g.trajectory()))
            dH.append(g.last_dH)
        assert np.max(np.abs(np.array(dH) - arr[f"{tag}_{integ}_dH"])) < 1e-9
        assert np.array_equal(np.array(acc), arr[f"{tag}_{integ}_acc"])
        ref_U = arr[f"{tag}_{integ}_U{ntr}"].view(np.complex128).reshape(-1, N, N)
        assert np.max(np.abs(g.get("U") - ref_U)) < 1e-10


FP32_CASES = [c for c in CASES if c[0] in ("su2_fund", "su3_sym2", "su3_pure", "su2_adj_d3")]


@pytest.mark.parametrize("case", FP32_CASES, ids=[c[0] for c in FP32_CASES])
def test_fp32_md_against_fp64(case):
    """FP32 MD: forces to ~1e-6 relative on the same (float-rounded) input;
    dH of a trajectory against the FP64 oracle within the stated 5e-3 absolute."""
    g, o = make_pair(case, nmd=10, precision=gha.FP32MD)
    U = g.get("U")
    o.set("U", U)
    if case[3]:
        o.set("phi", g.get("phi"))
    F, Fp = g.forces()
    Fo, Fpo = o.forces()
    assert relmax(F, Fo) < 2e-5
    if case[3]:
        assert relmax(Fp, Fpo) < 2e-5
    ob, obo = g.observables(), o.observables()
    assert abs(ob["Sg"] - obo["Sg"]) < 1e-12 * abs(obo["Sg"])  # H is reduced in FP64
    acc = g.trajectory()
    acc_o, dH_o, _, _ = o.trajectory()
    assert abs(g.last_dH - dH_o) < 5e-3


@pytest.mark.parametrize("N,spec", [(3, "2"), (2, "2")])
def test_sym2_path_equals_generic_tensor_path(N, spec, monkeypatch):
    """The direct U Phi U^T kernels for {2} against the generic lift/slot/project kernels."""
    res = []
    for generic in (False, True):
        if generic:
            monkeypatch.setenv("GHX_GENERIC_TENSOR", "1")
        else:
            monkeypatch.delenv("GHX_GENERIC_TENSOR", raising=False)
        g = gha.GaugeHiggsHMC(4, spec, seed=5, D=4, N=N).configure(beta=5.7, kappa=0.12, lam=0.5, nmd=6)
        g.hot(0.8, 12345, 0.5)
        F, Fp = g.forces()
        g.trajectory()
        res.append((F, Fp, g.last_dH, g.get("U")))
    monkeypatch.delenv("GHX_GENERIC_TENSOR", raising=False)
    assert relmax(res[0][0], res[1][0]) < 1e-13 and relmax(res[0][1], res[1][1]) < 1e-13
    assert abs(res[0][2] - res[1][2]) < 1e-10
    assert np.max(np.abs(res[0][3] - res[1][3])) < 1e-12


@pytest.mark.parametrize("eps", [0.05, 0.5, 3.0])
def test_fp32_su3_drift_exponential(eps):
    """FP32 SU(3) drift (Cayley-Hamilton series with scaling) against the FP64
    Taylor-18 oracle on identical float inputs; eps = 3 exercises the squaring."""
    g = gha.GaugeHiggsHMC(4, "", seed=3, D=4, N=3, precision=gha.FP32MD).configure(beta=5.7, nmd=4)
    g.hot(0.8)
    g.refresh_momenta()
    o = O.Oracle(4, 3, 4, "", seed=3)
    o.set("U", g.get("U"))
    o.set("P", g.get("P"))
    g.drift(eps)
    o.drift(eps)
    U = g.get("U")
    assert np.max(np.abs(U - o.get("U"))) < 2e-6 * max(1.0, eps)
    # stays on the group to float rounding
    assert np.max(np.abs(np.einsum("lji,ljk->lik", U.conj(), U) - np.eye(3))) < 2e-6 * max(1.0, eps)


@pytest.mark.parametrize("seg", [1, 3, 4, 16])
@pytest.mark.parametrize("prec", [gha.FP64, gha.FP32MD])
def test_streaming_kernel_equals_site_kernel(seg, prec):
    """The t-streaming launch shape performs the same per-site arithmetic: bitwise equal."""
    out = []
    for s in (0, seg):
        g = gha.GaugeHiggsHMC([4, 6, 4, 10], "2", seed=9, D=4, N=3, precision=prec).configure(beta=5.7, kappa=0.1, lam=0.5, nmd=4)
        g.lat.set_stream_segment(s)
        g.hot(0.8, 12345, 0.5)
        g.trajectory()
        out.append((g.last_dH, g.get("U"), g.get("phi"), g.get("P")))
    assert out[0][0] == out[1][0]
    for k in (1, 2, 3):
        assert np.array_equal(out[0][k], out[1][k])


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("prec", [gha.FP64, gha.FP32MD])
def test_block_walk_equals_storage_order(order, prec):
    """Walking the blocks in another order, or covering 4D tiles of 128 sites
    (orders 3-6: 16x4x2x1, 8x4x2x2, 16x2x2x2, 32x2x2x1), changes only the
    schedule: every site update reads the step's input fields and writes its
    own site, so the trajectory is bitwise equal. 32x8x4x8 admits every walk
    (whole-block slices and planes, Lt divisible by 8, extents divisible by
    every tile). The 4D tiles do not hold whole x-lines per wave, so there
    the x-lane hand-over of lower staples is switched off on both sides."""
    out = []
    for o in (0, order):
        g = gha.GaugeHiggsHMC([32, 8, 4, 8], "2", seed=9, D=4, N=3, precision=prec).configure(beta=5.7, kappa=0.1, lam=0.5, nmd=4)
        if order >= 3:
            g.lat.set_x_lanes(0)
        g.lat.set_block_order(o)
        assert g.lat.block_order() == o
        g.hot(0.8, 12345, 0.5)
        g.trajectory()
        g.lat.trajectories(2, use_graph=True)
        out.append((g.last_dH, g.get("U"), g.get("phi"), g.get("P")))
    assert out[0][0] == out[1][0]
    for k in (1, 2, 3):
        assert np.array_equal(out[0][k], out[1][k])


XL_CASES = [[4, 4, 4, 4], [8, 4, 2, 4], [2, 4, 8, 4], [16, 2, 2, 4], [32, 2, 2, 2], [64, 2, 2, 2]]


@pytest.mark.parametrize("dims", XL_CASES, ids=["x".join(map(str, d)) for d in XL_CASES])
@pytest.mark.parametrize("rep", ["2", "fund", ""])
def test_x_lanes_equal_own_staples(dims, rep):
    """SU(3) FP32 MD with the lower staple of plane (0, mu) handed from lane x-0
    to lane x (x-lines in lanes, links updated in the order 1, 2, 3, 0) against
    every site forming its own backward staples: forces, the one-site-per-thread
    step and the streaming step agree to FP32 rounding, for each L0 that fills
    waves with whole x-lines (2 to 64)."""
    out = []
    for xl, seg in ((0, 0), (1, 0), (1, 2)):
        g = gha.GaugeHiggsHMC(dims, rep, seed=9, D=4, N=3, precision=gha.FP32MD).configure(beta=5.7, kappa=0.2, lam=0.5, nmd=6)
        g.lat.set_x_lanes(xl)
        assert g.lat.x_lanes() == xl
        g.lat.set_stream_segment(seg)
        g.hot(0.8, 12345, 0.5)
        F, Fp = g.forces()
        g.refresh_momenta()  # MD without accept/reject: the comparison cannot hinge on a Metropolis decision
        h0 = g.hamiltonian()
        g.md_evolve()
        out.append((F, Fp, g.hamiltonian() - h0, g.get("U"), g.get("phi") if rep else None, g.get("P")))
    base = out[0]
    for b in out[1:]:
        assert relmax(b[0], base[0]) < 2e-5
        if rep:
            assert relmax(b[1], base[1]) < 2e-5
            assert relmax(b[4], base[4]) < 2e-3
        assert abs(b[2] - base[2]) < 2e-2
        for k in (3, 5):
            assert relmax(b[k], base[k]) < 2e-3
    # the one-site-per-thread and the streaming kernel run the same arithmetic
    assert out[1][2] == out[2][2]
    for k in (3, 5):
        assert np.array_equal(out[1][k], out[2][k])


def test_x_lanes_with_the_symmetric_power_engine():
    """SU(3) + {6,1} (d = 48) on an 8 x 4 x 2 x 4 lattice: the gauge part of the
    step runs with x-lanes and adds the engine's matter link force G. Forces
    and the MD evolution agree with the step that forms every staple itself."""
    out = []
    for xl in (0, 1):
        g = gha.GaugeHiggsHMC([8, 4, 2, 4], "6,1", seed=9, D=4, N=3, precision=gha.FP32MD).configure(beta=5.7, kappa=0.05, lam=0.5,
                                                                                                       nmd=4)
        g.lat.set_x_lanes(xl)
        assert g.lat.x_lanes() == xl and g.lat.sym_engine() == 1
        g.hot(0.4, 12345, 0.27)
        F, Fp = g.forces()
        g.refresh_momenta()
        g.md_evolve()
        out.append((F, Fp, g.get("U"), g.get("phi"), g.get("P")))
    assert relmax(out[1][0], out[0][0]) < 2e-5 and relmax(out[1][1], out[0][1]) < 2e-5
    for k in (2, 3, 4):
        assert relmax(out[1][k], out[0][k]) < 2e-3


@pytest.mark.parametrize("dims,prec", [([6, 4, 4, 4], gha.FP32MD), ([4, 6, 4, 10], gha.FP32MD), ([4, 4, 4, 4], gha.FP64)])
def test_x_lanes_off_where_waves_do_not_hold_x_lines(dims, prec):
    g = gha.GaugeHiggsHMC(dims, "2", seed=9, D=4, N=3, precision=prec)
    assert g.lat.x_lanes() == 0


def test_x_lanes_report_the_launch_shape():
    g = gha.GaugeHiggsHMC([32, 8, 4, 8], "2", seed=9, D=4, N=3, precision=gha.FP32MD)
    assert g.lat.x_lanes() == 1
    g.lat.set_link_lanes(1)
    assert g.lat.x_lanes() == 0
    g.lat.set_link_lanes(0)
    g.lat.set_block_order(3)  # 4D tiles of 16 x 4 x 2 x 1 sites
    assert g.lat.block_order() == 3 and g.lat.x_lanes() == 0
    g.lat.set_block_order(1)
    assert g.lat.x_lanes() == 1


def test_x_lanes_with_uneven_replicas():
    """Three SU(3) + {2} replicas of different extents (and x-lines of 8, 4 and
    16 sites) in one launch per MD step, FP32 MD: the MD evolution (refresh, 6
    Omelyan steps; no accept/reject, so the comparison cannot hinge on a
    Metropolis decision) with the x-lane hand-over agrees with every site
    forming its own staples, replica by replica."""
    from gauge_higgs_amd.replicas import ReplicaSet
    ext = [[8, 4, 4, 4], [4, 4, 4, 4], [16, 2, 2, 4]]
    out = []
    for xl in (0, 1):
        rs = ReplicaSet(4, 3, ext, "2", [5.7] * 3, [0.1, 0.15, 0.2], [0.5] * 3, nmd=6, precision=gha.FP32MD)
        rs.lat.set_x_lanes(xl)
        assert rs.lat.x_lanes() == xl
        rs.lat.refresh_momenta()
        rs.lat.md_evolve()
        out.append([(rs.lat.get("U", r), rs.lat.get("phi", r), rs.lat.get("P", r)) for r in range(3)])
        rs.lat.close()
    for r in range(3):
        for k in range(3):
            assert relmax(out[1][r][k], out[0][r][k]) < 2e-3


def test_block_walk_falls_back_where_it_does_not_apply():
    g = gha.GaugeHiggsHMC([4, 6, 4, 10], "2", seed=9, D=4, N=3)
    g.lat.set_block_order(1)
    assert g.lat.block_order() == 0


LINK_CASES = [("fund", 2, 4), ("2", 3, 4), ("adj", 2, 3), ("", 2, 4), ("adj", 3, 4), ("1,1", 3, 3), ("4,2", 3, 4), ("6,1", 3, 4)]


@pytest.mark.parametrize("rep,N,D", LINK_CASES, ids=[f"{r or 'pure'}-N{n}-D{d}" for r, n, d in LINK_CASES])
@pytest.mark.parametrize("prec,tol", [(gha.FP64, 1e-12), (gha.FP32MD, 2e-5)], ids=["fp64", "fp32"])
def test_link_lanes_equal_site_kernel(rep, N, D, prec, tol):
    """The link-parallel MD step (one lane per link) does the per-link
    arithmetic of the site kernel and sums the scalar force in another order:
    forces, and fields after a trajectory, agree to rounding."""
    out = []
    for lanes in (0, 1):
        g = gha.GaugeHiggsHMC([4, 6, 4, 5][:D], rep, seed=9, D=D, N=N, precision=prec)
        g.configure(beta=2.3 if N == 2 else 5.7, kappa=0.2, lam=0.5, nmd=6)
        g.lat.set_link_lanes(lanes)
        assert g.lat.link_lanes() == lanes
        g.hot(0.8, 12345, 0.5)
        F, Fp = g.forces()
        g.trajectory()
        out.append((F, Fp, g.last_dH, g.get("U"), g.get("phi") if rep else None, g.get("P")))
    a, b = out
    assert relmax(b[0], a[0]) < tol
    if rep:
        assert relmax(b[1], a[1]) < tol
    assert abs(b[2] - a[2]) < 1e3 * tol
    for k in (3, 5):
        assert relmax(b[k], a[k]) < 1e2 * tol
    if rep:
        assert relmax(b[4], a[4]) < 1e2 * tol


def test_link_lanes_with_frozen_multi_nested():
    """The link-parallel step with the frozen scalar, the multi-invariant
    potential and the multi-timescale split (which passes F_NO_HOP /
    F_NO_GAUGE / F_NO_ONSITE): a reversible trajectory, close to the site
    kernel's."""
    res = []
    for lanes in (0, 1):
        g = gha.GaugeHiggsHMC(4, "2", seed=3, D=4, N=3).configure(beta=5.7, kappa=0.2, lam=0.5, nmd=4, reunit_each_traj=False)
        g.lat.set_link_lanes(lanes)
        g.lat.set_multi_potential([0.2, 0.35, 0.5], 0.4)
        g.hot(0.8, 12345, 0.5)
        g.frozen_phi = True
        g.normalize_phi()
        g.lat.set_nested(3, 2)
        g.refresh_momenta()
        U0, phi0 = g.get("U"), g.get("phi")
        g.md_evolve()
        res.append((g.get("U"), g.get("phi")))
        g.set("P", -g.get("P"))
        g.set("pi", -g.get("pi"))
        g.md_evolve()
        assert np.max(np.abs(g.get("U") - U0)) < 1e-10 and np.max(np.abs(g.get("phi") - phi0)) < 1e-10
    assert relmax(res[1][0], res[0][0]) < 1e-11 and relmax(res[1][1], res[0][1]) < 1e-11


@pytest.mark.parametrize("case", [("2", 8), ("6,1", 4)], ids=["su3-sextet", "su3-61-engine"])
def test_autotune_restores_fields(case):
    """The trial steps of every candidate shape (and, for a large irrep, the
    gauge-only launches that decide the staple split) leave the fields as they
    were."""
    rep, L = case
    g = gha.GaugeHiggsHMC(L, rep, seed=2, D=4, N=3, precision=gha.FP32MD).configure(beta=5.7, kappa=0.1, lam=0.5)
    g.hot(0.8, 12345, 0.5)
    g.refresh_momenta()
    before = [g.get(k) for k in ("U", "P", "phi", "pi")]
    seg, ms = g.lat.autotune(steps=3)
    assert seg in (0, 1, 2, 4, 8) and ms > 0 and g.lat.stream_segment() == seg
    assert g.lat.link_split() in (0, 1)
    after = [g.get(k) for k in ("U", "P", "phi", "pi")]
    for a, b in zip(before, after):
        assert np.array_equal(a, b)
    assert g.lat.store_policy() in (0, 1)


@pytest.mark.parametrize("case", [("fund", 2, 6, 0), ("fund", 2, 6, 1), ("2", 3, 8, 0), ("6,1", 3, 4, 1)],
                         ids=["su2-site", "su2-lanes", "su3-sextet-xl", "su3-61-lanes"])
def test_store_policy_changes_no_value(case):
    """The MD step's stores streamed past the caches or stored with the default
    policy (GHX_TUNE_STORE_POLICY): the fields after MD agree bit for bit."""
    rep, N, L, lanes = case
    out = []
    for pol in (0, 1):
        g = gha.GaugeHiggsHMC(L, rep, seed=2, D=4, N=N, precision=gha.FP32MD).configure(beta=2.3 if N == 2 else 5.7, kappa=0.1,
                                                                                       lam=0.5, nmd=4, reunit_each_traj=False)
        g.hot(0.8, 12345, 0.5)
        g.lat.set_link_lanes(lanes)
        g.lat.set_store_policy(pol)
        assert g.lat.store_policy() == pol
        g.refresh_momenta()
        g.md_evolve()
        out.append([g.get(k) for k in ("U", "P", "phi", "pi")])
    for a, b in zip(*out):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("case", [("", 2, 4, gha.FP64, 1e-12), ("", 2, 3, gha.FP64, 1e-12), ("", 3, 4, gha.FP32MD, 2e-5),
                                  ("6,1", 3, 4, gha.FP32MD, 2e-5), ("6,1", 3, 4, gha.FP64, 1e-12)],
                         ids=["su2-pure-fp64", "su2-pure-d3", "su3-pure-fp32", "su3-61-fp32", "su3-61-fp64"])
def test_link_split_equals_link_lanes(case):
    """The link-parallel step with each link's staples split over its D - 1
    planes (md_link3_kernel; pure gauge, or the gauge part of a large irrep)
    against one lane per link (md_link_kernel): forces, and the links and
    momenta after MD, agree to the rounding of the staple sum."""
    rep, N, D, prec, tol = case
    out = []
    for split in (0, 1):
        g = gha.GaugeHiggsHMC(4, rep, seed=4, D=D, N=N, precision=prec).configure(beta=2.3 if N == 2 else 5.7, kappa=0.05, lam=0.5,
                                                                                  nmd=4, reunit_each_traj=False)
        g.hot(0.8, 12345, 0.3)
        g.lat.set_link_lanes(1)
        g.lat.set_link_split(split)
        assert g.lat.link_split() == split
        F, _ = g.forces()
        g.refresh_momenta()
        g.md_evolve()
        out.append((F, g.get("U"), g.get("P")))
    a, b = out
    assert relmax(b[0], a[0]) < tol
    assert relmax(b[1], a[1]) < 1e2 * tol and relmax(b[2], a[2]) < 1e2 * tol
    h = gha.GaugeHiggsHMC(4, "fund", seed=4, D=4, N=2)
    h.lat.set_link_lanes(1)
    assert h.lat.link_split() == 0  # a scalar in the kernel: one lane per link
