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


# the last case sits at the engine's limits: d = 63 (of 64) and dim Sym^7 = 36
# {6}, {5,1} and {4} take the column kernel's product build of the top level
# (k = 6 as 3 + 3, k = 5 as 3 + 2 with the lower level kept, k = 4 as 2 + 2)
ENGINE_CASES = [("6,1", 3, 4, 3), ("12", 2, 3, 4), ("2,1", 3, 3, 4), ("4,2", 3, 4, 3), ("4,4", 3, 3, 3), ("7,1", 3, 3, 3),
                ("6", 3, 4, 3), ("5,1", 3, 3, 3), ("4", 3, 4, 3)]


@pytest.mark.parametrize("rep,N,D,L", ENGINE_CASES, ids=[c[0] for c in ENGINE_CASES])
@pytest.mark.parametrize("prec,tol", [(gha.FP64, 1e-12), (gha.FP32MD, 2e-5)], ids=["fp64", "fp32"])
def test_sym_engine_equals_tensor_apply(rep, N, D, L, prec, tol):
    """The symmetric-power engine D(U) = Yc^T (Sym^k1(U) x Sym^k2(U)) Yc / c
    against the slot-by-slot tensor apply on the same fields: link and scalar
    forces, the action and a trajectory agree to rounding."""
    out = []
    for engine in (0, 1):
        g = gha.GaugeHiggsHMC(L, rep, seed=2, D=D, N=N, precision=prec)
        g.configure(beta=5.7 if N == 3 else 2.3, kappa=0.3, lam=0.5, nmd=6)
        g.lat.set_sym_engine(engine)
        assert g.lat.sym_engine() == engine
        g.hot(0.8, 12345, 0.4)
        F, Fp = g.forces()
        sh = g.observables()["SH"]
        g.trajectory()
        out.append((F, Fp, sh, g.last_dH, g.get("phi"), g.get("U")))
    a, b = out
    assert relmax(b[0], a[0]) < tol and relmax(b[1], a[1]) < tol
    assert abs(b[2] - a[2]) < tol * max(1.0, abs(a[2]))
    assert abs(b[3] - a[3]) < 1e3 * tol
    assert relmax(b[4], a[4]) < 1e2 * tol and relmax(b[5], a[5]) < 1e2 * tol


@pytest.mark.parametrize("rep,N,D,L", ENGINE_CASES, ids=[c[0] for c in ENGINE_CASES])
@pytest.mark.parametrize("prec,tol", [(gha.FP64, 1e-12), (gha.FP32MD, 1e-5)], ids=["fp64", "fp32"])
def test_paired_link_kernels_equal_program_kernel(rep, N, D, L, prec, tol):
    """The paired engine kernels (two links per wave) against the
    one-link-per-wave LDS program on the same fields: link and scalar forces,
    the action and an MD trajectory agree to rounding. Mode 1 builds Sym^k(U) a
    column per lane into LDS; mode 2 keeps each lane's column in registers,
    taking the previous level's column from another lane, and forms the forward
    hop from the columns of Sym^k(U^T). Each runs exactly where it applies:
    SU(3) with first Young row <= 6 (so not SU(2) spin 6 nor {7,1}), mode 2 in
    FP32 with a second row <= 1 (mode 1 otherwise); with D = 3 one half of the
    last wave idles."""
    rows = [int(x) for x in rep.split(",")]
    k1, k2 = rows[0], rows[1] if len(rows) > 1 else 0
    paired = N == 3 and k1 <= 6
    out = []
    for mode in (0, 1, 2):
        g = gha.GaugeHiggsHMC(L, rep, seed=3, D=D, N=N, precision=prec)
        g.configure(beta=5.7 if N == 3 else 2.3, kappa=0.3, lam=0.5, nmd=4, reunit_each_traj=False)
        g.lat.set_sym_pairs(mode)
        assert g.lat.sym_engine() == 1
        want = 0 if not paired or mode == 0 else (2 if mode == 2 and prec == gha.FP32MD and k2 <= 1 else 1)
        assert g.lat.sym_pairs() == want
        g.hot(0.8, 12345, 0.4)
        F, Fp = g.forces()
        sh = g.observables()["SH"]
        g.refresh_momenta()
        g.md_evolve()
        out.append((F, Fp, sh, g.get("U"), g.get("phi")))
    a = out[0]
    for b in out[1:]:
        assert relmax(b[0], a[0]) < tol and relmax(b[1], a[1]) < tol
        assert abs(b[2] - a[2]) < tol * max(1.0, abs(a[2]))
        assert relmax(b[3], a[3]) < 1e2 * tol and relmax(b[4], a[4]) < 1e2 * tol


def test_irrep_beyond_one_component_per_lane_is_refused():
    """The large-irrep kernels hold one component per lane of a 64-wide wave
    (d <= 64). SU(3) {7,2} (d = 81, dim S = 216) would fit the engine's tables,
    but it is refused at ghx_set_rep with GHX_E_UNSUPPORTED, in both
    precisions, instead of running with components dropped."""
    for prec in (gha.FP64, gha.FP32MD):
        with pytest.raises(gha.GhxError) as e:
            gha.GaugeHiggsHMC(4, "7,2", seed=1, D=3, N=3, precision=prec)
        assert e.value.code == -4 and "no compiled kernels" in str(e.value)


def test_sym_engine_default_and_speed():
    """{6,1} uses the engine by default; one MD step is faster than with the
    tensor apply."""
    g = gha.GaugeHiggsHMC(8, "6,1", seed=1, D=4, N=3, precision=gha.FP32MD).configure(beta=5.7, kappa=0.05, lam=0.5)
    g.hot(0.8, 12345, 0.3)
    g.refresh_momenta()
    assert g.lat.sym_engine() == 1
    ms = []
    for engine in (1, 0):
        g.lat.set_sym_engine(engine)
        g.lat.time_md_steps(2)
        ms.append(g.lat.time_md_steps(6) / 6)
    print(f"{{6,1}} 8^4 FP32 MD step: engine {ms[0]:.3f} ms, tensor {ms[1]:.3f} ms")
    assert ms[0] < ms[1]


def test_engine_uneven_replicas_match_oracle():
    """Two replicas of SU(3) {4,2} (d = 27, symmetric-power engine) with
    different extents and couplings in one launch per MD step, against
    independent oracle chains (the reference's tensor apply)."""
    ext = [[3, 3, 3, 3], [4, 3, 3, 3]]
    kap, bet = [0.2, 0.3], [5.7, 5.5]
    rs = gha.ReplicaSet(4, 3, ext, "4,2", bet, kap, [0.5] * 2, seed0=5, nmd=4, precision=gha.FP64)
    assert rs.lat.sym_engine() == 1
    rs.enabled = False
    ors = []
    for r in range(2):
        o = O.Oracle(4, 3, ext[r], "4,2", seed=5 + r * 1000003)
        o.couplings(beta=bet[r], kappa=kap[r], lam=0.5, nmd=4)
        o.hot(0.8)
        o.phi_gaussian(12345, 0.3)
        ors.append(o)
    acc, dH = rs.lat.trajectory_all()
    for r, o in enumerate(ors):
        a, d, _, _ = o.trajectory()
        assert abs(dH[r] - d) < 1e-8 and acc[r] == a
        assert np.max(np.abs(rs.lat.get("U", r) - o.get("U"))) < 1e-10
        assert np.max(np.abs(rs.lat.get("phi", r) - o.get("phi"))) < 1e-10
