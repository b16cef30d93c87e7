"""Pin the FP64 oracle to the reference's own outputs.

The fixtures in tests/golden/ were written by oracle/_ref/ref_harness, which
compiles the reference headers in place (tests/golden/make_golden.py). Integer
RNG streams, uniforms and Gaussians must be bit-identical; forces, actions and
fields agree to rounding (the oracle mirrors the reference expression by
expression, but g++ contracts FMAs per call site, so a few ulps differ).
"""
import numpy as np
import pytest
import scipy.linalg

import loop_cases as LC
import oracle_lib as O

# (tag, D, N, rep, L, beta, kappa, nmd, with_traj)
CASES = [
    ("su2_pure", 4, 2, "", 4, 2.3, 0.0, 20, True),
    ("su3_pure", 4, 3, "", 4, 5.7, 0.0, 12, True),
    ("su2_fund", 4, 2, "fund", 4, 2.3, 0.265, 20, True),
    ("su2_adj", 3, 2, "adj", 4, 2.3, 0.3, 12, True),
    ("su3_fund", 3, 3, "fund", 4, 5.7, 0.2, 12, True),
    ("su3_sym2", 4, 3, "2", 4, 5.7, 0.1, 20, True),
    ("su3_61", 4, 3, "6,1", 3, 5.7, 0.05, 16, False),
]


def rel(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    if np.iscomplexobj(a):
        a = a.view(np.float64)
    if np.iscomplexobj(b):
        b = b.view(np.float64)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def test_splitmix64_known_answers(goldens):
    g, _ = goldens
    for x, y in zip(g["rng_mix_in"], g["rng_mix_out"]):
        assert O.splitmix64(int(x)) == int(y)
    # published SplitMix64 values (Vigna): first outputs of the generator seeded at 0
    assert O.splitmix64(0) == 0xE220A8397B1DCDAF
    assert O.splitmix64(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4


def test_key_uniform_gauss_bit_exact(goldens):
    g, _ = goldens
    keys = g["rng_keys"]
    expect = []
    for t in range(4):
        expect += [O.key(0xA1, t), O.key(0xC3, t), O.key(0xD4, t), O.key(0xB2, t), O.key(0xE5, t)]
    for s in range(8):
        for mu in range(4):
            expect.append(O.key(1, s, mu, 2, 0))
    assert [int(k) for k in keys] == expect
    gk = g["rng_gauss_keys"]
    uni = np.array([O.uniform(1, int(k)) for k in gk] + [O.uniform(12345, int(k)) for k in gk])
    assert np.array_equal(uni, g["rng_uniform"])
    gau = np.concatenate([O.gauss_many(1, gk), O.gauss_many(12345, gk)])
    assert np.array_equal(gau, g["rng_gauss"])
    assert np.all((uni > 0) & (uni < 1))


@pytest.mark.parametrize("N", [2, 3])
def test_expi_and_reunitarize(goldens, N):
    g, _ = goldens
    H = g[f"expi{N}_H"].view(np.complex128).reshape(-1, N, N)
    E = g[f"expi{N}_expi"].view(np.complex128).reshape(-1, N, N)
    for h, e in zip(H, E):
        assert rel(O.expi(h), e) < 1e-14
    # reunitarize is exercised through the engine below; here check its fixed point
    Ri = g[f"expi{N}_reunit_in"].view(np.complex128).reshape(-1, N, N)
    Ro = g[f"expi{N}_reunit_out"].view(np.complex128).reshape(-1, N, N)
    for r in Ro:
        assert np.allclose(r.conj().T @ r, np.eye(N), atol=1e-14)
        assert abs(np.linalg.det(r) - 1) < 1e-14
    o = O.Oracle(4, N, 2, "", seed=1)
    U = np.tile(np.eye(N, dtype=np.complex128), (o.V * 4, 1, 1))
    U[: len(Ri)] = Ri
    o.set("U", U)
    o.reunitarize()
    assert rel(o.get("U")[: len(Ro)], Ro) < 1e-14


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_engine_matches_reference(goldens, case):
    g, sc = goldens
    tag, D, N, rep, L, beta, kappa, nmd, with_traj = case
    o = O.Oracle(D, N, L, rep, seed=1)
    o.couplings(beta=beta, kappa=kappa, lam=0.5, tau=1.0, nmd=nmd, omelyan=False)
    o.hot(0.8)
    if rep:
        o.phi_gaussian(12345, 0.3)
    assert rel(o.get("U"), g[tag + "_U0"]) < 1e-15
    if rep:
        assert np.array_equal(o.get("phi").view(np.float64).ravel(), g[tag + "_phi0"])
    F, Fp = o.forces()
    assert rel(F, g[tag + ("_Flink" if rep else "_F")]) < 1e-13
    if rep:
        assert rel(Fp, g[tag + "_Fphi"]) < 1e-13
    ob = o.observables()
    for k, key in [("Sg", "_Sg"), ("plaq", "_plaq"), ("poly", "_poly")]:
        assert abs(ob[k] - sc[tag + key]) <= 1e-13 * max(1, abs(sc[tag + key]))
    if rep:
        for k, key in [("SH", "_SH"), ("Llink", "_Llink"), ("Lphi", "_Lphi")]:
            assert abs(ob[k] - sc[tag + key]) <= 1e-13 * max(1, abs(sc[tag + key]))
        o.refresh()
        assert np.array_equal(o.get("P").ravel(), g[tag + "_P0"])
        assert np.array_equal(o.get("pi").view(np.float64).ravel(), g[tag + "_pi0"])
        assert abs(o.ham_parts()[4] - sc[tag + "_H0"]) <= 1e-12 * abs(sc[tag + "_H0"])
        o.kick(0.05)
        assert rel(o.get("P"), g[tag + "_P_kick"]) < 1e-14
        assert rel(o.get("pi"), g[tag + "_pi_kick"]) < 1e-14
        o.drift(0.05)
        assert rel(o.get("U"), g[tag + "_U_drift"]) < 1e-14
        assert rel(o.get("phi"), g[tag + "_phi_drift"]) < 1e-14
    else:
        o2 = O.Oracle(D, N, L, rep, seed=1)
        o2.refresh()
        assert np.array_equal(o2.get("P").ravel(), g[tag + "_P0"])
    if not with_traj:
        return
    for integ in ("lf", "omel"):
        o = O.Oracle(D, N, L, rep, seed=1)
        o.couplings(beta=beta, kappa=kappa, lam=0.5, tau=1.0, nmd=nmd, omelyan=(integ == "omel"))
        o.hot(0.8)
        if rep:
            o.phi_gaussian(12345, 0.3)
        ntr = 2 if rep else 3
        res = [o.trajectory() for _ in range(ntr)]
        dH = np.array([r[1] for r in res])
        acc = np.array([float(r[0]) for r in res])
        ref_dH = g[f"{tag}_{integ}_dH"]
        # |H| ~ 1e3: the summation differs by a few ulps of H
        assert np.max(np.abs(dH - ref_dH)) < 1e-10
        assert np.array_equal(acc, g[f"{tag}_{integ}_acc"])
        assert rel(o.get("U"), g[f"{tag}_{integ}_U{ntr}"]) < 1e-13
        if rep:
            assert rel(o.get("phi"), g[f"{tag}_{integ}_phi2"]) < 1e-13


@pytest.mark.parametrize("tag,N,spec", [("rep_su3_2", 3, "2"), ("rep_su3_11", 3, "1,1"), ("rep_su2_3", 2, "3"), ("rep_su3_61", 3, "6,1")])
def test_oracle_rep_matches_reference(goldens, tag, N, spec):
    g, sc = goldens
    o = O.Oracle(4, N, 2, spec, seed=1)
    assert o.d == int(sc[tag + "_d"]) and o.dimT == int(sc[tag + "_dimT"]) and o.nality == int(sc[tag + "_nality"])
    assert sc[tag + "_W_max_imag"] == 0.0
    assert np.max(np.abs(o.rep_W().ravel() - g[tag + "_W"])) < 1e-15
    assert rel(o.rep_T(), g[tag + "_T"]) < 1e-14
    U = g[tag + "_app_U"].view(np.complex128).reshape(-1, N, N)
    px = g[tag + "_app_px"].view(np.complex128).reshape(len(U), -1)
    py = g[tag + "_app_py"].view(np.complex128).reshape(len(U), -1)
    rot = g[tag + "_app_rot"].view(np.complex128).reshape(len(U), -1)
    rotd = g[tag + "_app_rotdag"].view(np.complex128).reshape(len(U), -1)
    hop = g[tag + "_app_hop"].reshape(len(U), -1)
    for t in range(len(U)):
        assert rel(o.rep_apply(U[t], py[t]), rot[t]) < 1e-13
        assert rel(o.rep_apply(U[t], py[t], dag=True), rotd[t]) < 1e-13
        assert rel(o.rep_hop_g(U[t], px[t], py[t]), hop[t]) < 1e-12


FROZEN = [("su2_fund_frozen", 4, 2, "fund", 2.3, 0.6, 20, 4), ("su2_adj_frozen", 3, 2, "adj", 2.3, 0.5, 12, 4),
          ("su2_4_frozen", 3, 2, "4", 2.3, 0.5, 10, 3), ("su3_21_frozen", 3, 3, "2,1", 5.7, 0.3, 10, 3)]


@pytest.mark.parametrize("tag,D,N,rep,beta,kappa,nmd,L", FROZEN, ids=[c[0] for c in FROZEN])
def test_frozen_scalar_matches_reference(goldens, tag, D, N, rep, beta, kappa, nmd, L):
    """|phi_x| = 1 scalar (gauge_higgs_hmc.hpp:61-103, 133, 144, 172): normalisation,
    tangent-projected momenta and kicks, geodesic drift, trajectories."""
    g, sc = goldens
    o = O.Oracle(D, N, L, rep, seed=1)
    o.couplings(beta=beta, kappa=kappa, lam=0.5, tau=1.0, nmd=nmd, omelyan=False)
    o.hot(0.8)
    o.phi_gaussian(12345, 0.3)
    o.frozen(True)
    o.normalize_phi()
    assert rel(o.get("phi"), g[tag + "_phi0"]) < 1e-15
    assert np.allclose(np.linalg.norm(o.get("phi"), axis=1), 1.0, atol=1e-14)
    o.refresh()
    assert rel(o.get("pi"), g[tag + "_pi0"]) < 1e-15
    o.kick(0.05)
    assert rel(o.get("pi"), g[tag + "_pi_kick"]) < 1e-14
    o.drift(0.05)
    assert rel(o.get("phi"), g[tag + "_phi_drift"]) < 1e-14
    assert rel(o.get("pi"), g[tag + "_pi_drift"]) < 1e-14
    for integ in ("lf", "omel"):
        o = O.Oracle(D, N, L, rep, seed=1)
        o.couplings(beta=beta, kappa=kappa, lam=0.5, tau=1.0, nmd=nmd, omelyan=(integ == "omel"))
        o.hot(0.8)
        o.phi_gaussian(12345, 0.3)
        o.frozen(True)
        o.normalize_phi()
        res = [o.trajectory() for _ in range(2)]
        assert np.max(np.abs(np.array([r[1] for r in res]) - g[f"{tag}_{integ}_dH"])) < 1e-10
        assert np.array_equal(np.array([float(r[0]) for r in res]), g[f"{tag}_{integ}_acc"])
        assert rel(o.get("U"), g[f"{tag}_{integ}_U2"]) < 1e-13
        assert rel(o.get("phi"), g[f"{tag}_{integ}_phi2"]) < 1e-13


# (tag, D, N, rep, beta, kappa, mu2, nmd, L, s): s scales the tolerances. At d = 13
# the degree-12 Lagrange projectors lose ~1e-10 relative to roundoff, in the
# reference as here (scalar_invariants.hpp:96-99 quotes ~2e-8 for its cache).
MULTI = [("multi_su2_fund", 4, 2, "fund", 2.3, 0.2, 0.4, 12, 4, 1.0), ("multi_su2_adj", 3, 2, "adj", 2.3, 0.2, 0.4, 12, 4, 1.0),
         ("multi_su3_2", 4, 3, "2", 5.7, 0.1, 0.5, 12, 4, 1.0), ("multi_su3_21", 3, 3, "2,1", 5.7, 0.1, 0.5, 10, 3, 1.0),
         ("multi_su2_12", 3, 2, "12", 2.3, 0.1, 0.4, 10, 3, 1e3)]


@pytest.mark.parametrize("case", MULTI, ids=[c[0] for c in MULTI])
def test_multi_invariant_potential_matches_reference(goldens, case):
    """Multi-invariant quartic potential (scalar_invariants.hpp:20-208) plugged
    into the HMC: channel Casimirs, V and dV/dphi^* on random phi, the scalar
    force and action, and trajectories of both integrators."""
    g, sc = goldens
    tag, D, N, rep, beta, kappa, mu2, nmd, L, s = case
    o = O.Oracle(D, N, L, rep, seed=1)
    lam = o.channels()
    assert len(lam) == sc[tag + "_nch"]
    assert np.max(np.abs(lam - g[tag + "_lambda"])) < 1e-9 * s
    f = g[tag + "_f"]
    o.multi_potential(f, mu2)
    px = g[tag + "_px"].view(np.complex128).reshape(3, -1)
    dv = g[tag + "_dV"].view(np.complex128).reshape(3, -1)
    dvu = g[tag + "_dV_uncached"].view(np.complex128).reshape(3, -1)
    for t in range(3):
        V, dV = o.multi_eval(px[t])
        assert abs(V - g[tag + "_V"][t]) < 1e-12 * s * max(1.0, abs(V))
        assert np.max(np.abs(dV - dvu[t])) < 1e-12 * s * np.max(np.abs(dvu[t]))
        assert np.max(np.abs(dV - dv[t])) < 1e-10 * s * np.max(np.abs(dv[t]))  # reference's superoperator cache: roundoff-level
    o.couplings(beta=beta, kappa=kappa, lam=0.5, tau=1.0, nmd=nmd, omelyan=False)
    o.hot(0.8)
    o.phi_gaussian(12345, 0.3)
    F, Fp = o.forces()
    assert rel(Fp, g[tag + "_Fphi"]) < 1e-10 * s
    assert abs(o.observables()["SH"] - sc[tag + "_SH"]) < 1e-10 * s * max(1, abs(sc[tag + "_SH"]))
    for integ in ("lf", "omel"):
        o = O.Oracle(D, N, L, rep, seed=1)
        o.multi_potential(f, mu2)
        o.couplings(beta=beta, kappa=kappa, lam=0.5, tau=1.0, nmd=nmd, omelyan=(integ == "omel"))
        o.hot(0.8)
        o.phi_gaussian(12345, 0.3)
        res = [o.trajectory() for _ in range(2)]
        assert np.max(np.abs(np.array([r[1] for r in res]) - g[f"{tag}_{integ}_dH"])) < 1e-8 * max(1.0, s / 100)
        assert np.array_equal(np.array([float(r[0]) for r in res]), g[f"{tag}_{integ}_acc"])
        assert rel(o.get("phi"), g[f"{tag}_{integ}_phi2"]) < 1e-11 * s


@pytest.mark.parametrize("case", LC.CASES, ids=[c[0] for c in LC.CASES])
def test_wilson_and_polyakov_loops_match_reference(goldens, case):
    """Measurement layer (observables.hpp:36-101): the rectangle path product,
    wilson_loop_fund, wilson_loop_rep and polyakov_loop_rep in every probe irrep
    on the reference's thermalised field. The d = 48 probe {6,1} costs ~15 s per
    Wilson-loop size in the tensor restatement, so only its Polyakov loop runs
    here; the GPU test covers its Wilson loops against the same fixtures."""
    g, sc = goldens
    tag, D, N, ext, labels = case
    o = O.Oracle(D, N, list(ext))
    o.set("U", g[tag + "_U"].view(np.complex128).reshape(-1, N, N))
    r = o.wilson_rect(5, 0, D - 1, 2, 3)
    assert np.max(np.abs(r.ravel() - g[tag + "_rect"].view(np.complex128))) < 1e-14
    for Ra, Rb in LC.SIZES:
        assert abs(o.wilson_loop("fund", Ra, Rb) - sc[f"{tag}_Wfund_{Ra}x{Rb}"]) < 1e-14
    for lab in labels:
        spec = LC.SPEC[lab]
        if lab != "61":
            for Ra, Rb in LC.sizes_for(lab):
                assert abs(o.wilson_loop(spec, Ra, Rb) - sc[f"{tag}_W_{lab}_{Ra}x{Rb}"]) < 1e-14, (lab, Ra, Rb)
        for t in range(D) if lab != "61" else [D - 1]:
            assert abs(o.polyakov_rep(spec, t) - sc[f"{tag}_P_{lab}_t{t}"]) < 1e-14, (lab, t)


def test_loops_closed_forms():
    """What the mathematics fixes: on the cold lattice every loop in every probe
    is 1; the 1x1 fundamental loop is the average plaquette; a random gauge
    transformation leaves the loops unchanged; the conjugate probe {1,1} of
    SU(3) gives the same real part as the fundamental."""
    o = O.Oracle(4, 3, [3, 4, 3, 4])
    o.cold()
    for spec in ("fund", "adj", "2", "2,1", "3"):
        assert abs(o.wilson_loop(spec, 2, 1) - 1.0) < 1e-14
        assert abs(o.polyakov_rep(spec, 1) - 1.0) < 1e-14
    o.hot(0.4)
    assert abs(o.wilson_loop("fund", 1, 1) - o.observables()["plaq"]) < 1e-14
    U = o.get("U")
    before = [o.wilson_loop(s, 2, 2) for s in ("fund", "adj", "2")] + [o.polyakov_rep("2,1", 0)]
    assert abs(o.wilson_loop("1,1", 2, 1) - o.wilson_loop("fund", 2, 1)) < 1e-14
    # gauge transformation U_mu(x) -> G(x) U_mu(x) G(x+mu)^dag with random SU(3) G
    rng = np.random.default_rng(5)
    ext = [3, 4, 3, 4]
    V = int(np.prod(ext))
    Gs = []
    for _ in range(V):
        H = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        H = 0.5 * (H + H.conj().T)
        Gs.append(scipy.linalg.expm(1j * H))
    strides = np.cumprod([1] + ext[:-1])
    Ut = U.reshape(V, 4, 3, 3).copy()
    for s in range(V):
        x = [(s // strides[m]) % ext[m] for m in range(4)]
        for mu in range(4):
            y = list(x)
            y[mu] = (y[mu] + 1) % ext[mu]
            sy = sum(y[m] * strides[m] for m in range(4))
            Ut[s, mu] = Gs[s] @ U.reshape(V, 4, 3, 3)[s, mu] @ Gs[sy].conj().T
    o.set("U", Ut.reshape(-1, 3, 3))
    after = [o.wilson_loop(s, 2, 2) for s in ("fund", "adj", "2")] + [o.polyakov_rep("2,1", 0)]
    assert np.max(np.abs(np.array(before) - np.array(after))) < 1e-13
