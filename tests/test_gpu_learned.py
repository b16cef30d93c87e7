"""Multi-timescale integrator and the learned surrogate of the matter link force.

Pins: the nested schedule with one inner step reduces to the reference's
leapfrog (gauge_higgs_hmc.hpp:149-165) up to rounding; it is reversible and
second order (dH falls ~4x when the outer step halves). The surrogate features
are gauge invariant (checked under a random gauge transformation on the host)
and satisfy the sum rules of the plaquette and the charge norms; its MFMA
forward and backward agree with a numpy model that rounds the same operands to
bfloat16, and that numpy gradient is itself checked by finite differences.
Whatever the surrogate weights, the split K_in(F_gauge + S) K_out(F_hop - S)
keeps the exact total force, so HMC stays exact: checked by reversibility.
"""
import numpy as np
import pytest

import gauge_higgs_amd as gha
from surrogate_model import bf16, model

pytestmark = pytest.mark.gpu


def make(L=4, rep="fund", N=2, kappa=0.265, prec=gha.FP64, nmd=8, beta=2.3, phi_sigma=0.3):
    g = gha.GaugeHiggsHMC(L, rep, seed=3, D=4, N=N, precision=prec)
    g.configure(beta=beta, kappa=kappa, lam=0.5, nmd=nmd, integ=gha.LEAPFROG, reunit_each_traj=False)
    g.hot(0.8, 12345, phi_sigma)
    return g


# ------------------------------------------------------------------ nested integrator
def test_nested_m1_is_leapfrog():
    a, b = make(), make()
    a.refresh_momenta()
    b.set("P", a.get("P"))
    b.set("pi", a.get("pi"))
    b.lat.set_nested(8, 1)
    a.md_evolve()
    b.md_evolve()
    for k in ("U", "P", "phi", "pi"):
        assert np.max(np.abs(a.get(k) - b.get(k))) < 1e-11, k


def test_nested_reversible():
    g = make()
    g.lat.set_nested(4, 3)
    g.refresh_momenta()
    U0, phi0 = g.get("U"), g.get("phi")
    g.md_evolve()
    g.set("P", -g.get("P"))
    g.set("pi", -g.get("pi"))
    g.md_evolve()
    assert np.max(np.abs(g.get("U") - U0)) < 1e-10
    assert np.max(np.abs(g.get("phi") - phi0)) < 1e-10


def test_nested_second_order():
    dh = []
    for n in (8, 16):
        g = make()
        g.lat.set_nested(n, 2)
        g.refresh_momenta()
        h0 = g.hamiltonian()
        g.md_evolve()
        dh.append(abs(g.hamiltonian() - h0))
    assert 2.5 < dh[0] / dh[1] < 6.0, dh


def test_nested_bigrep_reversible():
    g = gha.GaugeHiggsHMC(3, "6,1", seed=1, D=4, N=3).configure(beta=5.7, kappa=0.05, lam=0.5, nmd=4, reunit_each_traj=False)
    g.hot(0.8, 12345, 0.3)
    g.lat.set_nested(2, 3)
    g.refresh_momenta()
    U0, phi0 = g.get("U"), g.get("phi")
    g.md_evolve()
    g.set("P", -g.get("P"))
    g.set("pi", -g.get("pi"))
    g.md_evolve()
    assert np.max(np.abs(g.get("U") - U0)) < 1e-10
    assert np.max(np.abs(g.get("phi") - phi0)) < 1e-10


# ------------------------------------------------------------------ surrogate: features
def su2_random(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    a, b = q[:, 0] + 1j * q[:, 3], q[:, 2] + 1j * q[:, 1]
    m = np.zeros((n, 2, 2), complex)
    m[:, 0, 0], m[:, 0, 1], m[:, 1, 0], m[:, 1, 1] = a, b, -b.conj(), a.conj()
    return m


def neighbours(L, D=4):
    """Forward neighbour table of the reference geometry (direction 0 fastest)."""
    V = L ** D
    s = np.arange(V)
    x = np.stack([(s // L ** mu) % L for mu in range(D)], 1)
    up = np.zeros((V, D), int)
    for mu in range(D):
        y = x.copy()
        y[:, mu] = (y[:, mu] + 1) % L
        up[:, mu] = sum(y[:, nu] * L ** nu for nu in range(D))
    return up


def snapshot(g):
    X, B, S, F = g.lat.surrogate_snapshot()
    return X, B, S, F


def test_features_gauge_invariant():
    L = 4
    g = make(L=L)
    g.lat.surrogate_init()
    X0, B0, S0, F0 = snapshot(g)
    U, phi = g.get("U").reshape(L ** 4, 4, 2, 2), g.get("phi")
    rng = np.random.default_rng(5)
    Om = su2_random(rng, L ** 4)
    up = neighbours(L)
    Ut = np.einsum("sij,smjk,smlk->smil", Om, U, Om[up].conj())
    g.set("U", Ut.reshape(-1, 2, 2))
    g.set("phi", np.einsum("sij,sj->si", Om, phi))
    X1, B1, S1, F1 = snapshot(g)
    # features are computed in FP32: invariant up to FP32 rounding
    assert np.max(np.abs(X1 - X0)) < 2e-5 * max(1.0, np.max(np.abs(X0)))
    # covariant quantities: the adjoint action preserves norms per link
    for k in range(4):
        n0, n1 = np.linalg.norm(B0[:, k], axis=1), np.linalg.norm(B1[:, k], axis=1)
        assert np.max(np.abs(n1 - n0)) < 2e-5 * max(1e-3, np.max(n0)), k
    s0, s1 = np.linalg.norm(S0, axis=1), np.linalg.norm(S1, axis=1)
    assert np.max(np.abs(s1 - s0)) < 2e-2 * np.max(s0)
    assert np.allclose(np.linalg.norm(F1, axis=1), np.linalg.norm(F0, axis=1), atol=1e-12, rtol=1e-9)


def test_feature_sum_rules():
    L = 4
    g = make(L=L)
    g.lat.surrogate_init()
    X, B, S, F = snapshot(g)
    ob = g.observables()
    V = L ** 4
    # every plaquette is seen by its four links: mean feature = plaquette
    assert abs(X[:, :6].mean() - ob["plaq"]) < 1e-6
    phi = g.get("phi")
    n2 = np.sum(np.abs(phi) ** 2, axis=1)
    up = neighbours(L)
    assert np.allclose(X[:, 6].reshape(V, 4), n2[:, None], rtol=1e-6)
    assert np.allclose(X[:, 7].reshape(V, 4), n2[up], rtol=1e-6)
    # B1 = Q_x = phi^dag T^a phi with T^a = sigma^a / 2 in the fundamental
    sig = np.array([[[0, 1], [1, 0]], [[0, -1j], [1j, 0]], [[1, 0], [0, -1]]]) / 2
    Q = np.einsum("si,aij,sj->sa", phi.conj(), sig, phi).real
    Bq = B.reshape(V, 4, 4, 3)
    assert np.allclose(Bq[:, :, 1, :], Q[:, None, :], atol=1e-6)
    # Ad(U) is orthogonal; B3 = f B1 B2 is orthogonal to both
    assert np.allclose(np.linalg.norm(Bq[:, :, 2, :], axis=2), np.linalg.norm(Q[up], axis=2), atol=1e-6)
    assert np.max(np.abs(np.einsum("smk,smk->sm", Bq[:, :, 3], Bq[:, :, 1]))) < 1e-6
    assert np.max(np.abs(np.einsum("smk,smk->sm", Bq[:, :, 3], Bq[:, :, 2]))) < 1e-6
    # B0 is the gauge force: total force - matter link force
    Ft, _ = g.forces()
    assert np.max(np.abs(B[:, 0, :] - (Ft - F))) < 1e-5 * np.max(np.abs(Ft))


# ------------------------------------------------------------------ surrogate: MLP on MFMA
@pytest.mark.parametrize("prec", [gha.FP64, gha.FP32MD])
def test_mfma_forward_and_gradient(prec):
    g = make(prec=prec)
    g.lat.surrogate_init(seed=11)
    P = g.lat.surrogate_params()
    rng = np.random.default_rng(1)
    P = (P + rng.normal(size=P.size) * 0.05).astype(np.float32)  # make every layer matter
    P[5248:6272].reshape(64, 16)[:, 4:] = 0.0
    P[6276:] = 0.0
    g.lat.surrogate_params(P)
    X, B, S, F = snapshot(g)
    grad, loss = g.lat.surrogate_gradient()
    F32 = F.astype(np.float32).astype(np.float64)
    Sn, ln, gn = model(P, X, B, F32, bf16)
    assert np.linalg.norm(S - Sn) < 2e-3 * np.linalg.norm(Sn)
    assert abs(loss - ln) < 2e-3 * ln
    assert np.linalg.norm(grad - gn) < 1e-2 * np.linalg.norm(gn)
    # each parameter block separately (a transposed operand would fail here)
    for lo, hi in [(0, 1024), (1024, 1088), (1088, 5184), (5184, 5248), (5248, 6272), (6272, 6288)]:
        assert np.linalg.norm(grad[lo:hi] - gn[lo:hi]) < 2e-2 * np.linalg.norm(gn[lo:hi]) + 1e-9, (lo, hi)
    # the bf16 rounding is a small perturbation of the FP32 gradient
    _, _, g32 = model(P, X, B, F32, lambda a: a)
    assert np.linalg.norm(grad - g32) < 0.1 * np.linalg.norm(g32)


def test_training_lowers_loss():
    g = make()
    g.lat.surrogate_init(seed=2)
    l0, l1 = g.lat.surrogate_train(iters=300, lr=3e-3)
    assert l1 < 0.9 * l0, (l0, l1)


def test_surrogate_hmc_is_exact_split():
    # trained or not, K_in(F_gauge + S) K_out(F_hop - S) integrates the exact force
    g = make()
    g.lat.surrogate_init(seed=2)
    g.lat.surrogate_train(iters=50, lr=3e-3)
    g.lat.surrogate_enable(True)
    g.lat.set_nested(4, 3)
    g.refresh_momenta()
    U0, phi0 = g.get("U"), g.get("phi")
    g.md_evolve()
    g.set("P", -g.get("P"))
    g.set("pi", -g.get("pi"))
    g.md_evolve()
    assert np.max(np.abs(g.get("U") - U0)) < 1e-9
    assert np.max(np.abs(g.get("phi") - phi0)) < 1e-9
    # with m = 1 and S switched on, the split is again exactly the leapfrog up to rounding
    a, b = make(), make()
    b.lat.surrogate_init(seed=2)
    b.lat.surrogate_train(iters=20, lr=3e-3)
    b.lat.surrogate_enable(True)
    b.lat.set_nested(8, 1)
    a.refresh_momenta()
    b.set("P", a.get("P"))
    b.set("pi", a.get("pi"))
    a.md_evolve()
    b.md_evolve()
    assert np.max(np.abs(a.get("U") - b.get("U"))) < 1e-9
    assert np.max(np.abs(a.get("phi") - b.get("phi"))) < 1e-9
    # and dH -> 0 at second order: the integrated force is the exact one
    dh = []
    for n in (8, 16):
        c = make()
        c.lat.surrogate_init(seed=2)
        c.lat.surrogate_train(iters=20, lr=3e-3)
        c.lat.surrogate_enable(True)
        c.lat.set_nested(n, 2)
        c.refresh_momenta()
        h0 = c.hamiltonian()
        c.md_evolve()
        dh.append(abs(c.hamiltonian() - h0))
    assert 2.5 < dh[0] / dh[1] < 6.0, dh


def test_surrogate_trajectories_bigrep():
    g = gha.GaugeHiggsHMC(4, "6,1", seed=1, D=4, N=3, precision=gha.FP32MD).configure(beta=5.7, kappa=0.05, lam=0.5, nmd=16)
    g.hot(0.8, 12345, 0.3)
    g.lat.surrogate_init(seed=2)
    l0, l1 = g.lat.surrogate_train(iters=100, lr=3e-3)
    assert np.isfinite(l1) and l1 < l0
    g.lat.surrogate_enable(True)
    g.lat.set_nested(4, 4)
    g.lat.trajectories(4, use_graph=True)
    t, a, dh = g.lat.counters()
    assert t == 4 and np.isfinite(dh)


def test_nested_omelyan_outer():
    # Omelyan on the outer scale: reversible and second order
    g = make()
    g.configure(integ=gha.OMELYAN)
    g.lat.set_nested(3, 2)
    g.refresh_momenta()
    U0, phi0 = g.get("U"), g.get("phi")
    g.md_evolve()
    g.set("P", -g.get("P"))
    g.set("pi", -g.get("pi"))
    g.md_evolve()
    assert np.max(np.abs(g.get("U") - U0)) < 1e-10
    assert np.max(np.abs(g.get("phi") - phi0)) < 1e-10
    dh = []
    for n in (6, 12):
        c = make()
        c.configure(integ=gha.OMELYAN)
        c.lat.set_nested(n, 2)
        c.refresh_momenta()
        h0 = c.hamiltonian()
        c.md_evolve()
        dh.append(abs(c.hamiltonian() - h0))
    assert 2.5 < dh[0] / dh[1] < 6.0, dh


@pytest.mark.parametrize("inner", [gha.LEAPFROG, gha.OMELYAN])
def test_nested_inner_schedule_at_kappa0(inner):
    # at kappa = 0 the hopping forces vanish and the outer kicks do nothing:
    # n outer steps of m inner ones are then the single-scale integrator with
    # n*m (leapfrog outer) or 2*n*m (Omelyan outer) steps, up to the rounding
    # of merged kicks. Pins the inner step coefficients of both schemes.
    n, m = 2, 3
    a, b = make(kappa=0.0), make(kappa=0.0)
    for g in (a, b):
        g.configure(integ=inner)
    a.configure(nmd=n * m * (2 if inner == gha.OMELYAN else 1))
    b.lat.set_nested(n, m, inner=inner)
    a.refresh_momenta()
    b.set("P", a.get("P"))
    b.set("pi", a.get("pi"))
    a.md_evolve()
    b.md_evolve()
    for k in ("U", "P", "phi", "pi"):
        assert np.max(np.abs(a.get(k) - b.get(k))) < 1e-11, k


@pytest.mark.parametrize("fuse", [0, 1])
def test_nested_omelyan_inner_bigrep_at_kappa0(fuse):
    # the same pin for the {6,1} irrep, whose inner steps run the on-site
    # kernel (fused with the gauge update or not)
    def mk():
        g = gha.GaugeHiggsHMC(4, "6,1", seed=1, D=4, N=3).configure(beta=5.7, kappa=0.0, lam=0.5, nmd=8, integ=gha.OMELYAN,
                                                                    reunit_each_traj=False)
        g.hot(0.8, 12345, 0.3)
        g.lat.set_link_lanes(1)
        g.lat.set_fuse_inner(fuse)
        return g
    a, b = mk(), mk()
    b.lat.set_nested(2, 2, inner=gha.OMELYAN)
    a.refresh_momenta()
    b.set("P", a.get("P"))
    b.set("pi", a.get("pi"))
    a.md_evolve()
    b.md_evolve()
    for k in ("U", "P", "phi", "pi"):
        assert np.max(np.abs(a.get(k) - b.get(k))) < 1e-10, k


def test_nested_omelyan_inner_reversible_second_order():
    g = make()
    g.configure(integ=gha.OMELYAN)
    g.lat.set_nested(3, 2, inner=gha.OMELYAN)
    g.refresh_momenta()
    U0, phi0 = g.get("U"), g.get("phi")
    g.md_evolve()
    g.set("P", -g.get("P"))
    g.set("pi", -g.get("pi"))
    g.md_evolve()
    assert np.max(np.abs(g.get("U") - U0)) < 1e-10
    assert np.max(np.abs(g.get("phi") - phi0)) < 1e-10
    dh = []
    for n in (6, 12):
        c = make()
        c.configure(integ=gha.OMELYAN)
        c.lat.set_nested(n, 2, inner=gha.OMELYAN)
        c.refresh_momenta()
        h0 = c.hamiltonian()
        c.md_evolve()
        dh.append(abs(c.hamiltonian() - h0))
    assert 2.5 < dh[0] / dh[1] < 6.0, dh


def test_nested_m1_is_leapfrog_bigrep():
    # inner steps of a large irrep run the per-site on-site kernel; with one inner
    # step the split must still be the reference leapfrog up to rounding
    def mk():
        g = gha.GaugeHiggsHMC(3, "6,1", seed=1, D=4, N=3).configure(beta=5.7, kappa=0.05, lam=0.5, nmd=6, integ=gha.LEAPFROG,
                                                                    reunit_each_traj=False)
        g.hot(0.8, 12345, 0.3)
        return g
    a, b = mk(), mk()
    a.refresh_momenta()
    b.set("P", a.get("P"))
    b.set("pi", a.get("pi"))
    b.lat.set_nested(6, 1)
    a.md_evolve()
    b.md_evolve()
    for k in ("U", "P", "phi", "pi"):
        assert np.max(np.abs(a.get(k) - b.get(k))) < 1e-11, k


@pytest.mark.parametrize("prec", [gha.FP64, gha.FP32MD])
@pytest.mark.parametrize("surr", [False, True])
def test_fused_inner_step_equals_two_launches(prec, surr):
    """The inner scale of a large irrep ({6,1}) under the multi-timescale
    integrator, with the links link-parallel: the on-site and gauge updates in
    one launch (md_link_onsite_kernel) give the fields of the two separate
    launches bit for bit, with and without the surrogate force."""
    out = []
    for fuse in (0, 1):
        g = gha.GaugeHiggsHMC(4, "6,1", seed=1, D=4, N=3, precision=prec).configure(beta=5.7, kappa=0.05, lam=0.5, nmd=4,
                                                                                    reunit_each_traj=False)
        g.hot(0.8, 12345, 0.3)
        g.lat.set_link_lanes(1)
        g.lat.set_link_split(0)  # the fused launch runs one lane per link; the plane split rounds the staple sum otherwise
        g.lat.set_fuse_inner(fuse)
        g.lat.set_nested(3, 4)
        assert g.lat.fuse_inner() == fuse
        if surr:
            g.lat.surrogate_init(seed=2)
            g.lat.surrogate_enable(True)
        g.refresh_momenta()
        g.md_evolve()
        out.append({k: g.get(k) for k in ("U", "P", "phi", "pi")})
    for k in out[0]:
        assert np.array_equal(out[0][k], out[1][k]), k


def test_fused_inner_step_reported_only_where_it_runs():
    g = gha.GaugeHiggsHMC(4, "6,1", seed=1, D=4, N=3, precision=gha.FP32MD)
    assert g.lat.fuse_inner() == 0  # one thread per site: two launches
    g.lat.set_link_lanes(1)
    assert g.lat.fuse_inner() == 1
    h = gha.GaugeHiggsHMC(4, "2", seed=1, D=4, N=3, precision=gha.FP32MD)
    h.lat.set_link_lanes(1)
    assert h.lat.fuse_inner() == 0  # small irreps fuse the whole step already


@pytest.mark.parametrize("prec", [gha.FP64, gha.FP32MD])
@pytest.mark.parametrize("inner", [gha.LEAPFROG, gha.OMELYAN])
@pytest.mark.parametrize("lanes", [0, 1])
def test_split_inner_scale_equals_interleaved(prec, inner, lanes):
    """A large irrep ({6,1}) under the multi-timescale integrator: all on-site
    sub-steps of an inner block in one launch (the site's components in
    registers) plus a gauge-only launch per sub-step give the fields of the
    interleaved launches bit for bit, against both interleaved forms (one thread
    per site: two launches per sub-step; link lanes: one fused launch); the
    same holds for whole trajectories replayed from a captured graph."""
    out = []
    for split in (0, 1):
        g = gha.GaugeHiggsHMC(4, "6,1", seed=1, D=4, N=3, precision=prec).configure(beta=5.7, kappa=0.05, lam=0.5, nmd=4,
                                                                                    integ=gha.OMELYAN, reunit_each_traj=False)
        g.hot(0.8, 12345, 0.3)
        g.lat.set_link_lanes(lanes)
        g.lat.set_link_split(0)  # one lane per link on both sides, as in the fused launch
        g.lat.set_split_inner(split)
        g.lat.set_nested(3, 2, inner=inner)
        assert g.lat.split_inner() == split
        g.refresh_momenta()
        g.md_evolve()
        fields = {k: g.get(k) for k in ("U", "P", "phi", "pi")}
        g.lat.trajectories(2, use_graph=True)
        out.append((fields, g.lat.counters(), g.get("U"), g.get("phi")))
    a, b = out
    for k in a[0]:
        assert np.array_equal(a[0][k], b[0][k]), k
    assert a[1] == b[1]
    assert np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])


def test_split_inner_scale_not_with_surrogate():
    """With the surrogate force on, the inner links' kick depends on the scalar
    at every sub-step, so the inner scale stays interleaved (reported 0) and
    the switch changes nothing."""
    out = []
    for split in (0, 1):
        g = gha.GaugeHiggsHMC(4, "6,1", seed=1, D=4, N=3, precision=gha.FP32MD).configure(beta=5.7, kappa=0.05, lam=0.5, nmd=4,
                                                                                          reunit_each_traj=False)
        g.hot(0.8, 12345, 0.3)
        g.lat.set_split_inner(split)
        g.lat.set_nested(2, 3, inner=gha.OMELYAN)
        g.lat.surrogate_init(seed=2)
        g.lat.surrogate_enable(True)
        assert g.lat.split_inner() == 0
        g.refresh_momenta()
        g.md_evolve()
        out.append({k: g.get(k) for k in ("U", "P", "phi", "pi")})
    for k in out[0]:
        assert np.array_equal(out[0][k], out[1][k]), k
    h = gha.GaugeHiggsHMC(4, "2", seed=1, D=4, N=3, precision=gha.FP32MD)
    assert h.lat.split_inner() == 0  # small irreps update the scalar inside the step kernel
