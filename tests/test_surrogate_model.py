"""CPU checks of the numpy surrogate model the MFMA kernels are tested against:
bfloat16 rounding and the analytic gradient (finite differences)."""
import numpy as np

from surrogate_model import bf16, model


def test_bf16_rounding():
    x = np.array([1.0, -2.5, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -9, 3.0e38, 1e-30], dtype=np.float32)
    r = bf16(x)
    # ties go to even mantissa; everything else to nearest
    assert r[0] == 1.0 and r[1] == -2.5
    assert r[2] == 1.0 and r[3] == 1.0 + 4 * 2.0 ** -8
    assert r[4] == 1.0
    assert abs(r[5] - 3.0e38) < 3.0e38 * 2.0 ** -8
    assert abs(r[6] - 1e-30) < 1e-30 * 2.0 ** -8


def test_numpy_gradient_by_finite_differences():
    rng = np.random.default_rng(0)
    P = rng.normal(size=6288) * 0.2
    P[5248:6272].reshape(64, 16)[:, 4:] = 0.0
    X, B, F = rng.normal(size=(32, 16)), rng.normal(size=(32, 4, 3)), rng.normal(size=(32, 3))
    ident = lambda a: a
    _, l0, grad = model(P, X, B, F, ident)
    for p in [3, 700, 1030, 2000, 5000, 5200, 5250, 5251 + 16 * 7, 6272, 6275]:
        e = np.zeros_like(P)
        e[p] = 1e-6
        fd = (model(P + e, X, B, F, ident)[1] - model(P - e, X, B, F, ident)[1]) / 2e-6
        assert abs(fd - grad[p]) < 1e-6 * max(1.0, abs(fd)), p


def test_basis_bound_limits():
    """The per-link projection bound: 0 when every force lies in the span of
    its link's bases, 1 when it is orthogonal to them, and between the two
    (never above the global-coefficient residual) otherwise."""
    from gauge_higgs_amd.learned import basis_bound
    rng = np.random.default_rng(3)
    B = rng.normal(size=(500, 4, 8))
    c = rng.normal(size=(500, 4))
    F_in = np.einsum("lk,lka->la", c, B)
    r = basis_bound(B, F_in)
    assert r["B0123"]["per_link_optimum"] < 1e-20
    assert r["B0"]["per_link_optimum"] > 0.1
    # orthogonal complement of each link's span
    F_out = rng.normal(size=(500, 8))
    for l in range(500):
        q, _ = np.linalg.qr(B[l].T)
        F_out[l] -= q @ (q.T @ F_out[l])
    r = basis_bound(B, F_out)
    assert abs(r["B0123"]["per_link_optimum"] - 1.0) < 1e-12
    F = F_in + F_out
    r = basis_bound(B, F)
    for v in r.values():
        assert 0.0 <= v["per_link_optimum"] <= v["global_coefficients"] + 1e-12
