"""The device's Gaussian stream against the reference's (rng.hpp:42-46).

The device draws g = sqrt(-2 log u1) cos(2 pi u2) with correctly rounded log and
cos (gauss_cr.hpp), the reference with host glibc's. Integer keys and the
uniforms are exact on both sides, so every draw is equal except where glibc
misrounds log or cos; there the device value is the correctly rounded draw
(quad precision, rounded to double). Checked here:

- 10^8 consecutive draws of two seeds: every device draw equals the correctly
  rounded draw (all of the first 2*10^5, and each one that differs from
  glibc's), and the fraction that differs from glibc is stated;
- the start fields and momenta of the reference's own dumps (P0, pi0, phi0 of
  seven lattices, tests/golden): equal, apart from glibc misroundings listed
  one by one.
"""
import numpy as np
import pytest

import gauge_higgs_amd as gha
import oracle_lib as O
from gauss_policy import TAG_P, TAG_P_PURE, TAG_PI, assert_equal_but_glibc_misrounds, momentum_keys, scalar_keys

pytestmark = pytest.mark.gpu

CHUNK = 10_000_000


@pytest.mark.parametrize("seed", [12345, 1])
def test_gaussian_stream_1e8_draws(seed):
    total, differ = 0, 0
    for c in range(10):
        k0 = c * CHUNK
        dev = gha.rng_gauss(seed, k0, CHUNK)
        glibc, _ = O.gauss_range(seed, k0, CHUNK)
        idx = np.nonzero(dev != glibc)[0]
        q = O.gauss_quad_many(seed, (k0 + idx).astype(np.uint64))
        assert np.array_equal(dev[idx], q), f"chunk {c}: a device draw differs from both glibc and the correctly rounded draw"
        if c == 0:
            _, q0 = O.gauss_range(seed, 0, 200_000, glibc=False, quad=True)
            assert np.array_equal(dev[:200_000], q0)
        total += CHUNK
        differ += idx.size
    rate = differ / total
    print(f"seed {seed}: {differ} of {total} draws differ from glibc (rate {rate:.3e}); all are glibc misroundings")
    assert 1e-4 < rate < 5e-3


GOLDEN = [  # tag, D, N, rep, L
    ("su2_pure", 4, 2, "", 4),
    ("su3_pure", 4, 3, "", 4),
    ("su2_fund", 4, 2, "fund", 4),
    ("su2_adj", 3, 2, "adj", 4),
    ("su3_fund", 3, 3, "fund", 4),
    ("su3_sym2", 4, 3, "2", 4),
    ("su3_61", 4, 3, "6,1", 3),
]


@pytest.mark.parametrize("case", GOLDEN, ids=[c[0] for c in GOLDEN])
def test_start_fields_and_momenta_equal_reference_dumps(goldens, case):
    g, _ = goldens
    tag, D, N, rep, L = case
    h = gha.GaugeHiggsHMC(L, rep, seed=1, D=D, N=N, precision=gha.FP64)
    h.hot(0.8, 12345, 0.3)
    ngen = N * N - 1
    exc = {}
    if rep:
        d = h.d
        exc["phi0"] = assert_equal_but_glibc_misrounds(h.get("phi").view(np.float64), g[tag + "_phi0"],
                                                       lambda i: scalar_keys(i, d, 12345), 1, 0.3, tag + " phi0")
    h.refresh_momenta()
    sP = O.key(TAG_P if rep else TAG_P_PURE, 0)
    exc["P0"] = assert_equal_but_glibc_misrounds(h.get("P"), g[tag + "_P0"], lambda i: momentum_keys(i, D, ngen, sP), 1,
                                                 what=tag + " P0")
    if rep:
        sPi = O.key(TAG_PI, 0)
        exc["pi0"] = assert_equal_but_glibc_misrounds(h.get("pi").view(np.float64), g[tag + "_pi0"],
                                                      lambda i: scalar_keys(i, d, sPi), 1, what=tag + " pi0")
    print(f"{tag}: glibc misroundings (the listed exceptions) {exc}")


@pytest.mark.parametrize("case", [("su3_sym2", 4, 3, "2", 4), ("su2_fund", 4, 2, "fund", (4, 3, 5, 6))], ids=["su3_sym2", "su2_fund"])
def test_fp32_refresh_is_the_rounded_draw(case):
    """FP32 contexts refresh with draw_gauss_f32: the device's fast double draw,
    rounded to float, with the correctly rounded draw deciding near float
    rounding boundaries. The stored momenta must be exactly float(FP64 draw)."""
    _, D, N, rep, L = case
    out = {}
    for prec in (gha.FP64, gha.FP32MD):
        h = gha.GaugeHiggsHMC(L, rep, seed=11, D=D, N=N, precision=prec)
        h.hot(0.8, 12345, 0.3)
        h.refresh_momenta()
        out[prec] = (h.get("P"), h.get("pi"))
    P64, pi64 = out[gha.FP64]
    P32, pi32 = out[gha.FP32MD]
    assert np.array_equal(P32, P64.astype(np.float32).astype(np.float64))
    assert np.array_equal(pi32.view(np.float64), pi64.view(np.float64).astype(np.float32).astype(np.float64))
