"""Monte Carlo series analysis in the library (ghx_series_analysis,
ghx_autocorr_function; host code, no GPU): the reference's Stats accumulator
(observables.hpp:104-145) and autocorrelation module (autocorr.hpp:55-212).

Pinned three ways: the reference's own outputs on fixed AR(1) and constant
series (tests/golden, written by oracle/_ref/ref_harness), the closed forms its
test suite uses (AR(1): rho(t) = a^t, tau_int = (1+a)/(2(1-a)); the exact
inflation identity; the blocking plateau), and hand-checkable small cases."""
import numpy as np
import pytest

import gauge_higgs_amd as gha

KEYS = ("mean", "var", "binder", "binned_error", "tau_int", "window", "n_eff", "naive_error", "autocorr_error")


@pytest.mark.parametrize("tag", ["series_white", "series_ar05", "series_ar09", "series_ar099", "series_const"])
def test_matches_reference(goldens, tag):
    g, _ = goldens
    x = g[tag + "_x"]
    ref = g[tag + "_out"]
    got = gha.series_analysis(x)
    for k, r in zip(KEYS, ref):
        assert abs(got[k] - r) <= 1e-13 * max(1.0, abs(r)), (k, got[k], r)
    for bs, r in zip((1, 3, 16, 50), g[tag + "_blocking"]):
        assert abs(gha.series_analysis(x, block=bs)["blocking_error"] - r) <= 1e-15 * max(1.0, abs(r))
    rho = gha.autocorr_function(x, 12)
    assert rho.shape == g[tag + "_rho"].shape
    assert np.max(np.abs(rho - g[tag + "_rho"])) < 1e-14


def ar1(a, n, seed):
    rng = np.random.default_rng(seed)
    eta = rng.normal(size=n)
    x = np.empty(n)
    x[0] = eta[0]
    s = np.sqrt(1.0 - a * a)
    for t in range(1, n):
        x[t] = a * x[t - 1] + s * eta[t]
    return x


@pytest.mark.parametrize("a,n,rtol", [(0.0, 200000, 0.06), (0.5, 200000, 0.06), (0.9, 600000, 0.18)])
def test_ar1_closed_forms(a, n, rtol):
    """Statistical recovery of rho(1) = a, rho(2) = a^2 and tau_int, and
    Sokal's self-consistency W >= 6 tau_int, with the reference test's
    tolerances."""
    x = ar1(a, n, 11)
    rho = gha.autocorr_function(x, 8)
    assert abs(rho[0] - 1.0) < 1e-12
    assert abs(rho[1] - a) < 0.02
    assert abs(rho[2] - a * a) < 0.02
    s = gha.series_analysis(x)
    tau = (1 + a) / (2 * (1 - a))
    assert abs(s["tau_int"] - tau) < rtol * tau
    assert s["window"] >= 6.0 * s["tau_int"] - 1.0
    # exact identities between the outputs
    assert abs(s["autocorr_error"] / s["naive_error"] - np.sqrt(2 * s["tau_int"])) < 1e-10
    assert abs(s["n_eff"] - n / (2 * s["tau_int"])) < 1e-9 * n
    assert abs(s["naive_error"] - np.sqrt(np.var(x) / n)) < 1e-9 * s["naive_error"]


def test_white_noise_and_blocking_plateau():
    x = ar1(0.0, 200000, 12)
    s = gha.series_analysis(x)
    assert abs(s["tau_int"] - 0.5) < 0.04
    assert abs(s["autocorr_error"] / s["naive_error"] - 1.0) < 0.05
    assert abs(s["n_eff"] / x.size - 1.0) < 0.06
    x = ar1(0.9, 600000, 13)
    s = gha.series_analysis(x)
    be = [gha.series_analysis(x, block=bs)["blocking_error"] for bs in (1, 4, 16, 64, 256, 1024, 4096)]
    for p, q in zip(be, be[1:]):
        assert q >= p - 0.05 * q  # rises into a plateau
    assert abs(be[5] - s["autocorr_error"]) < 0.15 * s["autocorr_error"]
    assert be[0] < 0.6 * be[5]
    assert abs(be[0] - s["naive_error"]) < 0.02 * s["naive_error"]
    assert abs(s["autocorr_error"] / s["naive_error"] - np.sqrt(19.0)) < 0.1 * np.sqrt(19.0)


def test_small_exact_cases():
    v = np.array([1.0, 2.0, 3.0, 4.0])
    s = gha.series_analysis(v)
    assert s["mean"] == 2.5 and s["var"] == 1.25
    rho = gha.autocorr_function(v)
    assert len(rho) == 4 and abs(rho[1] - 0.3125 / 1.25) < 1e-15
    # binned error with n < 32: nbins = max(2, n/2) = 2 blocks of 2 -> means 1.5, 3.5
    assert abs(s["binned_error"] - np.sqrt(np.var([1.5, 3.5], ddof=1) / 2)) < 1e-15
    c = gha.series_analysis(np.full(1000, 3.7))
    assert c["tau_int"] == 0.5 and c["autocorr_error"] == 0.0 and c["window"] == 3
    assert np.array_equal(gha.autocorr_function(np.full(50, 3.7), 5), [1, 0, 0, 0, 0, 0])
    e = gha.series_analysis(np.zeros(0))
    assert e["tau_int"] == 0.5 and e["autocorr_error"] == 0.0 and e["n_eff"] == 0.0 and e["window"] == 0
    assert gha.autocorr_function(np.zeros(0)).size == 0
    o = gha.series_analysis(np.array([42.0]))
    assert o["tau_int"] == 0.5 and o["autocorr_error"] == 0.0 and o["binned_error"] == 0.0
