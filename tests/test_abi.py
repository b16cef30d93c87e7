"""CPU checks of the drop-in boundary: libghx.so loads without a GPU, exports
every symbol declared in include/ghx.h, its host-side representation builder
reproduces the reference basis, and bad input comes back as error codes."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "ghx.h")).read()
    return sorted(set(re.findall(r"\b(ghx_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from gauge_higgs_amd import _lib
    L = _lib.load()
    syms = declared_symbols()
    assert len(syms) >= 30
    for s in syms:
        assert hasattr(L, s), s


def test_clis_built():
    for name in ("gh_hmc", "hmc_higgs", "hmc_higgs_multi"):
        assert os.access(os.path.join(ROOT, "gauge_higgs_amd", "bin", name), os.X_OK)


@pytest.mark.parametrize("tag,N,spec", [("rep_su3_2", 3, "2"), ("rep_su3_11", 3, "1,1"), ("rep_su2_3", 2, "3"), ("rep_su3_61", 3, "6,1")])
def test_product_rep_basis_matches_reference(goldens, tag, N, spec):
    from gauge_higgs_amd import rep_query
    g, sc = goldens
    r = rep_query(N, spec)
    assert r["d"] == int(sc[tag + "_d"]) and r["dimT"] == int(sc[tag + "_dimT"]) and r["nality"] == int(sc[tag + "_nality"])
    assert np.max(np.abs(r["W"].ravel() - g[tag + "_W"])) < 1e-15
    T = g[tag + "_T"].view(np.complex128).reshape(r["T"].shape)
    assert np.max(np.abs(r["T"] - T)) < 1e-13


def test_product_rep_fund_adj():
    from gauge_higgs_amd import rep_query
    import oracle_lib as O
    for N in (2, 3):
        f = rep_query(N, "fund")
        assert f["d"] == N and f["nality"] == 1 and not f["real"]
        assert np.allclose(f["T"], O.generators(N))
        a = rep_query(N, "adj")
        o = O.Oracle(3, N, 2, "adj", seed=1)
        assert a["d"] == N * N - 1 and a["real"]
        assert np.max(np.abs(a["T"] - o.rep_T())) < 1e-14


def test_invalid_specs_return_error_codes():
    from gauge_higgs_amd import _lib
    L = _lib.load()
    d = C.c_int()
    for N, spec in [(2, b"1,1"), (3, b"1,2"), (3, b"0"), (3, b"")]:
        rc = L.ghx_rep_query(N, spec, C.byref(d), None, None, None, None, None)
        if spec == b"":
            assert rc == 0 and d.value == 0  # pure gauge
        else:
            assert rc == -5
    assert L.ghx_rep_query(4, b"fund", None, None, None, None, None, None) == -1
    out = C.c_void_p()
    ext = (C.c_int * 4)(4, 4, 4, 4)
    assert L.ghx_create(C.byref(out), 5, 2, 1, ext, 0) == -1  # D out of range, no device touched
    assert L.ghx_create(C.byref(out), 4, 4, 1, ext, 0) == -1  # N out of range
    bad = (C.c_int * 4)(4, 1, 4, 4)
    assert L.ghx_create(C.byref(out), 4, 2, 1, bad, 0) == -1  # extent < 2


def test_python_layer_fails_loudly_without_library(tmp_path, monkeypatch):
    from gauge_higgs_amd import _lib
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    with pytest.raises(RuntimeError):
        _lib.load()


@pytest.mark.parametrize("tag,N,spec", [("multi_su2_fund", 2, "fund"), ("multi_su2_adj", 2, "adj"), ("multi_su3_2", 3, "2")])
def test_casimir_channels_match_reference(goldens, tag, N, spec):
    """Host construction of the multi-invariant potential's channels against the
    reference's own Lanczos result (scalar_invariants.hpp:150-179)."""
    arr, sc = goldens
    import gauge_higgs_amd as gha

    lam = gha.casimir_channels(N, spec)
    assert len(lam) == sc[tag + "_nch"]
    assert np.max(np.abs(lam - arr[tag + "_lambda"])) < 1e-9
