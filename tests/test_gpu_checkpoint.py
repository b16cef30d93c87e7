"""Checkpoint / resume: a chain stopped, written and reloaded into a fresh
context continues bit for bit like the uninterrupted chain (fields, counters,
dH, replica swaps); a checkpoint never loads into a mismatching lattice."""
import numpy as np
import pytest

import gauge_higgs_amd as gha

pytestmark = pytest.mark.gpu


def chain(prec):
    g = gha.GaugeHiggsHMC(4, "fund", seed=5, D=4, N=2, precision=prec).configure(beta=2.3, kappa=0.26, lam=0.5, nmd=8)
    g.hot(0.8, 12345, 0.3)
    return g


@pytest.mark.parametrize("prec", [gha.FP64, gha.FP32MD])
def test_resume_is_bitwise(tmp_path, prec):
    path = str(tmp_path / "c.ghx")
    a = chain(prec)
    a.lat.trajectories(3, use_graph=True)
    a.lat.save_checkpoint(path)
    a.lat.trajectories(3, use_graph=True)
    b = chain(prec)
    b.configure(kappa=0.1)  # overwritten by the checkpoint
    b.lat.load_checkpoint(path)
    b.lat.trajectories(3, use_graph=True)
    assert a.lat.counters() == b.lat.counters()
    assert np.array_equal(a.get("U"), b.get("U")) and np.array_equal(a.get("phi"), b.get("phi"))


def test_resume_replica_exchange(tmp_path):
    path = str(tmp_path / "r.ghx")
    def rs():
        return gha.ReplicaSet(4, 2, [[4, 4, 4, 4]] * 3, "fund", [2.3] * 3, [0.22, 0.23, 0.24], [0.5] * 3, seed0=9, nmd=8,
                              ladders=[[0, 1, 2]])
    a = rs()
    a.step(4)
    a.lat.save_checkpoint(path)
    a.step(4)
    b = rs()
    b.lat.load_checkpoint(path)
    b.step(4)
    for k in range(3):
        assert a.lat.counters(k) == b.lat.counters(k)
        assert np.array_equal(a.lat.get("U", k), b.lat.get("U", k))
    sa, sb = a.swap_stats(), b.swap_stats()
    assert sa[0] == sb[0] and np.array_equal(sa[1], sb[1]) and np.array_equal(sa[2], sb[2])


def test_mismatch_refused(tmp_path):
    path = str(tmp_path / "m.ghx")
    a = chain(gha.FP64)
    a.lat.save_checkpoint(path)
    b = gha.GaugeHiggsHMC(6, "fund", seed=5, D=4, N=2)
    with pytest.raises(gha.GhxError):
        b.lat.load_checkpoint(path)
    c = gha.GaugeHiggsHMC(4, "adj", seed=5, D=4, N=2)
    with pytest.raises(gha.GhxError):
        c.lat.load_checkpoint(path)


def test_nested_inner_integrator_in_checkpoint(tmp_path):
    # the inner integrator of the multi-timescale scheme is part of the
    # integrator settings a checkpoint must match
    path = str(tmp_path / "n.ghx")
    a = chain(gha.FP64)
    a.lat.set_nested(2, 3, inner=gha.OMELYAN)
    a.lat.trajectories(2, use_graph=True)
    a.lat.save_checkpoint(path)
    a.lat.trajectories(2, use_graph=True)
    b = chain(gha.FP64)
    b.lat.set_nested(2, 3, inner=gha.LEAPFROG)
    with pytest.raises(gha.GhxError):
        b.lat.load_checkpoint(path)
    b.lat.set_nested(2, 3, inner=gha.OMELYAN)
    b.lat.load_checkpoint(path)
    b.lat.trajectories(2, use_graph=True)
    assert a.lat.counters() == b.lat.counters()
    assert np.array_equal(a.get("U"), b.get("U")) and np.array_equal(a.get("phi"), b.get("phi"))


def test_version2_checkpoint_loads(tmp_path):
    # a version-2 file (header without the inner integrator word) still loads
    p3, p2 = str(tmp_path / "v3.ghx"), str(tmp_path / "v2.ghx")
    a = chain(gha.FP64)
    a.lat.trajectories(2, use_graph=True)
    a.lat.save_checkpoint(p3)
    raw = bytearray(open(p3, "rb").read())
    assert raw[:8] == b"GHXCKPT1" and int.from_bytes(raw[8:12], "little") == 3
    raw[8:12] = (2).to_bytes(4, "little")
    del raw[8 + 4 * 12:8 + 4 * 13]
    open(p2, "wb").write(bytes(raw))
    b = chain(gha.FP64)
    b.lat.load_checkpoint(p2)
    assert a.lat.counters() == b.lat.counters()
    assert np.array_equal(a.get("U"), b.get("U")) and np.array_equal(a.get("phi"), b.get("phi"))
