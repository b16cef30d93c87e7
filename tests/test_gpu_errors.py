"""Error paths and state rules of the C ABI on the GPU.

- A rejected trajectory restores U and phi bit for bit and counts the
  trajectory but not an acceptance (gauge_higgs_hmc.hpp:185-191).
- A non-finite Hamiltonian comes back as GHX_E_NAN.
- ghx_set_rep drops what belonged to the previous irrep (plugged-in
  potential, frozen flag): the forces afterwards are those of the new irrep
  with the quartic potential.
- An irrep whose tensor apply needs more LDS per workgroup than the device
  allows is refused at ghx_set_rep (GHX_E_UNSUPPORTED) instead of failing at
  its first launch.
- Replica exchange refuses ladders whose replicas no longer share lambda.
- Checkpoints carry the swap statistics and refuse a context with another
  frozen flag or other integrator settings.
"""
import ctypes as C

import numpy as np
import pytest

import gauge_higgs_amd as gha
import oracle_lib as O

pytestmark = pytest.mark.gpu

GHX_E_INVALID, GHX_E_NAN, GHX_E_UNSUPPORTED = -1, -3, -4


def fund_chain(prec=gha.FP64, nmd=8, tau=1.0):
    g = gha.GaugeHiggsHMC(4, "fund", seed=3, D=4, N=2, precision=prec).configure(beta=2.3, kappa=0.265, lam=0.5, nmd=nmd,
                                                                                    tau=tau)
    g.hot(0.8, 12345, 0.3)
    return g


@pytest.mark.parametrize("prec", [gha.FP64, gha.FP32MD])
def test_forced_reject_restores_fields_bitwise(prec):
    # one MD step over a trajectory of length 3 from a hot start: dH in the
    # hundreds, so exp(-dH) is below every uniform the RNG can draw (>= 2^-54)
    g = fund_chain(prec, nmd=1, tau=3.0)
    U0, phi0 = g.get("U"), g.get("phi")
    t0, a0, _ = g.lat.counters()
    acc = g.trajectory()
    t1, a1, dH = g.lat.counters()
    print(f"forced reject: prec={prec} dH={dH:.3e}")
    assert dH > 40.0 and not acc
    assert (t1, a1) == (t0 + 1, a0)
    assert np.array_equal(g.get("U"), U0) and np.array_equal(g.get("phi"), phi0)


def test_forced_reject_matches_oracle_decision():
    g = fund_chain(gha.FP64, nmd=1, tau=3.0)
    o = O.Oracle(4, 2, 4, "fund", seed=3)
    o.couplings(beta=2.3, kappa=0.265, lam=0.5, tau=3.0, nmd=1)
    o.hot(0.8)
    o.phi_gaussian(12345, 0.3)
    acc = g.trajectory()
    acc_o, dH_o, _, _ = o.trajectory()
    assert not acc and not acc_o
    assert abs(g.last_dH - dH_o) <= 1e-9 * max(1.0, abs(dH_o))
    assert np.max(np.abs(g.get("U") - o.get("U"))) < 1e-14


def test_nonfinite_hamiltonian_is_reported():
    g = fund_chain(gha.FP64)
    phi = g.get("phi")
    phi[5, 0] = np.nan
    g.set("phi", phi)
    with pytest.raises(gha.GhxError) as e:
        g.lat.trajectory_all()
    assert e.value.code == GHX_E_NAN


def test_set_rep_drops_the_previous_irreps_state():
    lat = gha.Lattices(4, 2, [[4, 4, 4, 4]], rep="fund", precision=gha.FP64)
    lat.set_couplings(0, 2.3, 0.265, 0.5, 1.0)
    f = np.ones(len(gha.casimir_channels(2, "fund")))
    lat.set_multi_potential(f, 1.0)
    lat.set_frozen_phi(True)
    d = C.c_int()
    lat._check(lat._lib.ghx_set_rep(lat.h, b"adj", C.byref(d)))
    assert d.value == 3
    lat.d = 3
    lat.set_couplings(0, 2.3, 0.3, 0.5, 1.0)
    lat.set_seed(0, 1)
    lat.init_fields(0, gha.HOT, 0.8, 1, 12345, 0.3)
    o = O.Oracle(4, 2, 4, "adj", seed=1)
    o.couplings(beta=2.3, kappa=0.3, lam=0.5, nmd=8)
    o.hot(0.8)
    o.phi_gaussian(12345, 0.3)
    F, Fp = lat.forces()
    Fo, Fpo = o.forces()
    assert np.max(np.abs(F - Fo)) < 1e-12 * np.max(np.abs(Fo))
    assert np.max(np.abs(Fp - Fpo)) < 1e-12 * np.max(np.abs(Fpo))
    # not frozen any more: the momenta keep their radial part
    lat.refresh_momenta()
    ph, pi = lat.get("phi"), lat.get("pi")
    assert np.max(np.abs(np.sum((ph.conj() * pi).real, axis=1))) > 1e-3
    lat.close()


def test_irrep_beyond_the_lds_limit_is_refused():
    # SU(3) {9}: d = 55, outside the symmetric-power engine (dim Sym^9 = 55 > 36);
    # its tensor apply holds 3^9 = 19683 complex entries: 315 KB in FP64
    with pytest.raises(gha.GhxError) as e:
        gha.Lattices(4, 3, [[2, 2, 2, 2]], rep="9", precision=gha.FP64)
    print(e.value)
    assert e.value.code == GHX_E_UNSUPPORTED and "LDS" in str(e.value)


def test_ladder_with_unequal_lambda_refuses_to_swap():
    rs = gha.ReplicaSet(4, 2, [[4, 4, 4, 4]] * 2, "fund", [2.3] * 2, [0.22, 0.24], [0.5] * 2, seed0=9, nmd=4, ladders=[[0, 1]])
    rs.step(1)
    rs.lat.set_couplings(1, 2.3, 0.24, 0.6, 1.0)
    with pytest.raises(gha.GhxError) as e:
        rs.step(1)
    assert e.value.code == GHX_E_INVALID
    rs.lat.set_couplings(1, 2.3, 0.24, 0.5, 1.0)
    rs.step(1)


def rs3():
    return gha.ReplicaSet(4, 2, [[4, 4, 4, 4]] * 3, "fund", [2.3] * 3, [0.22, 0.23, 0.24], [0.5] * 3, seed0=9, nmd=8,
                          ladders=[[0, 1, 2]])


def test_checkpoint_carries_swap_statistics(tmp_path):
    path = str(tmp_path / "s.ghx")
    a = rs3()
    a.step(6)
    a.lat.save_checkpoint(path)
    sa = a.swap_stats()
    b = rs3()
    b.lat.load_checkpoint(path)
    sb = b.swap_stats()
    assert sa[0] == sb[0] and np.array_equal(sa[1], sb[1]) and np.array_equal(sa[2], sb[2])
    assert sa[1].sum() > 0
    a.step(3)
    b.step(3)
    assert np.array_equal(a.swap_stats()[2], b.swap_stats()[2])


def test_checkpoint_refuses_other_frozen_flag_or_integrator(tmp_path):
    path = str(tmp_path / "f.ghx")
    a = fund_chain(gha.FP64)
    a.lat.save_checkpoint(path)
    b = fund_chain(gha.FP64)
    b.frozen_phi = True
    with pytest.raises(gha.GhxError):
        b.lat.load_checkpoint(path)
    c = fund_chain(gha.FP64, nmd=10)
    with pytest.raises(gha.GhxError):
        c.lat.load_checkpoint(path)
    d = fund_chain(gha.FP64)
    d.lat.load_checkpoint(path)
