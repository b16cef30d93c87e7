"""The bench workload at its full size (SU(3) + sextet {2}, 32^4, FP32 MD, the
C3 configuration), checked through properties that do not depend on size:
FP32 forces against FP64 forces on the same fields, the thread-per-site step
(shared plaquettes, two-row products) against the link-parallel step (one
staple at a time) on the same fields, and MD reversibility."""
import numpy as np
import pytest

import gauge_higgs_amd as gha

pytestmark = pytest.mark.gpu
L = 32


def relmax(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def make(prec):
    # 4 MD steps of the bench's step size 1/20
    return gha.GaugeHiggsHMC(L, "2", seed=1, D=4, N=3, precision=prec).configure(beta=5.7, kappa=0.1, lam=0.5, nmd=4, tau=0.2,
                                                                                  reunit_each_traj=False)


def test_full_size_forces_fp32_fp64_and_both_step_kernels():
    gd = make(gha.FP64)
    gd.hot(0.8, 12345, 1.0)
    F64, Fp64 = gd.forces()
    U, phi = gd.get("U"), gd.get("phi")
    del gd
    gf = make(gha.FP32MD)
    gf.set("U", U)
    gf.set("phi", phi)
    assert gf.lat.link_lanes() == 0
    F32, Fp32 = gf.forces()
    print(f"32^4 FP32 vs FP64: link force {relmax(F32, F64):.2e}, scalar force {relmax(Fp32, Fp64):.2e}")
    assert relmax(F32, F64) < 2e-5 and relmax(Fp32, Fp64) < 2e-5
    gf.lat.set_link_lanes(1)
    F32b, Fp32b = gf.forces()
    assert relmax(F32b, F32) < 2e-5 and relmax(Fp32b, Fp32) < 2e-5


def test_full_size_reversibility_fp32():
    g = make(gha.FP32MD)
    g.hot(0.8, 12345, 1.0)
    g.refresh_momenta()
    U0, phi0 = g.get("U"), g.get("phi")
    g.md_evolve()
    g.set("P", -g.get("P"))
    g.set("pi", -g.get("pi"))
    g.md_evolve()
    dU, dphi = np.max(np.abs(g.get("U") - U0)), np.max(np.abs(g.get("phi") - phi0))
    print(f"32^4 FP32 reversibility over 4 MD steps: U {dU:.2e}, phi {dphi:.2e}")
    assert dU < 2e-4 and dphi < 2e-4


def test_full_size_c5_engine_matches_tensor_apply():
    """SU(3) + {6,1} (d = 48) at 12^4, the learned-force workload: the
    symmetric-power engine and the slot-by-slot tensor apply give the same
    forces on the same fields."""
    out = []
    for engine in (1, 0):
        g = gha.GaugeHiggsHMC(12, "6,1", seed=1, D=4, N=3, precision=gha.FP32MD).configure(beta=5.7, kappa=0.05, lam=0.5)
        g.lat.set_sym_engine(engine)
        assert g.lat.sym_engine() == engine
        g.hot(0.4, 12345, 0.27)
        out.append(g.forces())
    (Fe, Fpe), (Ft, Fpt) = out
    print(f"12^4 {{6,1}} engine vs tensor: link force {relmax(Fe, Ft):.2e}, scalar force {relmax(Fpe, Fpt):.2e}")
    assert relmax(Fe, Ft) < 2e-5 and relmax(Fpe, Fpt) < 2e-5
