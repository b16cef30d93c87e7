"""The reference's finite-size validation of the SU(2)-fundamental Higgs
transition, rerun on the GPU (validate_su2higgs.sh:10-26, results in
validation_su2higgs.dat:3-18).

Settings as the script's: hmc_higgs fund, L in {6, 8}, beta = 2.3, lambda =
0.5, kappa in {0.18, 0.21, 0.23, 0.245, 0.255, 0.265, 0.28, 0.31}, 120
thermalisation + 300 measured trajectories, nmd = 18 (Omelyan, tau = 1), seed
1234, hot start (sigma 0.8) with phi from stream 12345 (sigma 0.3), and the
trajectory counter reset after thermalisation as hmc_higgs.cpp:72-76 does. All
16 chains run as replicas of one context in FP64, one launch per MD step.

Compared with the reference's table:
- <L_link> within z < 4 of the reference, with the combined error of the
  reference's 16-bin error and this run's autocorrelation-aware one;
- the HMC acceptance within 0.05;
- chi_link = V Var(L_link) at kappa = 0.265 (1.78 at 6^4, 2.42 at 8^4) within
  z < 3, the error from a 16-block jackknife of this run counted for both;
- on each volume the largest chi_link at kappa 0.265 or 0.28, next to the
  reference's peak.
"""
import numpy as np
import pytest

import gauge_higgs_amd as gha

pytestmark = pytest.mark.gpu

KAPPAS = [0.18, 0.21, 0.23, 0.245, 0.255, 0.265, 0.28, 0.31]
# validation_su2higgs.dat:3-18: L, kappa, L_link, err, chi_link, chi_phi, Binder, acc
REF = """
6 0.18 0.166848 0.000733 0.15767 0.48571 0.05518 0.9867
6 0.21 0.210403 0.000572 0.21243 0.55682 0.01629 0.9900
6 0.23 0.249748 0.000870 0.22317 0.53845 -0.20364 0.9833
6 0.245 0.287791 0.001649 0.35700 0.56607 0.05491 0.9867
6 0.255 0.323802 0.003075 0.55158 0.75050 0.02263 0.9767
6 0.265 0.403819 0.006560 1.78276 1.18294 -0.05913 0.9767
6 0.28 0.704686 0.004079 1.51707 1.16741 0.01804 0.9867
6 0.31 1.094488 0.003568 1.45503 1.18556 -0.03564 0.9900
8 0.18 0.167099 0.000467 0.15500 0.42963 -0.00658 0.9767
8 0.21 0.212366 0.000531 0.19078 0.50128 0.01906 0.9833
8 0.23 0.248609 0.000720 0.27844 0.58217 0.09453 0.9833
8 0.245 0.286491 0.000868 0.33610 0.57976 -0.03013 0.9667
8 0.255 0.319696 0.001335 0.56430 0.62426 -0.02497 0.9767
8 0.265 0.453005 0.005193 2.42477 1.38787 0.13916 0.9733
8 0.28 0.698610 0.003033 1.73318 1.23411 0.07314 0.9667
8 0.31 1.095014 0.001746 1.51161 1.15204 -0.00123 0.9667
"""
NTHERM, NMEAS, NMD, SEED = 120, 300, 18, 1234


def reference_table():
    rows = [line.split() for line in REF.strip().splitlines()]
    return np.array([[float(v) for v in r] for r in rows])


def jackknife_chi(x, V, nblk=16):
    n = len(x) // nblk * nblk
    x = x[:n]
    full = V * x.var()
    b = x.reshape(nblk, -1)
    est = []
    for i in range(nblk):
        y = np.delete(b, i, axis=0).ravel()
        est.append(V * y.var())
    est = np.array(est)
    return full, float(np.sqrt((nblk - 1) * np.mean((est - est.mean()) ** 2)))


def run_scan():
    ext = [[6] * 4] * len(KAPPAS) + [[8] * 4] * len(KAPPAS)
    lat = gha.Lattices(4, 2, ext, rep="fund", precision=gha.FP64)
    lat.set_integrator(NMD, gha.OMELYAN, True)
    kap = KAPPAS + KAPPAS
    for r in range(lat.nrep):
        lat.set_couplings(r, 2.3, kap[r], 0.5, 1.0)
        lat.set_seed(r, SEED)
        lat.init_fields(r, gha.HOT, 0.8, 1, 12345, 0.3)
    lat.trajectories(NTHERM, use_graph=True)
    for r in range(lat.nrep):
        lat.set_counters(r, 0, 0)
    series = np.zeros((NMEAS, lat.nrep))
    for t in range(NMEAS):
        lat.trajectories(1, use_graph=True)
        series[t] = lat.observables_all()[:, 3]
    acc = np.array([lat.counters(r)[1] / lat.counters(r)[0] for r in range(lat.nrep)])
    vols = np.array(lat.vols, dtype=float)
    lat.close()
    return series, acc, vols


@pytest.fixture(scope="module")
def scan():
    return run_scan()


def test_link_energy_and_acceptance_match_reference(scan):
    series, acc, vols = scan
    ref = reference_table()
    zs = []
    for r in range(len(ref)):
        L, kappa, ll, err, chi, _, _, racc = ref[r]
        st = gha.series_analysis(series[:, r])
        e = max(st["binned_error"], st["autocorr_error"])
        z = abs(st["mean"] - ll) / np.hypot(e, err)
        zs.append(z)
        print(f"L={int(L)} kappa={kappa}: L_link {st['mean']:.6f} +- {e:.6f} (tau_int {st['tau_int']:.1f}) "
              f"vs {ll:.6f} +- {err:.6f}  z={z:.2f}  acc {acc[r]:.4f} vs {racc:.4f}")
        assert abs(acc[r] - racc) < 0.05
    assert max(zs) < 4.0, zs


def test_susceptibility_peak_matches_reference(scan):
    series, _, vols = scan
    ref = reference_table()
    for L in (6, 8):
        sel = np.where(ref[:, 0] == L)[0]
        chis = []
        for r in sel:
            c, ce = jackknife_chi(series[:, r], vols[r])
            chis.append((c, ce))
        c = np.array([x[0] for x in chis])
        print(f"L={L}: chi_link {np.round(c, 3).tolist()}; reference {ref[sel, 4].tolist()}")
        i265 = KAPPAS.index(0.265)
        chi, err = chis[i265]
        chi_ref = ref[sel[i265], 4]
        z = abs(chi - chi_ref) / (np.sqrt(2.0) * err)
        print(f"L={L} kappa=0.265: chi_link {chi:.3f} +- {err:.3f} vs {chi_ref}  z={z:.2f}")
        assert z < 3.0
        assert KAPPAS[int(np.argmax(c))] in (0.265, 0.28)
