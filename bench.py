"""Benchmark of the HMC molecular-dynamics step on one MI355X.

Workload (default): 4D SU(3) + sextet scalar (Young rows {2}, d=6), L=32
(1,048,576 sites), beta=5.7, kappa=0.1, lambda=0.5, eps=tau/nmd=1/20, FP32 MD
(links, momenta, phi, pi, forces in FP32). A "step" is one fused MD step of
the whole lattice (kick then drift of every site), i.e. V site-updates; the
timed region runs K steps back to back with the fields resident in HBM.

Prints one JSON line: metric/value (site-updates/s), roofline of the MD-step
kernel (algorithmic bytes per launch / HIP-event launch time against 8 TB/s),
the CPU baseline (the reference compiled from its sources, or the FP64 oracle
restatement, timed on this host on a 2-step sample of the same lattice), and
"precision": FP32-MD against FP64 on SU(2)+fundamental 16^4 (forces, one
trajectory's dH and observables) and both MD throughputs;
"pure_gauge": SU(2) pure gauge 8^4 FP64 throughput and plaquette at beta=2.3;
"replicas": 32 SU(2)+fundamental chains (kappa 0.20-0.32, L 6-12) advanced by one
launch per MD step, their throughput and the measured L_link susceptibility peaks;
"learned_force": GPU seconds per accepted trajectory of exact-force HMC against
multi-timescale HMC with the learned surrogate on SU(3) + {6,1} (d = 48), L=12.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (D, N, rep, L, beta, kappa, lambda, precision, d)
    "su3_sextet_L32_fp32": (4, 3, "2", 32, 5.7, 0.1, 0.5, "fp32", 6),
    "su2_fund_L16_fp32": (4, 2, "fund", 16, 2.3, 0.265, 0.5, "fp32", 2),
    "su2_fund_L16_fp64": (4, 2, "fund", 16, 2.3, 0.265, 0.5, "fp64", 2),
    "su2_pure_L8_fp64": (4, 2, "", 8, 2.3, 0.0, 0.0, "fp64", 0),
    # the large irrep of the learned-force comparison (symmetric-power engine)
    "su3_61_L12_fp32": (4, 3, "6,1", 12, 5.7, 0.05, 0.5, "fp32", 48),
}


def algorithmic_bytes_per_site(D, N, d, w):
    """Every state element read once and written once per MD step:
    B = 2 [D 2N^2 w + D (N^2-1) w + 2 (2d) w] (U, P, phi, pi)."""
    return 2 * (D * 2 * N * N * w + D * (N * N - 1) * w + 2 * (2 * d) * w)


def stored_bytes_per_site(D, N, d, w):
    """What this implementation moves per site-update if nothing were re-read:
    in FP32 the links keep N-1 rows (the last is rebuilt on load)."""
    nr = N - 1 if w == 4 else N
    return 2 * (D * 2 * N * nr * w + D * (N * N - 1) * w + 2 * (2 * d) * w)


def load_traffic(workload, kernel, block_order=0):
    """HBM bytes per launch of `kernel` (launched with `block_order`) from the
    committed PMC summary (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by the
    copy-kernel calibration)."""
    path = os.path.join(ROOT, "profiles", f"md_step_{workload}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        entry = json.load(f).get(kernel, {})
    if "by_block_order" in entry:
        entry = entry["by_block_order"].get(str(block_order), {})
    return entry.get("hbm_bytes_per_launch")


def copy_calibration(mib):
    """A plain device copy of known size, for scaling the HBM byte counters."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.ones(n, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 3}))
    return 0


def physical_cores():
    """Physical cores of this host from lscpu (cores per socket x sockets), or None."""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout
        f = {k.strip(): v.strip() for k, v in (line.split(":", 1) for line in out.splitlines() if ":" in line)}
        return int(f["Core(s) per socket"]) * int(f["Socket(s)"])
    except Exception:
        return None


def cpu_baseline(workload, sample_steps=16, single_L=16, single_steps=8):
    """The reference (compiled from its own sources by oracle/Makefile) on this
    host's cores: OMP threads = the physical cores lscpu reports, capped by the
    CPUs this process may run on and by OMP_NUM_THREADS when the environment
    sets it; plus a one-thread sample on a smaller lattice."""
    D, N, rep, L, *_, d = WORKLOADS[workload]
    phys = physical_cores()
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env_cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(v for v in (phys, avail, env_cap or None) if v)
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if workload == "su3_sextet_L32_fp32" and os.access(harness, os.X_OK):
        def run(nthreads, lat, steps):
            env = dict(os.environ, OMP_NUM_THREADS=str(nthreads), OMP_PROC_BIND="close")
            out = subprocess.run([harness, "bench", str(lat), str(steps)], capture_output=True, text=True, env=env, timeout=600)
            return json.loads(out.stdout.strip().splitlines()[-1])
        r = run(threads, L, sample_steps)
        r1 = run(1, single_L, single_steps)
        return {"value": r["site_updates_per_s"], "unit": "site-updates/s", "cores": threads, "kind": "reference",
                "physical_cores_lscpu": phys, "cpus_available": avail, "omp_num_threads_env": env_cap or None,
                "one_thread": {"value": r1["site_updates_per_s"], "sample": f"{single_L}^4, {single_steps} drift+kick MD steps "
                                                                             f"({r1['seconds']:.1f} s), OMP_NUM_THREADS=1"},
                "sample": f"reference GaugeHiggsHMC<4,3> + GeneralRep{{2}} FP64, {L}^4, {sample_steps} drift+kick MD steps "
                          f"({r['seconds']:.1f} s), OMP_NUM_THREADS={threads} (lscpu physical cores {phys}, "
                          f"{avail} CPUs available to the process)"}
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    os.environ["OMP_NUM_THREADS"] = str(threads)  # read when the oracle library first starts OpenMP
    o = O.Oracle(D, N, L, rep, seed=1)
    o.couplings(beta=WORKLOADS[workload][4], kappa=WORKLOADS[workload][5], lam=0.5, nmd=20)
    o.hot(0.8)
    if rep:
        o.phi_gaussian(12345, 1.0)
    o.refresh()
    if d > 6:  # the tensor apply of a large irrep costs ~10^3 times a small one per site
        sample_steps = 1
    sec = O.lib().orc_time_kick_drift(o.h, sample_steps, 0.05)
    return {"value": o.V * sample_steps / sec, "unit": "site-updates/s", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle restatement, {L}^4, {sample_steps} kick+drift MD steps ({sec:.1f} s)"}


def replica_block(n_therm, n_meas, use_graph=True):
    """32 SU(2)+fundamental chains, kappa in [0.20, 0.32], L in {6, 8, 10, 12}, FP32 MD,
    one launch per MD step for all chains, graph-replayed trajectories and kappa
    swaps every trajectory: measured L_link susceptibility peaks and throughput."""
    from gauge_higgs_amd import replicas
    rs = replicas.kappa_scan()
    rs.lat.autotune(steps=20)  # launch shape (fields restored)
    launch = {"stream_seg": rs.lat.stream_segment(), "block_order": rs.lat.block_order(), "link_lanes": rs.lat.link_lanes(),
              "store_cached": rs.lat.store_policy(), "link_split": rs.lat.link_split()}
    m = replicas.measure_scan(rs, n_therm=n_therm, n_meas=n_meas, use_graph=use_graph)
    vtot = int(sum(rs.lat.vols))
    steps = 200
    ms = rs.lat.time_md_steps(steps)
    bpsite = algorithmic_bytes_per_site(4, 2, 2, 4)
    sa = list(m["swap_acceptance"].values())
    out = {"workload": "su2_fund_32chains_L6-12_fp32", "chains": rs.n, "sites": vtot, "beta": 2.3, "lambda": 0.5,
           "kappa": [0.20, 0.32], "nmd": 20, "integrator": "Omelyan",
           "md_site_updates_per_s": vtot * steps / (ms / 1e3), "ms_per_md_step": ms / steps,
           "roofline": {"bound": "hbm", "achieved": vtot * bpsite / (ms / steps / 1e3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": vtot * bpsite / (ms / steps / 1e3) / 1e9 / HBM_PEAK_GBS, "algorithmic_bytes_per_site": bpsite},
           "s_per_trajectory_cycle": m["s_per_cycle"], "trajectories_per_s": rs.n / m["s_per_cycle"],
           "hmc_acceptance_mean": float(m["hmc_acceptance"].mean()), "swap_acceptance_mean": float(np.mean(sa)) if sa else None,
           "measured_cycles": n_meas, "chi_link_peaks": replicas.chi_peaks(m), "launch": launch}
    rs.lat.close()
    return out


def precision_block():
    """SU(2)+fundamental, 16^4, beta=2.3, kappa=0.265, lambda=0.5, nmd=20 (Omelyan):
    FP32-MD against FP64 from the same start and momenta — forces, one trajectory's
    dH and the observables after it — plus the MD throughput of both."""
    import gauge_higgs_amd as gha
    out = {"workload": "su2_fund_L16", "sites": 16 ** 4}
    runs = {}
    for name, prec in (("fp64", gha.FP64), ("fp32", gha.FP32MD)):
        g = gha.GaugeHiggsHMC(16, "fund", seed=1, D=4, N=2, precision=prec)
        g.configure(beta=2.3, kappa=0.265, lam=0.5, nmd=20, reunit_each_traj=True)
        g.hot(0.8, 12345, 1.0)
        runs[name] = g
    a, b = runs["fp64"], runs["fp32"]
    b.set("U", a.get("U"))
    b.set("phi", a.get("phi"))
    Fa, Fpa = a.forces()
    Fb, Fpb = b.forces()
    out["force_rel_err"] = float(np.max(np.abs(Fb - Fa)) / np.max(np.abs(Fa)))
    out["scalar_force_rel_err"] = float(np.max(np.abs(Fpb - Fpa)) / np.max(np.abs(Fpa)))
    a.refresh_momenta()
    b.set("P", a.get("P"))
    b.set("pi", a.get("pi"))
    for g in (a, b):
        g.lat.set_integrator(20, gha.OMELYAN, True)
    ha, hb = a.hamiltonian(), b.hamiltonian()
    a.md_evolve()
    b.md_evolve()
    a.lat.reunitarize()
    b.lat.reunitarize()
    dHa, dHb = a.hamiltonian() - ha, b.hamiltonian() - hb
    oa, ob = a.observables(), b.observables()
    out["one_trajectory"] = {"dH_fp64": dHa, "dH_fp32": dHb, "dH_abs_err": abs(dHb - dHa), "H_fp64": ha,
                             **{f"{k}_abs_err": abs(ob[k] - oa[k]) for k in ("plaq", "Llink", "Lphi")},
                             **{f"{k}_fp64": oa[k] for k in ("plaq", "Llink", "Lphi")}}
    for name, g in runs.items():
        g.lat.autotune(steps=20)
        g.lat.time_md_steps(5)
        ms = g.lat.time_md_steps(100)
        out[f"md_site_updates_per_s_{name}"] = 16 ** 4 * 100 / (ms / 1e3)
        out[f"launch_{name}"] = {"block_order": g.lat.block_order(), "link_lanes": g.lat.link_lanes(), "store_cached": g.lat.store_policy(), "link_split": g.lat.link_split(),
                                 "stream_seg": g.lat.stream_segment()}
        g.lat.close()
    return out


def pure_gauge_block():
    """SU(2) pure gauge (GaugeHMC semantics), 8^4, beta=2.3, FP64, Omelyan nmd=20:
    MD throughput at the tuned launch shape and the plaquette after 60 trajectories
    from the hot start (published value for SU(2) at beta=2.3: 0.603)."""
    import gauge_higgs_amd as gha
    g = gha.GaugeHiggsHMC(8, "", seed=1, D=4, N=2, precision=gha.FP64)
    g.configure(beta=2.3, nmd=20, tau=1.0)
    g.hot(0.8, 12345, 1.0)
    g.lat.autotune(steps=20)
    g.lat.trajectories(20, use_graph=True)
    plaq = []
    for _ in range(40):
        g.lat.trajectories(1, use_graph=True)
        plaq.append(g.observables()["plaq"])
    g.refresh_momenta()
    g.lat.time_md_steps(20)
    steps = 500
    ms = g.lat.time_md_steps(steps)
    out = {"workload": "su2_pure_L8_fp64", "sites": 8 ** 4, "md_site_updates_per_s": 8 ** 4 * steps / (ms / 1e3),
           "ms_per_md_step": ms / steps, "plaquette": float(np.mean(plaq)), "plaquette_err": float(np.std(plaq) / np.sqrt(len(plaq) / 4)),
           "launch": {"block_order": g.lat.block_order(), "link_lanes": g.lat.link_lanes(), "stream_seg": g.lat.stream_segment(),
                      "store_cached": g.lat.store_policy(), "link_split": g.lat.link_split()}}
    g.lat.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="su3_sextet_L32_fp32", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-shape autotuner")
    ap.add_argument("--stream-seg", type=int, default=None, help="force a streaming segment (implies --no-tune)")
    ap.add_argument("--block-order", type=int, default=None, choices=(0, 1, 2, 3, 4, 5, 6),
                    help="force the block walk of the MD-step kernel (implies --no-tune)")
    ap.add_argument("--sym-pairs", type=int, default=None, choices=(0, 1, 2),
                    help="large irreps: the symmetric-power engine's link kernel: two links per wave with Sym^k(U) in registers "
                         "(2, the default where it applies), in LDS (1), or one link per wave (0)")
    ap.add_argument("--no-learned", action="store_true", help="skip the exact vs learned-force HMC comparison")
    ap.add_argument("--only-learned", action="store_true", help="run only the exact vs learned-force comparison")
    ap.add_argument("--learned-meas", type=int, default=200, help="measured trajectories per chain in the comparison")
    ap.add_argument("--no-graph", action="store_true", help="launch trajectories directly instead of replaying hipGraphs")
    ap.add_argument("--no-replicas", action="store_true", help="skip the 32-chain kappa-scan block")
    ap.add_argument("--no-precision", action="store_true", help="skip the FP32-vs-FP64 block (SU(2)+fund 16^4)")
    ap.add_argument("--replica-meas", type=int, default=1000, help="measured trajectory cycles of the kappa scan")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (calibrates the HBM byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.only_learned:
        from gauge_higgs_amd import learned
        print(json.dumps({"learned_force": learned.compare(n_meas=args.learned_meas, use_graph=not args.no_graph)}))
        return 0
    if args.stream_seg is not None or args.block_order is not None:
        args.no_tune = True
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the multi-GPU path is not built.", file=sys.stderr)
        return 2

    try:  # bring torch's device context up now, so the barrier before the timed region is only a barrier
        import torch
        torch.cuda.synchronize()
    except Exception:
        torch = None
    import gauge_higgs_amd as gha
    D, N, rep, L, beta, kappa, lam, prec, d = WORKLOADS[args.workload]
    precision = gha.FP32MD if prec == "fp32" else gha.FP64
    w = 4 if prec == "fp32" else 8
    g = gha.GaugeHiggsHMC(L, rep, seed=1, D=D, N=N, precision=precision)
    g.configure(beta=beta, kappa=kappa, lam=lam, tau=1.0, nmd=20)
    g.hot(0.8, 12345, 1.0)
    g.refresh_momenta()
    V = g.lat.vols[0]
    if args.sym_pairs is not None:
        g.lat.set_sym_pairs(args.sym_pairs)
    tuned = {"stream_seg": args.stream_seg or 0, "block_order": args.block_order or 0}
    if args.stream_seg is not None:
        g.lat.set_stream_segment(args.stream_seg)
    if args.block_order is not None:
        g.lat.set_block_order(args.block_order)
        tuned["block_order"] = g.lat.block_order()
    if not args.no_tune:
        seg, tms = g.lat.autotune(steps=50)  # launch-shape autotuner (fields restored)
        tuned = {"stream_seg": seg, "block_order": g.lat.block_order(), "link_lanes": g.lat.link_lanes(), "store_cached": g.lat.store_policy(), "link_split": g.lat.link_split(),
                 "ms_per_step_at_tune": tms}
    if args.warmup:
        g.lat.time_md_steps(args.warmup)
    g.lat.sync()
    if torch is not None:
        torch.cuda.synchronize()
    t0 = time.perf_counter()
    ms = g.lat.time_md_steps(args.steps)  # HIP events on the context stream, bracketing K launches
    g.lat.sync()
    if torch is not None:
        torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    sec = ms / 1e3
    value = V * args.steps / sec
    bpsite = algorithmic_bytes_per_site(D, N, d, w)
    launch_s = sec / args.steps
    achieved = V * bpsite / launch_s / 1e9
    kernel = "md_stream_kernel" if tuned.get("stream_seg") else ("md_link_kernel" if tuned.get("link_lanes") else "md_step_kernel")
    if d > 6:  # a large irrep: the matter part runs in the symmetric-power engine before the step kernel
        link = ("sym_link_kernel", "sym_link_pair_kernel", "sym_link_col_kernel")[g.lat.sym_pairs()]
        kernel = f"sym_lift_kernel + {link} + sym_finish_kernel + " + kernel
        tuned["sym_pairs"] = g.lat.sym_pairs()
    traffic = load_traffic(args.workload, kernel, tuned.get("block_order", 0))
    line = {
        "metric": "MD site-updates/s; GPU-s per accepted traj, learned vs exact force (1 GPU)",
        "value": value,
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec == "fp32" else "f64",
        "data": "synthetic (reference hot start sigma=0.8, Gaussian phi sigma=1, Gaussian momenta; seed 1)",
        "config": {"workload": args.workload, "D": D, "N": N, "rep": rep or "none", "L": L, "sites": V, "beta": beta,
                   "kappa": kappa, "lambda": lam, "eps": 0.05, "wall_s": wall, "launch": tuned},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": kernel, "algorithmic_bytes_per_site": bpsite, "algorithmic_bytes_per_launch": V * bpsite,
                     "stored_state_bytes_per_site": stored_bytes_per_site(D, N, d, w)},
    }
    g.lat.close()
    if not args.no_replicas:
        line["replicas"] = replica_block(200, args.replica_meas, use_graph=not args.no_graph)
    if not args.no_precision:
        line["precision"] = precision_block()
        line["pure_gauge"] = pure_gauge_block()
    if not args.no_learned:
        from gauge_higgs_amd import learned
        line["learned_force"] = learned.compare(n_meas=args.learned_meas, use_graph=not args.no_graph)
    if not args.no_cpu:
        try:
            line["cpu_baseline"] = cpu_baseline(args.workload)
        except Exception as e:  # the GPU figure stands on its own
            line["cpu_baseline"] = {"value": None, "error": str(e)[:200]}
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
