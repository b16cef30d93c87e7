"""The drop-in CLIs against the stdout of the reference's own drivers.

tests/golden/cli_stdout.json holds what `gh_hmc` (main.cpp:11-51) and
`hmc_higgs` (hmc_higgs.cpp:40-100), compiled from the reference sources, print
for four short runs (make_cli_golden.py). The GPU CLIs (FP64, the reference's
integrator and seeds) must print the same lines in the same order, the same
words, and numbers that agree to the last digit printed, one unit of it
allowed for rounding, except the unitarity violation, which only has to stay
at rounding level. With bit-exact RNG streams and FP64 trajectories that
agree to ~1e-10, the accept decisions, and so the acceptance, are identical.
A bad rep spec prints `rep error:` and exits 1, as the reference does."""
import json
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NUM = re.compile(r"[-+]?\d+\.\d+(?:e[-+]\d+)?|[-+]?\d+(?:e[-+]\d+)?")

with open(os.path.join(ROOT, "tests", "golden", "cli_stdout.json")) as f:
    RUNS = json.load(f)


def split(line):
    return NUM.findall(line), NUM.sub("#", line)


@pytest.mark.parametrize("run", RUNS, ids=[" ".join(r["argv"]) for r in RUNS])
def test_cli_stdout_matches_reference(run):
    argv = run["argv"]
    out = subprocess.run([os.path.join(ROOT, "gauge_higgs_amd", "bin", argv[0]), *argv[1:]], capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stderr
    print(out.stdout)
    got, ref = out.stdout.strip().splitlines(), run["stdout"].strip().splitlines()
    assert len(got) == len(ref)
    for g, r in zip(got, ref):
        gn, gw = split(g)
        rn, rw = split(r)
        assert gw == rw, (g, r)
        for a, b in zip(gn, rn):
            if r.startswith("max"):
                assert float(a) < 1e-13 and float(b) < 1e-13, (g, r)
                continue
            if "." in b and "e" not in b:
                unit = 10.0 ** -len(b.split(".")[1])
            else:
                unit = abs(float(b)) * 1e-3 if "e" in b else 0.0
            assert abs(float(a) - float(b)) <= 1.01 * unit, (g, r)


def test_cli_bad_rep_spec():
    bad = subprocess.run([os.path.join(ROOT, "gauge_higgs_amd", "bin", "hmc_higgs"), "1,2", "4", "2.3", "0.2", "0.5"],
                         capture_output=True, text=True, timeout=60)
    assert bad.returncode == 1 and "rep error:" in bad.stderr
