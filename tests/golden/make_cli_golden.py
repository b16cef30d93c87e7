"""Writes tests/golden/cli_stdout.json: the stdout of the reference's own
drivers (`gh_hmc` from main.cpp and `hmc_higgs` from hmc_higgs.cpp, built by
`make -C oracle ref` from the reference sources into oracle/_ref/) for a few
short runs, with OMP_NUM_THREADS=1. The drop-in CLIs are compared with these
line by line (tests/test_gpu_cli.py).

    python tests/golden/make_cli_golden.py
"""
import json
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF = os.path.join(ROOT, "oracle", "_ref")

RUNS = [
    ["gh_hmc", "4", "2.3", "5", "10", "10"],
    ["gh_hmc", "4", "2.3", "3", "6", "8", "1.0", "7"],
    ["hmc_higgs", "fund", "4", "2.3", "0.265", "0.5", "5", "10", "10"],
    ["hmc_higgs", "adj", "4", "2.3", "0.3", "0.5", "3", "6", "8", "1.0", "3"],
]


def main():
    env = dict(os.environ, OMP_NUM_THREADS="1")
    out = []
    for argv in RUNS:
        r = subprocess.run([os.path.join(REF, argv[0]), *argv[1:]], capture_output=True, text=True, env=env, check=True)
        out.append({"argv": argv, "stdout": r.stdout})
    with open(os.path.join(ROOT, "tests", "golden", "cli_stdout.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
