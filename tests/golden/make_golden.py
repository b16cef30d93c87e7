"""Regenerate the golden fixtures from the reference itself.

Runs oracle/_ref/ref_harness (built by `make -C oracle ref` from the reference
headers in place) single-threaded, and packs its raw outputs into
tests/golden/reference_goldens.npz (arrays) and reference_scalars.json.
Only this script and the tests read these files.
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")


def main():
    if not os.path.exists(HARNESS):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "ref"])
    with tempfile.TemporaryDirectory() as tmp:
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.check_call([HARNESS, "dump", tmp], env=env)
        arrays = {}
        for fn in sorted(os.listdir(tmp)):
            path = os.path.join(tmp, fn)
            if fn.endswith(".f64"):
                arrays[fn[:-4]] = np.fromfile(path, dtype="<f8")
            elif fn.endswith(".u64"):
                arrays[fn[:-4]] = np.fromfile(path, dtype="<u8")
        with open(os.path.join(tmp, "manifest.json")) as f:
            scalars = json.load(f)
    np.savez_compressed(os.path.join(HERE, "reference_goldens.npz"), **arrays)
    with open(os.path.join(HERE, "reference_scalars.json"), "w") as f:
        json.dump(scalars, f, indent=1, sort_keys=True)
    print(f"wrote {len(arrays)} arrays, {len(scalars)} scalars")


if __name__ == "__main__":
    sys.exit(main())
