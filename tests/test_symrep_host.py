"""Host side of the symmetric-power engine (gauge_higgs_amd/csrc/symrep.cpp),
compiled from the product sources with g++ and run on the CPU.

For every irrep the engine checks D(U) = Yc^T (Sym^k1 U (x) Sym^k2 U) Yc / c
against the slot-by-slot tensor apply W U^(x n) W^T of the reference
(rep_general.hpp:137-159) on generic SU(N) links, and then runs its own LDS
program on the host against those matrices, for D(U) and D(U)^dag. The GPU
tests compare the device engine with the device tensor path
(test_gpu_bigrep.py) and with the oracle (test_gpu_frozen.py,
test_gpu_multi.py)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "gauge_higgs_amd", "csrc")

ENGINE = ["3:6,1", "3:2,1", "3:4,2", "3:4,4", "3:7", "3:1,1", "3:6,2", "2:12", "2:4"]


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("symrep") / "symrep_check")
    subprocess.check_call([cxx, "-O2", "-std=c++20", "-I", CSRC, os.path.join(ROOT, "tests", "native", "symrep_check.cpp"),
                           os.path.join(CSRC, "symrep.cpp"), os.path.join(CSRC, "rep_build.cpp"), "-o", exe])
    return exe


def run(exe, specs):
    out = subprocess.run([exe, *specs], capture_output=True, text=True, timeout=300, check=True).stdout.split("\n")
    rows = {}
    for line in out:
        if line.strip():
            f = line.split()
            rows[f[0]] = dict(build=int(f[1]), build_err=float(f[2]), prog=int(f[3]), prog_err=float(f[4]), npass=int(f[5]),
                              lds=int(f[6]), d=int(f[7]), ynz=int(f[8]), s2=int(f[9]), nnz=int(f[10]))
    return rows


def test_engine_matches_tensor_apply_and_program_runs(checker):
    rows = run(checker, ENGINE)
    for spec in ENGINE:
        r = rows[spec]
        assert r["build"] == 1 and r["build_err"] < 1e-10, (spec, r)
        assert r["prog"] == 1 and r["prog_err"] < 1e-12, (spec, r)
        assert r["npass"] % 2 == 0  # the device loop runs passes in pairs
        # LDS of the link kernel in FP64 (16 B per complex slot) stays far below the 160 KiB of a CU
        assert 16 * r["lds"] < 64 * 1024, (spec, r)
    assert rows["3:6,1"]["d"] == 48 and rows["3:6,2"]["d"] == 60


def test_projection_rows_are_sparse(checker):
    """The irrep's basis vectors are weight vectors, so a row of the projection
    Yc F / c touches only the monomial pairs (a, b) that split its colour
    content: at most dim Sym^k2 of them. The program check above runs the
    device's sparse rows against the dense engine."""
    rows = run(checker, ENGINE)
    for spec in ENGINE:
        r = rows[spec]
        assert 1 <= r["ynz"] <= r["s2"], (spec, r)
    # {6,1}: 48 rows of the 84 lifted components, at most 3 entries each
    assert rows["3:6,1"]["ynz"] == 3 and rows["3:6,1"]["nnz"] <= 3 * 48


def test_engine_refuses_what_it_cannot_run(checker):
    """Three-row diagrams keep the tensor path, and so does SU(4), whose links
    have more columns than an entry of the LDS program addresses."""
    rows = run(checker, ["4:2,1,1", "4:2,1"])
    assert rows["4:2,1,1"]["build"] == 0
    assert rows["4:2,1"]["build"] == 1 and rows["4:2,1"]["prog"] == 0
