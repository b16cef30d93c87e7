"""AddressSanitizer and UndefinedBehaviorSanitizer build of the host code: the
library's host-only sources (rep_build.cpp, symrep.cpp, potential.cpp,
series.cpp) and the FP64 oracle (oracle/cpuref_capi.cpp), compiled with g++
and driven by tests/native/host_sanitize.cpp on the CPU. The device code is not
part of this build (GPU sanitizers are not available on the GPU hosts)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "gauge_higgs_amd", "csrc")


def test_host_code_under_asan_and_ubsan(tmp_path):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "host_sanitize")
    srcs = [os.path.join(ROOT, "tests", "native", "host_sanitize.cpp")] + [
        os.path.join(CSRC, f) for f in ("rep_build.cpp", "symrep.cpp", "potential.cpp", "series.cpp")] + [
        os.path.join(ROOT, "oracle", "cpuref_capi.cpp")]
    subprocess.check_call([cxx, "-std=c++20", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
                           "-fno-sanitize-recover=all", "-I", CSRC, "-I", os.path.join(ROOT, "include"), "-I",
                           os.path.join(ROOT, "oracle"), *srcs, "-o", exe, "-lquadmath"])
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="1")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    print(out.stdout[-2000:], out.stderr[-4000:])
    assert out.returncode == 0
    assert "runtime error" not in out.stderr and "AddressSanitizer" not in out.stderr
    assert out.stdout.count(" ok") >= 12
