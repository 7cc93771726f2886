"""The host code of graphops_amd/csrc/host.cpp (generators, CSR transpose
build, the sweeps' row planner, bf16 rounding) under AddressSanitizer and
UndefinedBehaviorSanitizer: `make -C graphops_amd/csrc sanitize` builds a
CPU-only driver (no HIP call) that checks the arrays' invariants and the
adapter's error paths, and runs it; any sanitizer report or failed check
fails the build target."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "graphops_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None,
                    reason="needs g++ and make")
def test_host_code_under_asan_ubsan():
    r = subprocess.run(["make", "-s", "-C", CSRC, "sanitize"], capture_output=True, text=True,
                       timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "host_check: 0 failure(s)" in out
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out
