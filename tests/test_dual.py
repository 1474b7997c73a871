"""CPU: the forward-mode dual numbers of the tangent (niamh-mpcd_amd/csrc/
mpcd_dual.h) and the geometry helpers they run through (mpcd_math.h), on the
host: every value equals the plain FP64 result bit for bit, every derivative
matches a central finite difference, and the 2x2 director turns at exactly
the rate of its eigenbasis (a closed form)."""
import json
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "niamh-mpcd_amd", "csrc")


def test_dual_numbers_and_helpers(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = tmp_path / "dual_check"
    # host code only: the program calls no HIP runtime function
    cmd = [hipcc, "-O2", "-std=c++17", "-ffp-contract=off", "--offload-arch=gfx950", "-x", "hip",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I", CSRC,
           os.path.join(HERE, "dual_check.cpp"), "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["checks"] >= 40
    assert out["value_mismatch"] == 0, r.stderr
    assert out["worst_fd_rel_err"] < 1e-6, r.stderr
    assert out["director_rate_err"] < 1e-12
