"""The hand-written reverse mode of the Roe point flux (roe_point_vjp in
hermeshd_amd/csrc/hd_physics.h) against central finite differences of the
forward point function, on the host (CPU): tests/native/point_vjp_check.cpp
is built with hipcc and calls the same host+device functions the kernels
use.  Random states in every direction, on the Roe branch and on the
Lax-Friedrichs fallback (a negative-pressure side)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def test_roe_point_vjp_matches_finite_differences(tmp_path):
    if not (os.path.exists(HIPCC) or shutil.which("hipcc")):
        pytest.skip("hipcc not available")
    exe = str(tmp_path / "point_vjp_check")
    subprocess.check_call([HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature",
                           "-Xclang", "-packed-fp32-ops", os.path.join(HERE, "native", "point_vjp_check.cpp"),
                           "-o", exe], stderr=subprocess.DEVNULL)
    out = subprocess.check_output([exe]).decode().split("\n")
    rows = {line.split()[0]: [float(v) for v in line.split()[1:]] for line in out if line.strip()}
    # central differences with h = 1e-6 (relative): truncation and rounding below 1e-6
    for name in ("roe", "roe_llf"):
        assert all(e < 1e-6 for e in rows[name]), (name, rows[name])
