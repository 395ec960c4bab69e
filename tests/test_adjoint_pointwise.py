"""Point-wise adjoint physics (flou.jl_amd/csrc/adjoint.hpp, built for the CPU) against the
oracle's dual-number JVPs: physical flux, the three two-point fluxes (derivative in the
first state, and both states at once, from the states or from the per-node primitives), LxF in the face frame (both states), 1-D to 3-D,
nearly equal and strongly different states (both branches of the log-mean)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_pointwise_vjps_match_oracle_jvps(tmp_path):
    exe = tmp_path / "adjoint_point_check"
    cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "oracle"),
           "-I", os.path.join(ROOT, "flou.jl_amd", "csrc"),
           os.path.join(HERE, "adjoint_point_check.cpp"), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail("build failed:\n" + r.stderr)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    rows = [line.split() for line in out.strip().splitlines()]
    # flux, 3 two-point fluxes (one side, both sides, both sides from primitives), LxF left/right,
    # x 3 spreads
    assert len(rows) == 36
    for name, tp, spread, e1, e2, e3 in rows:
        worst = max(float(e1), float(e2), float(e3))
        print(name, tp, spread, worst)
        assert worst < 1e-12, (name, tp, spread, worst)
