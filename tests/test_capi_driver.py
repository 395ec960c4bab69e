"""The C ABI from a C++ host (tests/capi_driver.cpp, built by __graft_entry__.build()):
dlopen + dlsym of every entry point it uses, its own operators, mesh and descriptor, no
Python host code.  On the CPU it checks loading, the version string and the argument
errors of flou_hip_create; on the GPU (test_gpu_capi_driver) the full run."""
import os
import subprocess

import pytest

from flouhip import _capi as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "bin", "capi_driver")

needs_driver = pytest.mark.skipif(not os.path.exists(DRIVER), reason="tests/bin/capi_driver not built (build())")


def run_driver(*args, timeout=120):
    p = subprocess.run([DRIVER, C.LIB_PATH, *args], capture_output=True, text=True, timeout=timeout)
    print(p.stdout)
    print(p.stderr)
    return p


@needs_driver
def test_driver_loads_library_and_checks_arguments_without_gpu():
    p = run_driver("--no-gpu")
    assert p.returncode == 0
    assert "gfx950" in p.stdout and "0 failure(s)" in p.stdout
    assert "gamma" in p.stdout.lower()


@needs_driver
@pytest.mark.gpu
def test_gpu_capi_driver_full_run():
    p = run_driver()
    assert p.returncode == 0, p.stdout + p.stderr
    assert "advection: full period returns" in p.stdout and "C ABI driver: 0 failure(s)" in p.stdout
