"""The C ABI from compiled C++ (tests/abi_cpp/abi_check.cpp): plain
column-major buffers through the adapter pattern a reference-side caller
uses, getCellNeighbors / getVelocities / stepRungeKuttaAdaptive on a
3,000-cell dome against the oracle (list identical, velocities and the RK
step within 1e-10). Runs as its own process with /opt/rocm's HIP runtime."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_caller_matches_oracle():
    exe = os.path.join(ROOT, "tests", "abi_cpp", "abi_check")
    print("abi_check 3000", flush=True)
    r = subprocess.run([exe, "3000"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and "ABI CHECK OK" in r.stdout
