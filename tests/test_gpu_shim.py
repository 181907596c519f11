"""The host-side bssnRHS / bssnrhs / physical_constraints mirror
(dendro-5.01_amd/host/), driven
from C++ like the reference's rhsTestBed: host arrays (staged through the
device), device arrays, and block-by-block bssnrhs, each against the CPU
oracle at 1e-11 relative per variable, with padding untouched."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))

pytestmark = pytest.mark.gpu


def test_host_shim_matches_oracle(cuda):
    exe = os.path.join(HERE, "native", "shim_test")
    assert os.path.exists(exe), "tests/native/shim_test not built (python dendro-5.01_amd/build.py --all)"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("max rel err") == 4
