"""The oracle's point equations against golden vectors written by the
reference's own generated equations (tests/golden/make_golden.py), and the
product's staged form of the same equations (host build of
dendro-5.01_amd/csrc/bssn_staged.h) against the same vectors."""
import ctypes
import glob
import os
import subprocess

import numpy as np
import pytest

import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = sorted(glob.glob(os.path.join(HERE, "golden", "point_eqs_*.npz")))
dp = ctypes.POINTER(ctypes.c_double)


def _staged_lib():
    path = os.path.join(HERE, "native", "libstaged_host.so")
    if not os.path.exists(path):
        subprocess.check_call(["g++", "-O2", "-ffp-contract=off", "-std=c++17", "-fPIC", "-shared", "-o", path,
                               os.path.join(HERE, "native", "staged_host.cpp")])
    L = ctypes.CDLL(path)
    L.staged_point_eqs.argtypes = [ctypes.c_long, dp, dp, dp, dp, dp]
    return L


def test_golden_files_present():
    assert len(GOLD) == 3


@pytest.mark.parametrize("path", GOLD, ids=lambda p: os.path.basename(p))
def test_oracle_point_eqs_match_reference(path):
    g = np.load(path)
    prm = po.params(lambda_=g["lam"], lambda_f=g["lamf"])
    out = po.point_eqs(g["inp"], g["eta"], prm)
    ref = g["out"]
    rel = np.abs(out - ref).max(1) / np.abs(ref).max(1)
    assert rel.max() < 1e-13, rel


@pytest.mark.parametrize("path", GOLD, ids=lambda p: os.path.basename(p))
def test_staged_point_eqs_match_reference(path):
    g = np.load(path)
    L = _staged_lib()
    n = g["inp"].shape[1]
    out = np.zeros((24, n))
    f = lambda a: np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(dp)
    L.staged_point_eqs(n, f(g["inp"]), f(g["eta"]), f(g["lam"]), f(g["lamf"]), out.ctypes.data_as(dp))
    ref = g["out"]
    rel = np.abs(out - ref).max(1) / np.abs(ref).max(1)
    assert rel.max() < 1e-13, rel


def test_oracle_detects_planted_error():
    """A dropped term in the expected values is caught: perturbing one input
    slot (grad_0_alpha) changes several right-hand sides by far more than
    the tolerance."""
    g = np.load(GOLD[0])
    inp = g["inp"].copy()
    inp[24] += 1e-3
    out = po.point_eqs(inp, g["eta"], po.params(lambda_=g["lam"], lambda_f=g["lamf"]))
    rel = np.abs(out - g["out"]).max(1) / np.abs(g["out"]).max(1)
    assert (rel > 1e-6).sum() >= 3
