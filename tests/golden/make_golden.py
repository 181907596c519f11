"""Write the golden point-equation vectors from the reference's own
generated equations, compiled by oracle/Makefile `make ref`:

  point_eqs_*.npz  BSSN_GR/src/bssneqs_eta_const_standard_gauge.cpp
                   (oracle/_ref/libref_eqs.so): the 207 point inputs in the
                   order of oracle/bssn_point_io.h, the per-point eta,
                   lambda / lambda_f, and the 24 right-hand sides;
  physcon_*.npz    BSSN_GR/src/physconeqs.cpp (oracle/_ref/libref_physcon.so):
                   the 113 point inputs in the order of
                   oracle/physcon_point_io.h and the 6 constraint outputs
                   (ham, mom0..2, psi4_real, psi4_img).

    python tests/golden/make_golden.py

Inputs are seeded and stored, so a fixture does not depend on a
random-number stream.
"""
import ctypes
import hashlib
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libref_eqs.so")
REF_PHYSCON = os.path.join(ROOT, "oracle", "_ref", "libref_physcon.so")
NIN = 207
NIN_PHYSCON = 113
dp = ctypes.POINTER(ctypes.c_double)


def ref_point_eqs(inp, eta, lam, lamf):
    L = ctypes.CDLL(REF_LIB)
    L.ref_point_eqs.argtypes = [ctypes.c_long, dp, dp, dp, dp, dp]
    n = inp.shape[1]
    out = np.zeros((24, n))
    f = lambda a: np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(dp)
    L.ref_point_eqs(n, f(inp), f(eta), f(np.asarray(lam, float)), f(np.asarray(lamf, float)), out.ctypes.data_as(dp))
    return out


def inputs(rng, n, scale, chi_lo, chi_hi):
    x = rng.standard_normal((NIN, n)) * scale
    x[0] += 1.0                                  # alpha
    x[1] = rng.uniform(chi_lo, chi_hi, n)        # chi
    for k in (12, 15, 17):                       # gt diagonal
        x[k] += 1.0
    return x


SETS = {
    # perturbed flat space, as in the flat + Gaussian-bump workload
    "near_flat": dict(seed=11, n=256, scale=1e-3, chi=(0.999, 1.001), lam=(1, 1, 1, 1), lamf=(1.0, 0.0), eta=2.0),
    # strongly curved: conformal factor down to 0.05, O(0.1) metric terms
    "strong": dict(seed=12, n=256, scale=1e-1, chi=(0.05, 1.0), lam=(1, 1, 1, 1), lamf=(1.0, 0.0), eta=2.0),
    # other gauge parameters and a point-dependent eta
    "params": dict(seed=13, n=128, scale=3e-2, chi=(0.3, 1.0), lam=(1, 0, 1, 0), lamf=(0.5, 0.75), eta=None),
}


def ref_physcon_point(inp):
    L = ctypes.CDLL(REF_PHYSCON)
    L.ref_physcon_point.argtypes = [ctypes.c_long, dp, dp]
    n = inp.shape[1]
    out = np.zeros((6, n))
    L.ref_physcon_point(n, np.ascontiguousarray(inp, dtype=np.float64).ctypes.data_as(dp), out.ctypes.data_as(dp))
    return out


def physcon_inputs(rng, n, scale, chi_lo, chi_hi, coord, axis_frac=0.0):
    x = rng.standard_normal((NIN_PHYSCON, n)) * scale
    x[0] = rng.uniform(chi_lo, chi_hi, n)        # chi
    for k in (5, 8, 10):                         # gt diagonal
        x[k] += 1.0
    x[17:20] = rng.uniform(-coord, coord, (3, n))  # point coordinates
    m = rng.random(n) < axis_frac                 # points on the z axis: psi4 = 0 there
    x[17, m] = 0.0
    x[18, m] = 0.0
    return x


PHYSCON_SETS = {
    # perturbed flat space at coordinates of a 64-block domain
    "near_flat": dict(seed=21, n=256, scale=1e-3, chi=(0.999, 1.001), coord=1.0, axis=0.1),
    # strong field near a puncture
    "strong": dict(seed=22, n=256, scale=1e-1, chi=(0.05, 1.0), coord=8.0, axis=0.1),
}


def main():
    manifest = {}
    for name, c in SETS.items():
        rng = np.random.default_rng(c["seed"])
        inp = inputs(rng, c["n"], c["scale"], *c["chi"])
        eta = np.full(c["n"], c["eta"]) if c["eta"] is not None else rng.uniform(0.5, 2.5, c["n"])
        out = ref_point_eqs(inp, eta, c["lam"], c["lamf"])
        path = os.path.join(HERE, "point_eqs_%s.npz" % name)
        np.savez(path, inp=inp, eta=eta, lam=np.asarray(c["lam"], float), lamf=np.asarray(c["lamf"], float),
                 out=out)
        manifest[os.path.basename(path)] = hashlib.sha256(open(path, "rb").read()).hexdigest()
        print(name, "max|out| per var", np.abs(out).max(1).round(4).tolist())
    for name, c in PHYSCON_SETS.items():
        rng = np.random.default_rng(c["seed"])
        inp = physcon_inputs(rng, c["n"], c["scale"], *c["chi"], c["coord"], c["axis"])
        out = ref_physcon_point(inp)
        path = os.path.join(HERE, "physcon_%s.npz" % name)
        np.savez(path, inp=inp, out=out)
        manifest[os.path.basename(path)] = hashlib.sha256(open(path, "rb").read()).hexdigest()
        print("physcon", name, "max|out| per output", np.abs(out).max(1).round(4).tolist())
    with open(os.path.join(HERE, "MANIFEST.json"), "w") as f:
        json.dump({"source": {"point_eqs_*": "BSSN_GR/src/bssneqs_eta_const_standard_gauge.cpp via "
                                             "oracle/_ref/libref_eqs.so",
                              "physcon_*": "BSSN_GR/src/physconeqs.cpp via oracle/_ref/libref_physcon.so"},
                   "files": manifest}, f, indent=1)


if __name__ == "__main__":
    main()
