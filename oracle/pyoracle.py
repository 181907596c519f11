"""ctypes loader of the CPU oracle (liboracle.so) -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, as the checker or as the timed CPU baseline.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NUM_VARS = 24
NIN = 207
dp = ctypes.POINTER(ctypes.c_double)


class Params(ctypes.Structure):
    _fields_ = [("lambda_", ctypes.c_double * 4), ("lambda_f", ctypes.c_double * 2),
                ("eta_const", ctypes.c_double), ("eta_r0", ctypes.c_double),
                ("eta_damping_exp", ctypes.c_double), ("ko_sigma", ctypes.c_double),
                ("deriv_order", ctypes.c_int), ("pad", ctypes.c_int)]


def params(lambda_=(1, 1, 1, 1), lambda_f=(1.0, 0.0), eta_const=2.0, eta_r0=30.0, eta_damping_exp=1.0,
           ko_sigma=0.4, deriv_order=42):
    p = Params()
    p.deriv_order = deriv_order
    p.lambda_[:] = [float(x) for x in lambda_]
    p.lambda_f[:] = [float(x) for x in lambda_f]
    p.eta_const, p.eta_r0, p.eta_damping_exp, p.ko_sigma = eta_const, eta_r0, eta_damping_exp, ko_sigma
    return p


_libs = {}


def load(fast=False):
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name not in _libs:
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            subprocess.check_call(["make", "-s", "all"] + (["ORACLE_FAST=1"] if fast else []), cwd=HERE)
        L = ctypes.CDLL(path)
        L.orc_deriv.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, ctypes.c_double,
                                ctypes.POINTER(ctypes.c_uint), ctypes.c_uint, dp]
        L.orc_bssnrhs_blocks.argtypes = [ctypes.POINTER(dp), ctypes.POINTER(dp), ctypes.c_uint32,
                                         ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint),
                                         ctypes.POINTER(ctypes.c_uint), dp, dp, ctypes.POINTER(Params), ctypes.c_int]
        L.orc_point_eqs.argtypes = [ctypes.c_long, dp, dp, ctypes.POINTER(Params), dp]
        L.orc_enforce_constraints.argtypes = [ctypes.POINTER(dp), ctypes.c_long, ctypes.c_double]
        L.orc_enforce_constraints.restype = ctypes.c_int
        L.orc_rk4_stage.argtypes = [dp, dp, dp, ctypes.c_double, ctypes.c_long]
        L.orc_rk4_final.argtypes = [dp, dp, ctypes.POINTER(dp), ctypes.c_double, ctypes.c_long]
        L.orc_physcon_point.argtypes = [ctypes.c_long, dp, dp]
        L.orc_puncture_points.argtypes = [ctypes.c_long, dp, dp, ctypes.c_double, dp]
        L.orc_physcon.argtypes = [ctypes.POINTER(dp), ctypes.POINTER(dp), ctypes.c_uint64, dp, dp,
                                  ctypes.POINTER(ctypes.c_uint), ctypes.c_uint]
        _libs[name] = L
    return _libs[name]


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(dp)


def _rows(U):
    return (dp * U.shape[0])(*[_p(U[v]) for v in range(U.shape[0])]) if U.ndim == 2 else None


def deriv(kind, axis, u, h, sz, bflag=0, beta=None, fill=np.nan):
    """One reference operator on one padded block (see bssn_oracle.h)."""
    L = load()
    u = np.ascontiguousarray(u, dtype=np.float64).ravel()
    out = np.full_like(u, fill)
    szc = (ctypes.c_uint * 3)(*sz)
    b = None if beta is None else _p(np.ascontiguousarray(beta, dtype=np.float64).ravel())
    L.orc_deriv(kind, axis, _p(out), _p(u), h, szc, bflag, b)
    return out


def rhs_blocks(U, blocks, prm=None, nthreads=0, fast=False, fill=np.nan, out=None):
    """bssnRHS over all blocks of a (24, dof) array; non-interior points of
    the result keep `fill` (or whatever `out` held)."""
    L = load(fast)
    prm = prm or params()
    U = np.ascontiguousarray(U, dtype=np.float64)
    R = np.full_like(U, fill) if out is None else out
    nb = len(blocks)
    offs = np.array([b["offset"] for b in blocks], dtype=np.uint64)
    sz = np.array([b["sz"] for b in blocks], dtype=np.uint32).ravel()
    bf = np.array([b["bflag"] for b in blocks], dtype=np.uint32)
    pmin = np.array([b["pmin"] for b in blocks], dtype=np.float64).ravel()
    pmax = np.array([b["pmax"] for b in blocks], dtype=np.float64).ravel()
    L.orc_bssnrhs_blocks(_rows(R), _rows(U), nb, offs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                         sz.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)),
                         bf.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)), _p(pmin), _p(pmax),
                         ctypes.byref(prm), nthreads)
    return R


def point_eqs(inp, eta, prm=None):
    L = load()
    prm = prm or params()
    inp = np.ascontiguousarray(inp, dtype=np.float64)
    n = inp.shape[1]
    out = np.zeros((NUM_VARS, n))
    L.orc_point_eqs(n, _p(inp), _p(np.ascontiguousarray(eta, dtype=np.float64)), ctypes.byref(prm), _p(out))
    return out


PHYSCON_NIN = 113
PHYSCON_NAMES = ["ham", "mom0", "mom1", "mom2", "psi4_real", "psi4_img"]


def physcon_point(inp):
    """Constraint point equations: inp (113, n) in physcon_point_io.h order
    -> (6, n) ham, mom0..2, psi4_real, psi4_img."""
    L = load()
    inp = np.ascontiguousarray(inp, dtype=np.float64)
    assert inp.shape[0] == PHYSCON_NIN
    out = np.zeros((6, inp.shape[1]))
    L.orc_physcon_point(inp.shape[1], _p(inp), _p(out))
    return out


def physcon_blocks(U, blocks, fill=np.nan):
    """physical_constraints over all blocks of a (24, dof) array -> (6, dof);
    non-interior points keep `fill`."""
    L = load()
    U = np.ascontiguousarray(U, dtype=np.float64)
    C = np.full((6, U.shape[1]), fill)
    for b in blocks:
        szc = (ctypes.c_uint * 3)(*b["sz"])
        L.orc_physcon(_rows(C), _rows(U), b["offset"], _p(np.asarray(b["pmin"], dtype=np.float64)),
                      _p(np.asarray(b["pmax"], dtype=np.float64)), szc, b["bflag"])
    return C


def puncture_holes(holes):
    """Two hole dicts (mass, pos, p, spin, spin_theta, spin_phi) -> the
    (20,) array orc_puncture_point takes."""
    a = []
    for h in holes:
        a += [h["mass"], *h["pos"], *h.get("p", (0.0, 0.0, 0.0)), h.get("spin", 0.0), h.get("spin_theta", 0.0),
              h.get("spin_phi", 0.0)]
    return np.asarray(a, dtype=np.float64)


def puncture_points(xyz, holes, chi_floor=0.1):
    """punctureData at points xyz (3, n) -> (24, n)."""
    L = load()
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    n = xyz.shape[1]
    out = np.zeros((24, n))
    L.orc_puncture_points(n, _p(xyz), _p(puncture_holes(holes)), chi_floor, _p(out))
    return out


def enforce_constraints(U, chi_floor=0.1):
    L = load()
    U = np.ascontiguousarray(U, dtype=np.float64).copy()
    bad = L.orc_enforce_constraints(_rows(U), U.shape[1], chi_floor)
    return U, bad


def rk4_stage(u_prev, k, cdt):
    L = load()
    out = np.empty_like(u_prev)
    L.orc_rk4_stage(_p(out), _p(u_prev), _p(k), cdt, u_prev.size)
    return out


def rk4_final(u_prev, ks, dt):
    L = load()
    out = np.empty_like(u_prev)
    arr = (dp * 4)(*[_p(k) for k in ks])
    L.orc_rk4_final(_p(out), _p(u_prev), arr, dt, u_prev.size)
    return out


def ghost_refill_uniform(U, nb, sz=13):
    """Same-level ghost refill on the nb^3 uniform block grid of
    pydgr.layout.uniform_grid (blocks in x-fastest order, corners sz-7 points
    apart): every ghost box of every block takes the neighbour's interior
    values; boxes without a neighbour (domain boundary) are left as they are."""
    E = sz - 7
    V = U.reshape(U.shape[0], nb, nb, nb, sz, sz, sz)  # [var, bz, by, bx, k, j, i]

    def rng(d):
        return {-1: (0, 3), 0: (3, sz - 3), 1: (sz - 3, sz)}[d]

    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                if dx == dy == dz == 0:
                    continue
                b = [slice(max(0, -d), nb - max(0, d)) for d in (dz, dy, dx)]
                s = [slice(max(0, d), nb - max(0, -d)) for d in (dz, dy, dx)]
                (k0, k1), (j0, j1), (i0, i1) = rng(dz), rng(dy), rng(dx)
                V[:, b[0], b[1], b[2], k0:k1, j0:j1, i0:i1] = \
                    V[:, s[0], s[1], s[2], k0 - dz * E:k1 - dz * E, j0 - dy * E:j1 - dy * E, i0 - dx * E:i1 - dx * E]
    return U


def ghost_refill_blocks(U, blocks):
    r"""Same-level ghost refill on any block list, found from the geometry
    alone: a ghost box of block b (direction d in {-1,0,1}^3 \ 0) takes the
    interior values of the block of equal size, spacing and group whose
    interior corner lies (sz-7) spacings away along every axis with d != 0
    (the copy Mesh::unzip makes between same-level neighbours,
    include/mesh.tcc:7436).  Boxes without such a neighbour are left as they
    are.  Blocks must be on a common lattice per size (no partial overlap)."""
    key = {}
    for n, b in enumerate(blocks):
        h = (b["pmax"][0] - b["pmin"][0]) / (b["sz"][0] - 1)
        c = tuple(int(round(b["pmin"][a] / h)) for a in range(3))
        key[(tuple(b["sz"]), round(h, 15), b.get("group", 0)) + c] = n
    for n, b in enumerate(blocks):
        s0, s1, s2 = b["sz"]
        h = (b["pmax"][0] - b["pmin"][0]) / (s0 - 1)
        c = [int(round(b["pmin"][a] / h)) for a in range(3)]
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    if dx == dy == dz == 0:
                        continue
                    d = (dx, dy, dz)
                    k = (tuple(b["sz"]), round(h, 15), b.get("group", 0)) + tuple(
                        c[a] + d[a] * (b["sz"][a] - 7) for a in range(3))
                    m = key.get(k)
                    if m is None:
                        continue
                    src = blocks[m]
                    V = U[:, b["offset"]:b["offset"] + s0 * s1 * s2].reshape(U.shape[0], s2, s1, s0)
                    W = U[:, src["offset"]:src["offset"] + s0 * s1 * s2].reshape(U.shape[0], s2, s1, s0)
                    rng = []
                    for a, sa in zip((dz, dy, dx), (s2, s1, s0)):
                        lo, hi = {-1: (0, 3), 0: (3, sa - 3), 1: (sa - 3, sa)}[a]
                        rng.append((lo, hi, a * (sa - 7)))
                    (k0, k1, ek), (j0, j1, ej), (i0, i1, ei) = rng
                    V[:, k0:k1, j0:j1, i0:i1] = W[:, k0 - ek:k1 - ek, j0 - ej:j1 - ej, i0 - ei:i1 - ei]
    return U


def rk4_step_blocks(U, blocks, dt, refill, prm=None, chi_floor=0.1, nthreads=0):
    """One RK4 step of rkBSSN.cpp:803-1011 on unzipped blocks:
    stage combines u_prev + (U[s+1]*dt)*k_s with the constraint projection,
    the final u_prev + sum_s (C[s]*dt)*k_s in order s = 0..3, and the ghost
    refill `refill(U)` after every stage.  Returns (u_new, bad)."""
    from pydgr import layout  # noqa: WPS433 (test infrastructure)
    Uc = [0.0, 0.5, 0.5, 1.0]
    Cc = [1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0]
    ii = layout.interior_index(blocks)
    uprev = np.ascontiguousarray(U, dtype=np.float64)
    stage = uprev.copy()
    cur = uprev
    acc = None
    bad = 0
    for s in range(4):
        k = rhs_blocks(cur, blocks, prm, nthreads=nthreads, fill=0.0)
        kk = k[:, ii]
        up = uprev[:, ii]
        if s == 0:
            acc = up + (Cc[0] * dt) * kk
        else:
            acc = acc + (Cc[s] * dt) * kk
        if s < 3:
            un = up + (Uc[s + 1] * dt) * kk
            un, b = enforce_constraints(np.ascontiguousarray(un), chi_floor)
            bad |= b
            stage[:, ii] = un
            refill(stage)
            cur = stage
    unew = uprev.copy()
    fin, b = enforce_constraints(np.ascontiguousarray(acc), chi_floor)
    bad |= b
    unew[:, ii] = fin
    refill(unew)
    return unew, bad


def rk4_step_uniform(U, blocks, nb, dt, prm=None, chi_floor=0.1, nthreads=0):
    """rk4_step_blocks on the nb^3 uniform grid of pydgr.layout.uniform_grid."""
    return rk4_step_blocks(U, blocks, dt, lambda A: ghost_refill_uniform(A, nb), prm, chi_floor, nthreads)
