"""The oracle's constraint monitor (oracle/physcon_oracle.cpp) pinned against
the reference: golden vectors written by the reference's own generated
equations (BSSN_GR/src/physconeqs.cpp, tests/golden/physcon_*.npz), and
closed forms that fix the block-level derivative wiring:

  * polynomial data (degree <= 3, no faces; degree <= 2 with every face
    flagged) where deriv42 and its closures are exact: the block result
    equals the point equations fed with the analytic derivatives;
  * flat space: every output exactly 0;
  * conformally flat, time-symmetric data (Brill-Lindquist): momentum
    constraints exactly 0, Hamiltonian constraint -> 0 at fourth order.
"""
import os

import numpy as np
import pytest

import pyoracle as po
from pydgr import data, layout

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


@pytest.mark.parametrize("name", ["near_flat", "strong"])
def test_point_equations_match_reference_golden(name):
    z = np.load(os.path.join(GOLD, "physcon_%s.npz" % name))
    got = po.physcon_point(z["inp"])
    ref = z["out"]
    for c in range(6):
        scale = np.abs(ref[c]).max()
        err = np.abs(got[c] - ref[c]).max() / scale
        assert err <= 1e-13, (name, po.PHYSCON_NAMES[c], err)


def test_psi4_zero_on_the_z_axis():
    z = np.load(os.path.join(GOLD, "physcon_strong.npz"))
    on_axis = (np.abs(z["inp"][17]) <= 1e-7) & (np.abs(z["inp"][18]) <= 1e-7)
    assert on_axis.sum() > 5
    got = po.physcon_point(z["inp"])
    assert np.all(got[4:6, on_axis] == 0.0) and np.all(z["out"][4:6, on_axis] == 0.0)
    assert np.all(got[4:6, ~on_axis] != 0.0)


# state order of the constraint inputs and their VAR slots
STATE = ["chi", "K", "Gt0", "Gt1", "Gt2", "gt0", "gt1", "gt2", "gt3", "gt4", "gt5",
         "At0", "At1", "At2", "At3", "At4", "At5"]
VAR_SLOT = {"chi": 1, "K": 2, "Gt0": 3, "Gt1": 4, "Gt2": 5, "gt0": 12, "gt1": 13, "gt2": 14, "gt3": 15,
            "gt4": 16, "gt5": 17, "At0": 18, "At1": 19, "At2": 20, "At3": 21, "At4": 22, "At5": 23}
PAIRS = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]


def _poly_state(rng, deg):
    """Per state variable: base value + small random polynomial of total
    degree <= deg; returns callables for value, gradient, Hessian."""
    base = {v: 0.0 for v in STATE}
    base["chi"] = 0.8
    base["gt0"] = base["gt3"] = base["gt5"] = 1.0
    terms = []
    for d in range(deg + 1):
        for a in range(d + 1):
            for b in range(d + 1 - a):
                terms.append((a, b, d - a - b))
    coef = {v: rng.standard_normal(len(terms)) * 0.002 for v in STATE}

    def value(v, x, y, z):
        s = base[v]
        for c, (a, b, e) in zip(coef[v], terms):
            s = s + c * x ** a * y ** b * z ** e
        return s

    def grad(v, x, y, z, i):
        s = 0.0
        for c, (a, b, e) in zip(coef[v], terms):
            p = [a, b, e]
            if p[i] == 0:
                continue
            q = list(p)
            q[i] -= 1
            s = s + c * p[i] * x ** q[0] * y ** q[1] * z ** q[2]
        return s

    def hess(v, x, y, z, i, j):
        s = 0.0
        for c, (a, b, e) in zip(coef[v], terms):
            q = [a, b, e]
            f = q[i]
            q[i] -= 1
            if f == 0 or q[j] == 0:
                continue
            f *= q[j]
            q[j] -= 1
            s = s + c * f * x ** q[0] * y ** q[1] * z ** q[2]
        return s

    return value, grad, hess


@pytest.mark.parametrize("deg,bflag", [(3, 0), (2, 0x3F)])
def test_block_wiring_on_polynomials(deg, bflag):
    rng = np.random.default_rng(7 + deg)
    blocks, dof = layout.uniform_grid(1, dx=0.1, center=False, outer_bflags=False)
    blocks[0]["bflag"] = bflag
    for d in range(3):  # off the z axis so that psi4 is evaluated
        blocks[0]["pmin"][d] += 0.7 + 0.3 * d
        blocks[0]["pmax"][d] += 0.7 + 0.3 * d
    value, grad, hess = _poly_state(rng, deg)
    X, Y, Z, idx = [c.numpy() for c in data.coords(blocks)]
    U = np.zeros((24, dof))
    U[0] = 1.0
    for v in STATE:
        U[VAR_SLOT[v], idx] = value(v, X, Y, Z)
    pos = np.empty(dof, dtype=np.int64)
    pos[idx] = np.arange(dof)
    X, Y, Z = X[pos], Y[pos], Z[pos]  # coordinates by flat index
    got = po.physcon_blocks(U, blocks)
    ii = layout.interior_index(blocks)
    x, y, z = X[ii], Y[ii], Z[ii]
    inp = [value(v, x, y, z) for v in STATE] + [x, y, z]
    inp += [grad(v, x, y, z, i) for v in STATE for i in range(3)]
    inp += [hess(v, x, y, z, i, j) for v in ["chi", "gt0", "gt1", "gt2", "gt3", "gt4", "gt5"] for (i, j) in PAIRS]
    inp = np.array([np.broadcast_to(np.asarray(a, dtype=np.float64), x.shape) for a in inp])
    ref = po.physcon_point(inp)
    for c in range(6):
        err = np.abs(got[c, ii] - ref[c]).max() / np.abs(ref[c]).max()
        assert err <= 1e-9, (deg, bflag, po.PHYSCON_NAMES[c], err)


def test_flat_space_is_zero():
    blocks, dof = layout.uniform_grid(2, dx=1.0 / 16, outer_bflags=True)
    U = data.flat_bumps(blocks, dof, amp=0.0).numpy()
    C = po.physcon_blocks(U, blocks)
    ii = layout.interior_index(blocks)
    assert np.all(C[:, ii] == 0.0)


def test_brill_lindquist_constraints():
    """Time-symmetric conformally flat data: mom = 0 exactly (At = K = Gt =
    0); ham is a discretisation error that falls ~16x per halving of h."""
    errs = []
    for dx in (0.1, 0.05):
        blocks, dof = layout.uniform_grid(1, dx=dx, center=False, outer_bflags=False)
        for d, c in enumerate((1.3, 0.9, 0.7)):
            blocks[0]["pmin"][d] += c
            blocks[0]["pmax"][d] += c
        U = data.brill_lindquist(blocks, dof, m1=0.5, m2=0.0, sep=0.0, offset=(0.0, 0.0, 0.0)).numpy()
        C = po.physcon_blocks(U, blocks)
        ii = layout.interior_index(blocks)
        assert np.all(C[1:4, ii] == 0.0)
        # the same physical point set: the block's first interior point
        errs.append(np.abs(C[0, ii]).max())
    ratio = errs[0] / errs[1]
    assert ratio > 10.0, (errs, ratio)
