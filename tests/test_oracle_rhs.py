"""Whole-block properties of the oracle's bssnrhs restatement
(orc_bssnrhs, BSSN_GR/src/rhs.cpp:86-392): flat-space invariance,
manufactured solutions that exercise the derivative wiring (first, mixed
second and upwinded derivatives), axis-permutation symmetry, the Sommerfeld
boundary formula, untouched padding, plus the constraint projection and the
RK4 combine."""
import numpy as np
import pytest

import pyoracle as po
from pydgr import data, layout

from helpers import rel_err_per_var

SG0, SA0 = 12, 18


def coords_np(blocks):
    """Coordinates of every padded point; data.coords lists blocks in offset
    order, so position k holds flat index k."""
    x, y, z, idx = data.coords(blocks)
    assert np.array_equal(idx.numpy(), np.arange(len(idx)))
    return x.numpy(), y.numpy(), z.numpy(), idx.numpy()


@pytest.mark.parametrize("order", [42, 64])
def test_flat_space_is_exactly_static(order):
    blocks, dof = layout.uniform_grid(2, dx=1.0 / 64)
    U = data.flat_bumps(blocks, dof, amp=0.0).numpy()
    R = po.rhs_blocks(U, blocks, po.params(deriv_order=order), fill=0.0)
    assert np.abs(R).max() == 0.0


def test_padding_untouched():
    blocks, dof = layout.uniform_grid(2, dx=1.0 / 64)
    U = data.flat_bumps(blocks, dof).numpy()
    R = po.rhs_blocks(U, blocks, fill=-7.0)
    ii = layout.interior_index(blocks)
    mask = np.ones(dof, bool)
    mask[ii] = False
    assert np.all(R[:, mask] == -7.0)
    assert np.all(np.isfinite(R[:, ii]))


@pytest.mark.parametrize("order", [42, 64])
def test_manufactured_shift_terms(order):
    """gt = delta, chi = alpha = 1, K = At = Gt = B = 0 and a bilinear shift
    beta = s (xy, yz, zx).  Then (CodeGen/bssn.py:60-82)
      b_rhs^i  = beta^k d_k beta^i
      gt_rhs_ij = d_j beta^i + d_i beta^j - 2/3 delta_ij div beta
      chi_rhs  = -2/3 div beta
      Gt_rhs^i = lap beta^i + 1/3 d_i div beta = s/3 (1,1,1)  (mixed derivatives)
      B_rhs^i  = Gt_rhs^i
    and everything else vanishes; all operators are exact on these
    polynomials."""
    s = 0.1
    blocks, dof = layout.uniform_grid(2, dx=1.0 / 64, outer_bflags=False)
    x, y, z, idx = coords_np(blocks)
    U = data.flat_bumps(blocks, dof, amp=0.0).numpy()
    U[6, idx], U[7, idx], U[8, idx] = s * x * y, s * y * z, s * z * x
    R = po.rhs_blocks(U, blocks, po.params(ko_sigma=0.0, deriv_order=order), fill=0.0)
    ii = layout.interior_index(blocks)
    X, Y, Z = x[ii], y[ii], z[ii]
    b = [s * X * Y, s * Y * Z, s * Z * X]
    db = [[s * Y, s * X, 0 * X], [0 * X, s * Z, s * Y], [s * Z, 0 * X, s * X]]  # db[i][k] = d_k beta^i
    div = db[0][0] + db[1][1] + db[2][2]
    want = np.zeros((24, len(ii)))
    for i in range(3):
        want[6 + i] = sum(b[k] * db[i][k] for k in range(3))
        want[3 + i] = s / 3.0
        want[9 + i] = s / 3.0
    for p, (i, j) in enumerate([(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]):
        want[SG0 + p] = db[i][j] + db[j][i] - (2.0 / 3.0) * (i == j) * div
    want[1] = -(2.0 / 3.0) * div
    np.testing.assert_allclose(R[:, ii], want, atol=1e-13)


@pytest.mark.parametrize("order", [42, 64])
def test_manufactured_lapse_terms(order):
    """alpha a cubic, K constant, beta constant, flat metric:
    a_rhs = lambda0 beta.grad(alpha) - 2 alpha K, K_rhs = -lap(alpha) + alpha K^2/3."""
    blocks, dof = layout.uniform_grid(2, dx=1.0 / 64, outer_bflags=False)
    x, y, z, idx = coords_np(blocks)
    U = data.flat_bumps(blocks, dof, amp=0.0).numpy()
    a = 1.0 + 0.01 * (x ** 3 + 2 * x * y * z - y ** 2 + 0.5 * z)
    U[0, idx] = a
    U[2, idx] = 0.2
    beta = (0.1, -0.2, 0.05)
    for k in range(3):
        U[6 + k, idx] = beta[k]
    R = po.rhs_blocks(U, blocks, po.params(ko_sigma=0.0, deriv_order=order), fill=0.0)
    ii = layout.interior_index(blocks)
    X, Y, Z = x[ii], y[ii], z[ii]
    A = 1.0 + 0.01 * (X ** 3 + 2 * X * Y * Z - Y ** 2 + 0.5 * Z)
    ga = 0.01 * np.array([3 * X ** 2 + 2 * Y * Z, 2 * X * Z - 2 * Y, 2 * X * Y + 0.5])
    lap = 0.01 * (6 * X - 2.0)
    np.testing.assert_allclose(R[0, ii], sum(beta[k] * ga[k] for k in range(3)) - 2 * A * 0.2, atol=1e-13)
    np.testing.assert_allclose(R[2, ii], -lap + A * 0.04 / 3.0, atol=1e-12)


def test_sixth_order_option_changes_operators():
    """deriv_order 64 switches the operator set: same data, different RHS,
    and for a degree-7 lapse only the 6th-order K_rhs is exact."""
    blocks, dof = layout.uniform_grid(2, dx=1.0 / 16, outer_bflags=False)
    x, y, z, idx = coords_np(blocks)
    U = data.flat_bumps(blocks, dof, amp=0.0).numpy()
    U[0, idx] = 1.0 + 0.01 * x ** 7
    ii = layout.interior_index(blocks)
    X = x[ii]
    want = -0.01 * 42 * X ** 5  # K_rhs = -lap(alpha) with K = 0
    got = {o: po.rhs_blocks(U, blocks, po.params(ko_sigma=0.0, deriv_order=o), fill=0.0)[2, ii] for o in (42, 64)}
    err42 = np.abs(got[42] - want).max()
    err64 = np.abs(got[64] - want).max()
    assert err64 < 1e-11 < err42


PERM_XY = {3: 4, 4: 3, 6: 7, 7: 6, 9: 10, 10: 9, 12: 15, 15: 12, 14: 16, 16: 14, 18: 21, 21: 18, 20: 22, 22: 20}


@pytest.mark.parametrize("order", [42, 64])
def test_axis_swap_symmetry(order):
    """Swapping x and y in the data (including vector/tensor components and
    the boundary flags) swaps them in the right-hand side."""
    blocks, dof = layout.uniform_grid(1, dx=1.0 / 64)
    blocks[0]["bflag"] = 0b000101  # -x, -y faces
    U = data.flat_bumps(blocks, dof, seed=3, amp=2e-2, width=0.2).numpy()
    # mirror the beta signs too so the upwind choice is exercised both ways
    prm = po.params(deriv_order=order)
    R = po.rhs_blocks(U, blocks, prm, fill=0.0)
    sz = 13
    perm = [PERM_XY.get(v, v) for v in range(24)]
    U2 = U[perm].reshape(24, sz, sz, sz).transpose(0, 1, 3, 2).reshape(24, -1).copy()
    b2 = [dict(blocks[0])]
    b2[0]["bflag"] = 0b000101
    R2 = po.rhs_blocks(U2, b2, prm, fill=0.0)
    R2b = R2.reshape(24, sz, sz, sz).transpose(0, 1, 3, 2).reshape(24, -1)[perm]
    ii = layout.interior_index(blocks)
    # mixed derivatives compose d_y(d_x u) in both runs, i.e. d_x(d_y u) of
    # the swapped data: equal up to rounding, larger for the 6th-order weights
    assert rel_err_per_var(R2b, R, ii).max() < (1e-13 if order == 42 else 1e-12)


def test_sommerfeld_faces():
    """Data linear in x, y, z (every derivative operator, closures included,
    is exact and the dissipation vanishes): on flagged faces the right-hand
    side is -(1/r)(x.grad f + falloff (f - f_inf)) (BSSN_GR/src/rhs.cpp:838-982)."""
    blocks, dof = layout.uniform_grid(1, dx=1.0 / 16, center=False)
    for d in range(3):
        blocks[0]["pmin"][d] += 0.3
        blocks[0]["pmax"][d] += 0.3
    blocks[0]["bflag"] = 63
    x, y, z, idx = coords_np(blocks)
    rng = np.random.default_rng(5)
    U = data.flat_bumps(blocks, dof, amp=0.0).numpy()
    coef = rng.uniform(-0.05, 0.05, size=(24, 4))
    for v in range(24):
        U[v, idx] += coef[v, 0] + coef[v, 1] * x + coef[v, 2] * y + coef[v, 3] * z
    R = po.rhs_blocks(U, blocks, fill=0.0)
    fall = np.array([1, 1, 1, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2], float)
    asym = np.array([1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0, 1, 0, 0, 0, 0, 0, 0], float)
    X, Y, Z = x, y, z
    sz = 13
    kk, jj, ii_ = np.meshgrid(np.arange(sz), np.arange(sz), np.arange(sz), indexing="ij")
    face = ((ii_ == 3) | (ii_ == sz - 4) | (jj == 3) | (jj == sz - 4) | (kk == 3) | (kk == sz - 4))
    inner = (ii_ >= 3) & (ii_ < sz - 3) & (jj >= 3) & (jj < sz - 3) & (kk >= 3) & (kk < sz - 3)
    sel = (face & inner).ravel()
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)[sel]
    for v in range(24):
        f = U[v][sel]
        want = -(X[sel] * coef[v, 1] + Y[sel] * coef[v, 2] + Z[sel] * coef[v, 3] + fall[v] * (f - asym[v])) / r
        np.testing.assert_allclose(R[v][sel], want, rtol=1e-10, atol=1e-13)


def test_constraint_projection_invariants():
    """After enforce_bssn_constraints (bssn_constraints.cpp:12-172):
    det(gt) = 1, gt is the input metric rescaled, At is trace-free and
    differs from the input by a multiple of gt, and the floors hold."""
    rng = np.random.default_rng(7)
    n = 500
    M = rng.standard_normal((n, 3, 3)) * 0.2
    G = np.eye(3)[None] * rng.uniform(0.5, 2.0, (n, 1, 1)) + M @ M.transpose(0, 2, 1)
    A = rng.standard_normal((n, 3, 3)) * 0.1
    A = A + A.transpose(0, 2, 1)
    U = np.zeros((24, n))
    pairs = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    for p, (i, j) in enumerate(pairs):
        U[SG0 + p] = G[:, i, j]
        U[SA0 + p] = A[:, i, j]
    U[0] = rng.uniform(-0.5, 1.5, n)
    U[1] = rng.uniform(-0.5, 1.5, n)
    V, bad = po.enforce_constraints(U, 0.1)
    assert bad == 0
    G2 = np.zeros((n, 3, 3))
    A2 = np.zeros((n, 3, 3))
    for p, (i, j) in enumerate(pairs):
        G2[:, i, j] = G2[:, j, i] = V[SG0 + p]
        A2[:, i, j] = A2[:, j, i] = V[SA0 + p]
    np.testing.assert_allclose(np.linalg.det(G2), 1.0, rtol=1e-12)
    np.testing.assert_allclose(G2, G / np.cbrt(np.linalg.det(G))[:, None, None], rtol=1e-12)
    tr = np.einsum("nij,nij->n", np.linalg.inv(G2), A2)
    assert np.abs(tr).max() < 1e-13
    D = A2 - A
    lam = D[:, 0, 0] / G2[:, 0, 0]
    np.testing.assert_allclose(D, lam[:, None, None] * G2, atol=1e-13)
    assert V[1].min() >= 0.1 and V[0].min() >= 0.1
    np.testing.assert_array_equal(V[0], np.maximum(U[0], 0.1))


def test_constraint_projection_flags_bad_metric():
    U = np.zeros((24, 1))
    U[SG0], U[SG0 + 3], U[SG0 + 5] = 1.0, 1.0, -1.0
    _, bad = po.enforce_constraints(U)
    assert bad == 1


def test_rk4_combine_matches_stability_polynomial():
    """RK4 (rkBSSN.cpp:895-1011, tableau rkBSSN.h:55-58) applied to
    du/dt = lam u reproduces 1 + z + z^2/2 + z^3/6 + z^4/24, z = lam dt."""
    lam, dt = -0.7, 0.3
    u0 = np.linspace(-1, 1, 11)
    U = [0.5, 0.5, 1.0]
    ks = []
    cur = u0.copy()
    for s in range(4):
        k = lam * cur
        ks.append(k)
        if s < 3:
            cur = po.rk4_stage(u0, k, U[s] * dt)
    u1 = po.rk4_final(u0, ks, dt)
    zz = lam * dt
    np.testing.assert_allclose(u1, u0 * (1 + zz + zz ** 2 / 2 + zz ** 3 / 6 + zz ** 4 / 24), rtol=1e-14)
