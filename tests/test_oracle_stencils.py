"""Pin the oracle's finite-difference operators (orc_deriv, restating
BSSN_GR/src/derivs.cpp) by their mathematics rather than by retyping them.

For each operator and each row class (interior, each face-closure row, each
upwind side) the coefficients are extracted from impulse responses and
checked for
  * support: exactly the grid points the reference's row reads, and
  * polynomial exactness: sum_m c_m (m-q)^k equals the k-th moment of the
    derivative being approximated for every k up to the row's order.
A stencil with fixed support and maximal exactness is unique, so a wrong
coefficient, sign, index or spacing fails here.  The written index ranges
(which the mixed derivatives by composition rely on) are checked too.
"""
import math

import numpy as np
import pytest

import pyoracle as po

FIRST, SECOND, ADV, KO, FIRST6, SECOND6, ADV6 = 0, 1, 2, 3, 4, 5, 6
SZ = 13
E = SZ - 3
SMR3, SMR2, SMR1 = 59.0 / 48.0 * 64, 43.0 / 48.0 * 64, 49.0 / 48.0 * 64  # derivs.cpp:918-925, h = 1


def line_index(axis, q, a=6, b=6):
    c = [a, b]
    c.insert(axis, q)
    i, j, k = c
    return i + SZ * (j + SZ * k)


def coeffs(kind, axis, q, bflag=0, beta=0.0, h=1.0):
    c = np.zeros(SZ)
    at = line_index(axis, q)
    for m in range(SZ):
        u = np.zeros(SZ ** 3)
        u[line_index(axis, m)] = 1.0
        beta_arr = np.full(SZ ** 3, beta)
        D = po.deriv(kind, axis, u, h, (SZ, SZ, SZ), bflag, beta_arr, fill=np.nan)
        c[m] = D[at]
    return c


def moments(c, q, kmax):
    m = np.arange(SZ) - q
    return np.array([np.sum(c * m.astype(float) ** k) for k in range(kmax + 1)])


def check_row(c, q, support, deg, order):
    got = set(np.nonzero(c)[0].tolist())
    assert got == set(support), (got, support)
    mom = moments(c, q, deg)
    want = np.zeros(deg + 1)
    if order is not None:
        want[order] = math.factorial(order)
    np.testing.assert_allclose(mom, want, atol=1e-11 * max(1.0, np.abs(c).sum()))


LOW = {0: 1, 1: 4, 2: 16}
HIGH = {0: 2, 1: 8, 2: 32}


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_first_derivative_rows(axis):
    for q in range(5, E - 2):
        check_row(coeffs(FIRST, axis, q), q, [q - 2, q - 1, q + 1, q + 2], 4, 1)
    bf = LOW[axis] | HIGH[axis]
    check_row(coeffs(FIRST, axis, 3, bf), 3, [3, 4, 5], 2, 1)
    check_row(coeffs(FIRST, axis, 4, bf), 4, [3, 5], 2, 1)
    check_row(coeffs(FIRST, axis, E - 2, bf), E - 2, [E - 3, E - 1], 2, 1)
    check_row(coeffs(FIRST, axis, E - 1, bf), E - 1, [E - 3, E - 2, E - 1], 2, 1)
    # without the face flag the rows next to the face use the interior formula
    check_row(coeffs(FIRST, axis, 3), 3, [1, 2, 4, 5], 4, 1)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_second_derivative_rows(axis):
    for q in range(5, E - 2):
        check_row(coeffs(SECOND, axis, q), q, [q - 2, q - 1, q, q + 1, q + 2], 5, 2)
    bf = LOW[axis] | HIGH[axis]
    check_row(coeffs(SECOND, axis, 3, bf), 3, [3, 4, 5, 6], 3, 2)
    check_row(coeffs(SECOND, axis, 4, bf), 4, [3, 4, 5], 2, 2)
    check_row(coeffs(SECOND, axis, E - 2, bf), E - 2, [E - 3, E - 2, E - 1], 2, 2)
    check_row(coeffs(SECOND, axis, E - 1, bf), E - 1, [E - 4, E - 3, E - 2, E - 1], 3, 2)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_upwind_rows(axis):
    for q in range(6, E - 3):
        check_row(coeffs(ADV, axis, q, beta=0.7), q, list(range(q - 1, q + 4)), 4, 1)
        check_row(coeffs(ADV, axis, q, beta=-0.7), q, list(range(q - 3, q + 2)), 4, 1)
        # beta == 0 takes the "not positive" side (derivs.cpp:517)
        check_row(coeffs(ADV, axis, q, beta=0.0), q, list(range(q - 3, q + 2)), 4, 1)
    bf = LOW[axis] | HIGH[axis]
    for b in (0.7, -0.7):
        check_row(coeffs(ADV, axis, 3, bf, b), 3, [3, 4, 5], 2, 1)
        check_row(coeffs(ADV, axis, E - 1, bf, b), E - 1, [E - 3, E - 2, E - 1], 2, 1)
    check_row(coeffs(ADV, axis, 4, bf, 0.7), 4, [4, 5, 6], 2, 1)
    check_row(coeffs(ADV, axis, 4, bf, -0.7), 4, [3, 5], 2, 1)
    check_row(coeffs(ADV, axis, 5, bf, 0.7), 5, [4, 5, 6, 7, 8], 4, 1)
    check_row(coeffs(ADV, axis, 5, bf, -0.7), 5, [3, 4, 5], 2, 1)
    check_row(coeffs(ADV, axis, E - 3, bf, 0.7), E - 3, [E - 3, E - 2, E - 1], 2, 1)
    check_row(coeffs(ADV, axis, E - 3, bf, -0.7), E - 3, list(range(E - 6, E - 1)), 4, 1)
    check_row(coeffs(ADV, axis, E - 2, bf, 0.7), E - 2, [E - 3, E - 1], 2, 1)
    check_row(coeffs(ADV, axis, E - 2, bf, -0.7), E - 2, [E - 4, E - 3, E - 2], 2, 1)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_kreiss_oliger_rows(axis):
    # interior: the sixth central difference over 64 h, which annihilates
    # polynomials of degree <= 5 and maps x^6 to 720 h^5 / 64
    for q in range(6, E - 3):
        c = coeffs(KO, axis, q)
        check_row(c, q, list(range(q - 3, q + 4)), 5, None)
        assert abs(moments(c, q, 6)[6] - 720.0 / 64.0) < 1e-12
    bf = LOW[axis] | HIGH[axis]
    # one-sided rows: annihilate quadratics; normalised by the smr constants
    # of derivs.cpp:918-925 (third moment relative to the face point)
    rows = [(3, [3, 4, 5, 6], 6.0 / SMR3), (4, [3, 4, 5, 6, 7], -12.0 / SMR2), (5, list(range(3, 9)), 6.0 / SMR1)]
    for q, sup, m3 in rows:
        c = coeffs(KO, axis, q, bf)
        assert set(np.nonzero(c)[0].tolist()) == set(sup)
        mom = moments(c, 3, 3)
        np.testing.assert_allclose(mom[:3], 0.0, atol=1e-14)
        assert abs(mom[3] - m3) < 1e-14
    # mirror rows at the far face
    for q, sup, m3 in rows:
        qq = E - 1 - (q - 3)
        c = coeffs(KO, axis, qq, bf)
        assert set(np.nonzero(c)[0].tolist()) == set(E - 1 - (s - 3) for s in sup)
        mom = moments(c, E - 1, 3)
        np.testing.assert_allclose(mom[:3], 0.0, atol=1e-14)
        assert abs(mom[3] + m3) < 1e-14


def sbp_dissipation(n):
    """-(D3^T D3) on a half line of n points: D3 is the undivided third
    difference (rows [-1, 3, -3, 1]), the construction of the narrow-stencil
    artificial dissipation of Mattsson, Svard and Nordstrom (J. Sci. Comput.
    21, 2004).  Its interior rows are minus the sixth difference, its first
    three rows the boundary-modified ones."""
    D3 = np.zeros((n - 3, n))
    for i in range(n - 3):
        D3[i, i:i + 4] = [-1.0, 3.0, -3.0, 1.0]
    return -(D3.T @ D3)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_kreiss_oliger_rows_by_construction(axis):
    """Every ko_deriv42 row (BSSN_GR/src/derivs.cpp:914-1077) equals a row of
    -(D3^T D3) built independently above: interior rows divided by 64 h, the
    three closure rows next to a flagged face divided by the weights
    smr3 / smr2 / smr1 (= 59/48, 43/48, 49/48 x 64 h, derivs.cpp:918-925),
    and the far-face rows the mirror images.  This fixes every coefficient,
    including the ones the moment conditions leave free in the 5- and 6-point
    closure rows."""
    n = E - 3  # the line of interior points [3, E)
    M = sbp_dissipation(n)
    for q in range(6, E - 3):
        want = np.zeros(SZ)
        want[3:E] = M[q - 3] / 64.0
        np.testing.assert_allclose(coeffs(KO, axis, q), want, rtol=1e-15, atol=0)
    bf = LOW[axis] | HIGH[axis]
    for r, smr in enumerate((SMR3, SMR2, SMR1)):
        want = np.zeros(SZ)
        want[3:E] = M[r] / smr
        np.testing.assert_allclose(coeffs(KO, axis, 3 + r, bf), want, rtol=1e-15, atol=0)
        mirror = np.zeros(SZ)
        mirror[3:E] = (M[r] / smr)[::-1]
        np.testing.assert_allclose(coeffs(KO, axis, E - 1 - r, bf), mirror, rtol=1e-15, atol=0)
    # the construction is symmetric and annihilates constants through
    # quadratics on every row, as a dissipation operator must
    assert np.array_equal(M, M.T)
    x = np.arange(n, dtype=float)
    for k in range(3):
        np.testing.assert_allclose(M @ x ** k, 0.0, atol=1e-9)
    # a flag on one face only: the other side keeps interior rows
    np.testing.assert_allclose(coeffs(KO, axis, E - 1, LOW[axis])[E - 4:E + 3],
                               np.array([-1, 6, -15, 20, -15, 6, -1]) / -64.0, rtol=1e-15, atol=0)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_sixth_order_first_rows(axis):
    """deriv64_x/y/z (QuadGrav/src/derivs.cpp:2141-2420): 6th-order centred
    interior, 4th-order one-sided rows at the three points next to a
    flagged face; deriv64adv is the same operator whatever the shift."""
    for q in range(6, E - 3):
        check_row(coeffs(FIRST6, axis, q), q, [q - 3, q - 2, q - 1, q + 1, q + 2, q + 3], 6, 1)
    bf = LOW[axis] | HIGH[axis]
    check_row(coeffs(FIRST6, axis, 3, bf), 3, [3, 4, 5, 6, 7], 4, 1)
    check_row(coeffs(FIRST6, axis, 4, bf), 4, [3, 4, 5, 6, 7], 4, 1)
    check_row(coeffs(FIRST6, axis, 5, bf), 5, [3, 4, 6, 7], 4, 1)
    check_row(coeffs(FIRST6, axis, E - 3, bf), E - 3, [E - 5, E - 4, E - 2, E - 1], 4, 1)
    check_row(coeffs(FIRST6, axis, E - 2, bf), E - 2, list(range(E - 5, E)), 4, 1)
    check_row(coeffs(FIRST6, axis, E - 1, bf), E - 1, list(range(E - 5, E)), 4, 1)
    check_row(coeffs(FIRST6, axis, 3), 3, [0, 1, 2, 4, 5, 6], 6, 1)
    for q in (3, 5, 7, E - 1):
        for b in (0.7, -0.7, 0.0):
            np.testing.assert_array_equal(coeffs(ADV6, axis, q, bf, b), coeffs(FIRST6, axis, q, bf))


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_sixth_order_second_rows(axis):
    """deriv64_xx/yy/zz (QuadGrav/src/derivs.cpp:2569-2910)."""
    for q in range(6, E - 3):
        check_row(coeffs(SECOND6, axis, q), q, list(range(q - 3, q + 4)), 7, 2)
    bf = LOW[axis] | HIGH[axis]
    check_row(coeffs(SECOND6, axis, 3, bf), 3, [3, 4, 5, 6, 7], 4, 2)
    check_row(coeffs(SECOND6, axis, 4, bf), 4, [3, 4, 5, 6, 7], 4, 2)
    check_row(coeffs(SECOND6, axis, 5, bf), 5, [3, 4, 5, 6, 7], 5, 2)
    check_row(coeffs(SECOND6, axis, E - 3, bf), E - 3, list(range(E - 5, E)), 5, 2)
    check_row(coeffs(SECOND6, axis, E - 2, bf), E - 2, list(range(E - 5, E)), 4, 2)
    check_row(coeffs(SECOND6, axis, E - 1, bf), E - 1, list(range(E - 5, E)), 4, 2)


@pytest.mark.parametrize("kind,axis,ranges", [
    (FIRST6, 0, [(3, SZ - 3), (0, SZ), (0, SZ)]),       # QuadGrav/src/derivs.cpp:2151-2158
    (FIRST6, 1, [(3, SZ - 3), (3, SZ - 3), (0, SZ)]),   # :2293-2300
    (FIRST6, 2, [(3, SZ - 3)] * 3),
    (SECOND6, 1, [(3, SZ - 3)] * 3),
    (FIRST, 0, [(3, SZ - 3), (1, SZ - 1), (1, SZ - 1)]),   # derivs.cpp:19-27
    (FIRST, 1, [(3, SZ - 3), (3, SZ - 3), (1, SZ - 1)]),   # derivs.cpp:96-104
    (FIRST, 2, [(3, SZ - 3)] * 3),
    (SECOND, 0, [(3, SZ - 3)] * 3),
    (ADV, 1, [(3, SZ - 3)] * 3),
    (KO, 2, [(3, SZ - 3)] * 3),
])
def test_written_ranges(kind, axis, ranges):
    u = np.random.default_rng(0).standard_normal(SZ ** 3)
    D = po.deriv(kind, axis, u, 0.1, (SZ, SZ, SZ), 0, np.ones(SZ ** 3), fill=np.nan).reshape(SZ, SZ, SZ)
    want = np.zeros((SZ, SZ, SZ), bool)
    (i0, i1), (j0, j1), (k0, k1) = ranges
    want[k0:k1, j0:j1, i0:i1] = True
    assert np.array_equal(np.isfinite(D), want)


def test_spacing_scaling():
    """Operators scale as h^-order: first derivative of x^1 is 1 for any h."""
    for h in (1.0 / 64, 0.37):
        x = (np.arange(SZ) * h)[None, None, :] * np.ones((SZ, SZ, 1))
        D = po.deriv(FIRST, 0, x.ravel(), h, (SZ, SZ, SZ)).reshape(SZ, SZ, SZ)
        np.testing.assert_allclose(D[3:-3, 3:-3, 3:-3], 1.0, rtol=1e-12)
        D2 = po.deriv(SECOND, 0, (x ** 2).ravel(), h, (SZ, SZ, SZ)).reshape(SZ, SZ, SZ)
        np.testing.assert_allclose(D2[3:-3, 3:-3, 3:-3], 2.0, rtol=1e-9)
