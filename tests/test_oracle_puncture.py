"""The oracle's restatement of punctureData (oracle/puncture_oracle.cpp)
against closed forms: Brill-Lindquist values, floors, a trace-free and
divergence-free Bowen-York curvature with the textbook momentum and spin
terms, and the limits of the conformal-factor corrections."""
import numpy as np

import pyoracle as po

BL = [dict(mass=0.5, pos=(4.0, 0.0, 0.0)), dict(mass=0.5, pos=(-4.0, 0.0, 0.0))]


def _pts(rng, n, lo=-6.0, hi=6.0):
    return rng.uniform(lo, hi, (3, n))


def test_brill_lindquist_values():
    rng = np.random.default_rng(1)
    x = _pts(rng, 500)
    out = po.puncture_points(x, BL, chi_floor=1e-4)
    r1 = np.sqrt(((x - np.array([[4.0], [0], [0]])) ** 2).sum(0))
    r2 = np.sqrt(((x - np.array([[-4.0], [0], [0]])) ** 2).sum(0))
    psi = 1 + 0.5 / (2 * r1) + 0.5 / (2 * r2)
    assert np.allclose(out[1], np.maximum(psi ** -4, 1e-4), rtol=1e-14, atol=0)
    assert np.allclose(out[0], np.maximum(psi ** -2, 1e-4), rtol=1e-14, atol=0)
    assert np.all(out[[12, 15, 17]] == 1.0)
    zero = [v for v in range(24) if v not in (0, 1, 12, 15, 17)]
    assert np.all(out[zero] == 0.0)


def test_floors_near_a_puncture():
    x = np.array([[4.001], [0.0], [0.0]])
    out = po.puncture_points(x, BL, chi_floor=0.1)
    assert out[1, 0] == 0.1 and out[0, 0] == 0.1


def _holes_by(rng):
    return [dict(mass=0.6, pos=(1.5, 0.2, -0.1), p=tuple(rng.uniform(-0.2, 0.2, 3)), spin=0.15,
                 spin_theta=0.7, spin_phi=1.9),
            dict(mass=0.4, pos=(-1.5, -0.3, 0.2), p=tuple(rng.uniform(-0.2, 0.2, 3)), spin=-0.1,
                 spin_theta=2.1, spin_phi=-0.4)]


def _Ahat(x, holes):
    """psi^6 At (the flat-space Bowen-York tensor) from the oracle."""
    out = po.puncture_points(x, holes, chi_floor=0.0)
    psi6 = out[1] ** -1.5  # chi = psi^-4 without floor
    A = np.zeros((3, 3) + x.shape[1:])
    pairs = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    for p, (i, j) in enumerate(pairs):
        A[i, j] = A[j, i] = out[18 + p] * psi6
    return A


def test_bowen_york_textbook_terms():
    """psi^6 At = sum over holes of 3/(2 r^2) (P_i n_j + P_j n_i - (delta_ij -
    n_i n_j) P.n) + 3/r^3 ((S x n)_i n_j + (S x n)_j n_i)."""
    rng = np.random.default_rng(2)
    holes = _holes_by(rng)
    x = _pts(rng, 200, -4, 4)
    got = _Ahat(x, holes)
    ref = np.zeros_like(got)
    for h in holes:
        d = x - np.array(h["pos"]).reshape(3, 1)
        r = np.sqrt((d ** 2).sum(0))
        n = d / r
        P = np.array(h["p"]).reshape(3, 1)
        th, ph = h["spin_theta"], h["spin_phi"]
        S = h["spin"] * np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)]).reshape(3, 1)
        sxn = np.cross(S, n, axis=0)
        pn = (P * n).sum(0)
        for i in range(3):
            for j in range(3):
                ref[i, j] += 3 / (2 * r ** 2) * (P[i] * n[j] + P[j] * n[i] - ((i == j) - n[i] * n[j]) * pn)
                ref[i, j] += 3 / r ** 3 * (sxn[i] * n[j] + sxn[j] * n[i])
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-14 * np.abs(ref).max())


def test_bowen_york_trace_and_divergence_free():
    rng = np.random.default_rng(3)
    holes = _holes_by(rng)
    x0 = np.array([[0.3, -0.4, 2.0], [1.1, 1.5, -0.2], [-0.9, 0.8, 0.9]]).T  # away from the holes
    A = _Ahat(x0, holes)
    assert np.abs(A[0, 0] + A[1, 1] + A[2, 2]).max() <= 1e-14 * np.abs(A).max()
    h = 1e-3
    c = np.array([1.0, -8.0, 8.0, -1.0]) / (12 * h)
    div = np.zeros((3, x0.shape[1]))
    scale = 0.0
    for j in range(3):
        e = np.zeros((3, 1))
        e[j] = 1.0
        dA = sum(cc * _Ahat(x0 + s * h * e, holes)[:, j] for cc, s in zip(c, (-2, -1, 1, 2)))
        div += dA
        scale = max(scale, np.abs(dA).max())
    assert np.abs(div).max() <= 1e-8 * scale, (np.abs(div).max(), scale)


def test_corrections_vanish_without_spin_or_momentum_and_decay():
    rng = np.random.default_rng(4)
    x = _pts(rng, 100)
    small = [dict(BL[0], spin=5e-7, p=(3e-7, 0.0, 0.0)), dict(BL[1])]
    assert np.array_equal(po.puncture_points(x, small, 0.0)[:2], po.puncture_points(x, BL, 0.0)[:2])
    spun = [dict(mass=1.0, pos=(0.0, 0.0, 0.0), spin=0.5, spin_theta=0.3, spin_phi=0.2, p=(0.2, 0.1, 0.0)),
            dict(mass=1e-12, pos=(1e6, 0.0, 0.0))]
    plain = [dict(mass=1.0, pos=(0.0, 0.0, 0.0)), dict(mass=1e-12, pos=(1e6, 0.0, 0.0))]
    d = []
    for r in (100.0, 200.0, 400.0):
        p = np.array([[r], [0.0], [0.0]])
        psi_s = po.puncture_points(p, spun, 0.0)[1, 0] ** -0.25
        psi_p = po.puncture_points(p, plain, 0.0)[1, 0] ** -0.25
        d.append(psi_s - psi_p)
    assert d[0] > d[1] > d[2] > 0.0
    assert 0.3 < d[1] / d[0] < 0.7 and 0.3 < d[2] / d[1] < 0.7  # ~1/r
