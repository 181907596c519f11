"""The TwoPunctures oracle (oracle/tp_oracle.py) pinned by what the
mathematics fixes: the ADM momenta of the Bowen-York curvature
(TPEquations.cpp:81-145) are the P and S it is built from, A_ij is
traceless and divergence-free away from the punctures, and the solved u
satisfies the Hamiltonian constraint at off-grid points to an accuracy
that improves spectrally with the resolution."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import tp_oracle as tpo  # noqa: E402


def params(**kw):
    p = dict(par_b=1.5, par_m_plus=0.5, par_m_minus=0.4, par_P_plus=(0.0, 0.12, 0.03),
             par_P_minus=(0.0, -0.12, 0.01), par_S_plus=(0.02, -0.01, 0.07), par_S_minus=(0.0, 0.05, -0.03),
             npoints_A=10, npoints_B=10, npoints_phi=6, TP_epsilon=1e-6, TP_Tiny=1e-15, TP_Extend_Radius=0.0,
             initial_lapse=2, initial_lapse_psi_exponent=-2.0)
    p.update(kw)
    return p


def test_bowen_york_adm_momenta_are_the_inputs():
    p = params()
    P, J = tpo.adm_momenta(p)
    b = p["par_b"]
    Pt = np.add(p["par_P_plus"], p["par_P_minus"])
    Jt = (np.cross([b, 0, 0], p["par_P_plus"]) + np.cross([-b, 0, 0], p["par_P_minus"])
          + np.add(p["par_S_plus"], p["par_S_minus"]))
    np.testing.assert_allclose(P, Pt, atol=1e-10)
    np.testing.assert_allclose(J, Jt, atol=1e-10)
    # a flipped spin sign would give -S: the spin term's convention is fixed
    q = params(par_P_plus=(0, 0, 0), par_P_minus=(0, 0, 0), par_S_minus=(0, 0, 0))
    _, J1 = tpo.adm_momenta(q)
    np.testing.assert_allclose(J1, q["par_S_plus"], atol=1e-10)


def test_bowen_york_traceless_and_divergence_free():
    p = params()
    rng = np.random.default_rng(1)
    pts = rng.uniform(-3, 3, size=(30, 3))
    b = p["par_b"]
    pts = pts[(np.linalg.norm(pts - [b, 0, 0], axis=1) > 0.5) & (np.linalg.norm(pts + [b, 0, 0], axis=1) > 0.5)]
    A = tpo.by_aij(p, pts[:, 0], pts[:, 1], pts[:, 2])
    scale = np.abs(A).max(axis=(0, 1))
    assert np.all(np.abs(A[0, 0] + A[1, 1] + A[2, 2]) <= 1e-14 * scale)
    np.testing.assert_allclose(A, A.transpose(1, 0, 2), atol=0)
    h = 1e-3
    div = np.zeros((3, len(pts)))
    for j in range(3):
        e = np.zeros(3)
        e[j] = h
        f = [tpo.by_aij(p, *(pts + s * e).T) for s in (-2, -1, 1, 2)]
        div += (f[0][:, j] - 8 * f[1][:, j] + 8 * f[2][:, j] - f[3][:, j]) / (12 * h)
    assert np.abs(div).max() <= 1e-8 * scale.max()


def ham_residual(o, pts, h=1e-2):
    """|Lap u + A^2/8 psi^-7| / |Lap u| at points (4th-order differences)."""
    out = []
    b = o.p["par_b"]
    for x, y, z in pts:
        u0 = o.u(x, y, z)[0]
        lap = 0.0
        for d in range(3):
            e = np.zeros(3)
            e[d] = h
            us = [o.u(*(np.array([x, y, z]) + s * e))[0] for s in (-2, -1, 1, 2)]
            lap += (-us[0] + 16 * us[1] - 30 * u0 + 16 * us[2] - us[3]) / (12 * h * h)
        psi = (1 + 0.5 * o.mp / np.sqrt((x - b) ** 2 + y * y + z * z) + 0.5 * o.mm / np.sqrt((x + b) ** 2 + y * y + z * z)
               + u0)
        AA = (tpo.by_aij(o.p, x, y, z) ** 2).sum()
        out.append(abs(lap + 0.125 * AA / psi ** 7) / abs(lap))
    return np.array(out)


def test_oracle_solution_satisfies_the_hamiltonian_constraint():
    rng = np.random.default_rng(2)
    pts = rng.uniform(-4, 4, size=(8, 3))
    b = 1.5
    pts = pts[(np.linalg.norm(pts - [b, 0, 0], axis=1) > 1.0) & (np.linalg.norm(pts + [b, 0, 0], axis=1) > 1.0)]
    res = []
    for n in (8, 16):
        o = tpo.TPOracle(params(npoints_A=n, npoints_B=n, npoints_phi=6))
        o.newton(1e-12, 20)
        res.append(np.median(ham_residual(o, pts)))
    print("median relative Hamiltonian residual:", res)
    # spectral convergence: 0.36 at n = 8, 1.6e-3 at n = 16 (2e-4 at n = 20)
    assert res[1] < 3e-3 and res[1] < res[0] / 100


def test_brill_lindquist_limit():
    o = tpo.TPOracle(params(par_P_plus=(0, 0, 0), par_P_minus=(0, 0, 0), par_S_plus=(0, 0, 0),
                            par_S_minus=(0, 0, 0)))
    o.newton(1e-12, 5)
    assert np.abs(o.V).max() == 0.0
    v = tpo.bssn_point(o.p, o.mp, o.mm, np.array([0.3]), np.array([0.7]), np.array([-0.2]), np.zeros(1))
    rp, rm = np.sqrt((0.3 - 1.5) ** 2 + 0.49 + 0.04), np.sqrt((0.3 + 1.5) ** 2 + 0.49 + 0.04)
    psi = 1 + 0.25 / rp + 0.2 / rm
    np.testing.assert_allclose(v["chi"], psi ** -4, rtol=1e-14)
    np.testing.assert_allclose(v["alpha"], psi ** -2, rtol=1e-14)
    assert v["K"][0] == 0.0 and all(v["At%d" % q][0] == 0.0 for q in range(6))
