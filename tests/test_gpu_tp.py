"""TwoPunctures initial data through the C ABI (dgr_tp_*): the spectral
solve (rocSOLVER LU on the GPU) against the numpy oracle of the same
collocation problem, the grid evaluation kernel against the oracle's
per-point formulas, the Brill-Lindquist limit, the bare-mass iteration, and
the Hamiltonian constraint of the evaluated grid data (this library's
constraint monitor) converging at the stencils' fourth order."""
import os
import sys

import numpy as np
import pytest
import torch

from pydgr import Plan, TwoPunctures, layout, make_params, tp_params
from pydgr.native import VAR_NAMES

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import tp_oracle as tpo  # noqa: E402

pytestmark = pytest.mark.gpu

BASE = dict(par_b=1.5, par_m_plus=0.5, par_m_minus=0.4, par_P_plus=(0.0, 0.12, 0.03),
            par_P_minus=(0.0, -0.12, 0.01), par_S_plus=(0.02, -0.01, 0.07), par_S_minus=(0.0, 0.05, -0.03),
            give_bare_mass=1, Newton_tol=1e-12, Newton_maxit=20)


def as_dict(p):
    d = {f: getattr(p, f) for f, _ in p._fields_}
    for k, v in d.items():
        if hasattr(v, "__len__"):
            d[k] = tuple(v)
    return d


def grid_points(blocks, dof):
    x = np.zeros((3, dof))
    for b in blocks:
        s0, s1, s2 = b["sz"]
        h = [(b["pmax"][a] - b["pmin"][a]) / (b["sz"][a] - 1) for a in range(3)]
        k, j, i = np.meshgrid(np.arange(s2), np.arange(s1), np.arange(s0), indexing="ij")
        o = b["offset"] + np.arange(s0 * s1 * s2)
        x[0, o] = b["pmin"][0] + i.ravel() * h[0]
        x[1, o] = b["pmin"][1] + j.ravel() * h[1]
        x[2, o] = b["pmin"][2] + k.ravel() * h[2]
    return x


def evaluate(cuda, tp, blocks, dof):
    plan = Plan(blocks, make_params())
    U = torch.full((24, dof), np.nan, dtype=torch.float64, device=cuda)
    tp.evaluate(plan, U)
    torch.cuda.synchronize()
    return plan, U


def test_collocation_solution_matches_oracle(cuda):
    p = tp_params(npoints_A=12, npoints_B=12, npoints_phi=6, **BASE)
    tp = TwoPunctures(p)
    info = tp.info()
    print("newton steps", info["newton_steps"], "last update", info["last_update"], flush=True)
    assert info["last_update"] < 1e-12
    o = tpo.TPOracle(as_dict(p))
    o.newton(1e-13, 30)
    V = tp.collocation()
    err = np.abs(V - o.V).max() / np.abs(o.V).max()
    print("collocation max rel diff vs oracle", err, flush=True)
    assert err < 1e-10
    # spectral interpolation on the host against the oracle's
    rng = np.random.default_rng(3)
    pts = rng.uniform(-5, 5, size=(40, 3))
    np.testing.assert_allclose(tp.u_at(pts), o.u(*pts.T), rtol=1e-9, atol=1e-12 * np.abs(o.V).max())
    tp.close()


def test_grid_evaluation_matches_oracle_formulas(cuda):
    p = tp_params(npoints_A=14, npoints_B=12, npoints_phi=8, center_offset=(0.1, -0.05, 0.02),
                  initial_lapse="averaged", **BASE)
    tp = TwoPunctures(p)
    blocks, dof = layout.uniform_grid(3, dx=0.25)
    plan, U = evaluate(cuda, tp, blocks, dof)
    got = U.cpu().numpy()
    assert np.isfinite(got).all()
    X = grid_points(blocks, dof)
    c = np.asarray(p.center_offset)
    u = tp.u_at(X.T)
    info = tp.info()
    ref = tpo.bssn_point(as_dict(p), info["mp"], info["mm"], X[0] - c[0], X[1] - c[1], X[2] - c[2], u)
    for k in ["alpha", "chi", "K"] + ["gt%d" % q for q in range(6)] + ["At%d" % q for q in range(6)]:
        v = VAR_NAMES.index(k)
        # K is the trace of a traceless tensor: rounding noise on the scale of At
        scale = max(np.abs(ref["At%d" % q]).max() for q in range(6)) if k == "K" else np.abs(ref[k]).max()
        if scale == 0.0:  # off-diagonal gt: exactly zero on both sides
            assert np.all(got[v] == 0.0), k
            continue
        err = np.abs(got[v] - ref[k]).max() / scale
        assert err < 1e-11, (k, err)
    for v in (3, 4, 5, 6, 7, 8, 9, 10, 11):  # Gt, beta, B
        assert np.all(got[v] == 0.0)
    tp.close()
    plan.close()


def test_brill_lindquist_limit(cuda):
    kw = dict(BASE, par_P_plus=(0, 0, 0), par_P_minus=(0, 0, 0), par_S_plus=(0, 0, 0), par_S_minus=(0, 0, 0))
    tp = TwoPunctures(tp_params(npoints_A=10, npoints_B=10, npoints_phi=4, **kw))
    assert np.abs(tp.collocation()).max() == 0.0
    info = tp.info()
    assert info["E"] == info["mp"] + info["mm"]
    blocks, dof = layout.uniform_grid(2, dx=0.29)  # no grid point on a puncture
    plan, U = evaluate(cuda, tp, blocks, dof)
    got = U.cpu().numpy()
    X = grid_points(blocks, dof)
    rp = np.sqrt((X[0] - 1.5) ** 2 + X[1] ** 2 + X[2] ** 2)
    rm = np.sqrt((X[0] + 1.5) ** 2 + X[1] ** 2 + X[2] ** 2)
    psi = 1 + 0.25 / rp + 0.2 / rm
    np.testing.assert_allclose(got[1], psi ** -4, rtol=1e-13)
    np.testing.assert_allclose(got[0], psi ** -2, rtol=1e-13)
    assert np.all(got[18:24] == 0.0) and np.all(got[2] == 0.0)
    tp.close()
    plan.close()


def test_bare_masses_reach_the_target_adm_masses(cuda):
    kw = dict(BASE, give_bare_mass=0, target_M_plus=0.55, target_M_minus=0.45, adm_tol=1e-10)
    tp = TwoPunctures(tp_params(npoints_A=12, npoints_B=12, npoints_phi=6, **kw))
    info = tp.info()
    print(info, flush=True)
    assert abs(info["mp_adm"] - 0.55) <= 1e-10 and abs(info["mm_adm"] - 0.45) <= 1e-10
    assert info["mp"] < 0.55 and info["mm"] < 0.45  # the binding and the momenta take mass
    b = 1.5
    Jt = (np.cross([b, 0, 0], BASE["par_P_plus"]) + np.cross([-b, 0, 0], BASE["par_P_minus"])
          + np.add(BASE["par_S_plus"], BASE["par_S_minus"]))
    np.testing.assert_allclose(info["J"], Jt, atol=1e-15)
    tp.close()


def test_hamiltonian_constraint_converges_on_the_grid(cuda):
    tp = TwoPunctures(tp_params(npoints_A=24, npoints_B=24, npoints_phi=8, **BASE))
    norms = []
    for nb in (8, 16):
        blocks, dof = layout.uniform_grid(nb, dx=8.0 / (6 * nb))
        plan, U = evaluate(cuda, tp, blocks, dof)
        con = torch.zeros((6, dof), dtype=torch.float64, device=cuda)
        plan.physcon(con, U)
        torch.cuda.synchronize()
        X = grid_points(blocks, dof)
        far = ((np.sqrt((X[0] - 1.5) ** 2 + X[1] ** 2 + X[2] ** 2) > 1.0) &
               (np.sqrt((X[0] + 1.5) ** 2 + X[1] ** 2 + X[2] ** 2) > 1.0) & (np.abs(X).max(axis=0) < 3.0))
        inner = np.zeros(dof, bool)
        inner[layout.interior_index(blocks)] = True
        ham = con[0].cpu().numpy()[far & inner]
        norms.append(np.sqrt(np.mean(ham ** 2)))
        plan.close()
    print("Hamiltonian constraint rms, h and h/2:", norms, "ratio", norms[0] / norms[1], flush=True)
    assert norms[0] / norms[1] > 8.0  # fourth-order stencils: 16
    tp.close()


def test_reference_binary_parameters_converge_at_default_resolution(cuda):
    # BSSN_GR/pars/bbh.par.json at the reference's 30 x 30 x 16 points, its
    # default Newton_tol (1e-10) and Newton_maxit (5): without the row
    # scaling of the Newton equations this case diverged
    E = []
    for nphi in (8, 16):
        tp = TwoPunctures(tp_params(par_b=4.0, par_m_plus=0.5, par_m_minus=0.5, give_bare_mass=1,
                                    par_P_plus=(0.0, 0.114, 0.0), par_P_minus=(0.0, -0.114, 0.0),
                                    npoints_A=30, npoints_B=30, npoints_phi=nphi))
        info = tp.info()
        print(nphi, info, flush=True)
        assert info["last_update"] < 1e-10 and info["newton_steps"] <= 5
        E.append(info["E"])
        tp.close()
    assert 1.0 < E[0] < 1.05 and abs(E[0] - E[1]) < 1e-10


def test_host_point_evaluation_matches_the_grid_fill(cuda):
    # dgr_tp_point (the reference's per-point call, on the host) against the
    # device fill at sample padded points: host libm vs device math
    p = tp_params(npoints_A=12, npoints_B=12, npoints_phi=6, center_offset=(0.05, 0.0, -0.03), **BASE)
    tp = TwoPunctures(p)
    blocks, dof = layout.uniform_grid(2, dx=0.29)
    plan, U = evaluate(cuda, tp, blocks, dof)
    got = U.cpu().numpy()
    X = grid_points(blocks, dof)
    rng = np.random.default_rng(5)
    for o in rng.choice(dof, size=40, replace=False):
        host = tp.point(*X[:, o])
        scale = np.abs(got[:, o]).max()
        assert np.abs(host - got[:, o]).max() <= 1e-13 * scale, o
    tp.close()
    plan.close()
