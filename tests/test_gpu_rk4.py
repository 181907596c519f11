"""GPU RK4 path (dgr_rk4_*, dgr_ghost_refill, dgr_enforce_constraints)
through the C ABI against the oracle."""
import numpy as np
import pytest
import torch

import pyoracle as po
from pydgr import RK4, Plan, data, layout, make_params

pytestmark = pytest.mark.gpu


def test_ghost_refill_matches_oracle_bitwise(cuda):
    nb = 3
    blocks, dof = layout.uniform_grid(nb, dx=1.0 / 32)
    U = data.flat_bumps(blocks, dof, seed=1, amp=0.05, width=0.3).numpy()
    ii = layout.interior_index(blocks)
    ghost = np.ones(dof, bool)
    ghost[ii] = False
    W = U.copy()
    W[:, ghost] = -1.0
    ref = po.ghost_refill_uniform(W.copy(), nb)
    Wd = torch.from_numpy(W).to(cuda)
    Plan(blocks, make_params()).ghost_refill(Wd)
    assert np.array_equal(Wd.cpu().numpy(), ref)


def test_enforce_constraints_matches_oracle(cuda):
    rng = np.random.default_rng(3)
    blocks, dof = layout.uniform_grid(2, dx=1.0 / 32)
    U = data.flat_bumps(blocks, dof, seed=2, amp=0.05, width=0.3).numpy()
    U[18:24] += rng.normal(0, 0.05, (6, dof))  # At with a trace
    U[12:18] *= 1.0 + rng.uniform(0, 0.1, (1, dof))  # det(gt) != 1
    U[1] = rng.uniform(0.0, 0.5, dof)  # chi below the floor in places
    ii = layout.interior_index(blocks)
    ref = U.copy()
    fixed, bad = po.enforce_constraints(np.ascontiguousarray(U[:, ii]))
    assert bad == 0
    ref[:, ii] = fixed
    Ud = torch.from_numpy(U).to(cuda)
    assert Plan(blocks, make_params(chi_floor=0.1)).enforce_constraints(Ud)
    G = Ud.cpu().numpy()
    scale = np.abs(ref[:, ii]).max(1)
    assert (np.abs(G[:, ii] - ref[:, ii]).max(1) / scale).max() < 1e-14
    mask = np.ones(dof, bool)
    mask[ii] = False
    assert np.array_equal(G[:, mask], U[:, mask])


def _evolve_gpu(cuda, blocks, U, dt, steps, use_graph, order=42):
    Ud = torch.from_numpy(U.copy()).to(cuda)
    plan = Plan(blocks, make_params(deriv_order=order))
    rk = RK4(plan, Ud, use_graph=use_graph)
    for _ in range(steps):
        rk.step(dt)
    ok = rk.status_ok()
    out = Ud.cpu().numpy()
    rk.close()
    return out, ok


@pytest.mark.parametrize("use_graph,order", [(False, 42), (True, 42), (True, 64)])
def test_rk4_steps_match_oracle(cuda, use_graph, order):
    nb = 3
    dx = 1.0 / 32
    blocks, dof = layout.uniform_grid(nb, dx=dx)
    U = data.flat_bumps(blocks, dof, seed=5, amp=1e-2, width=0.25).numpy()
    dt = 0.25 * dx
    ref = U.copy()
    for _ in range(2):
        ref, bad = po.rk4_step_uniform(ref, blocks, nb, dt, po.params(deriv_order=order))
        assert bad == 0
    got, ok = _evolve_gpu(cuda, blocks, U, dt, 2, use_graph, order)
    assert ok
    ii = layout.interior_index(blocks)
    change = np.abs(ref[:, ii] - U[:, ii]).max(1)
    err = np.abs(got[:, ii] - ref[:, ii]).max(1)
    rel = err / np.where(change > 0, change, 1.0)
    print("max error relative to the two-step change: %.3e" % rel.max(), flush=True)
    assert rel.max() <= 1e-11, rel


def test_graph_replays_match_eager_steps(cuda):
    """Several graph replays with unrelated eager GPU work and host
    synchronisation between them give bitwise the same state as eager
    stepping."""
    nb = 2
    dx = 1.0 / 32
    blocks, dof = layout.uniform_grid(nb, dx=dx)
    U = data.flat_bumps(blocks, dof, seed=6, amp=1e-2, width=0.25).numpy()
    plan = Plan(blocks, make_params())
    Ue = torch.from_numpy(U.copy()).to(cuda)
    Ug = torch.from_numpy(U.copy()).to(cuda)
    rke = RK4(plan, Ue, use_graph=False)
    rkg = RK4(plan, Ug, use_graph=True)
    scratch = torch.zeros(1 << 20, dtype=torch.float64, device=cuda)
    for n in range(4):
        rke.step(0.25 * dx)
        rkg.step(0.25 * dx)
        scratch.add_(1.0)
        torch.cuda.synchronize()
        assert torch.equal(Ue, Ug), n
    assert rke.status_ok() and rkg.status_ok()


@pytest.mark.parametrize("kind", ["uniform", "adaptive"])
def test_one_call_equals_split_calls(cuda, kind):
    """Within one call the state lives in the evolver's work memory and is
    written back (interiors, then the ghost refills) when the call ends: n
    steps in one call and the same steps over several calls give bitwise
    the same arrays, padding included; a call of 0 steps changes nothing."""
    if kind == "uniform":
        dx = 1.0 / 32
        blocks, dof = layout.uniform_grid(3, dx=dx)
        U = data.flat_bumps(blocks, dof, seed=7, amp=1e-2, width=0.25).to(cuda)
        prm = make_params()
        dt = 0.25 * dx
    else:
        blocks, dof, _ = layout.adaptive_octree(half_width=4.0, base_level=1, levels=4, centers=((0.7, 0.1, -0.2),),
                                                factor=1.5)
        U = data.flat_bumps(blocks, dof, seed=8, amp=1e-3, width=0.5, box=((-4.0,) * 3, (4.0,) * 3)).to(cuda)
        prm = make_params(ko_sigma=0.4, chi_floor=1e-4)
        dt = 0.25 * min((b["pmax"][0] - b["pmin"][0]) / 12 for b in blocks)
    plan = Plan(blocks, prm)
    plan.ghost_refill(U)  # consistent padding to start from
    U0 = U.clone()
    Ua, Ub = U.clone(), U.clone()
    rka = RK4(plan, Ua, use_graph=True)
    rkb = RK4(plan, Ub, use_graph=True)
    rka.step(dt, 0)
    torch.cuda.synchronize()
    assert torch.equal(Ua, U0)
    rka.step(dt, 5)
    for n in (2, 1, 2):
        rkb.step(dt, n)
    torch.cuda.synchronize()
    assert rka.status_ok() and rkb.status_ok()
    assert torch.isfinite(Ua).all()
    assert torch.equal(Ua, Ub)
    assert not torch.equal(Ua, U0)
    rka.close()
    rkb.close()
