"""Time-stepper context entry points through the C ABI: block-wise and
single-block RHS (BSSNCtx::rhs_blkwise / rhs_blk) against the batched RHS,
and the Butcher-tableau stepper (ETS::evolve driving BSSNCtx::rhs) against
an oracle restatement of the same loop and against the RK4 path."""
import numpy as np
import pytest
import torch

import pyoracle as po
from pydgr import ETS, ETS_RK3, ETS_RK4, RK4, Plan, data, layout, make_params

pytestmark = pytest.mark.gpu

SENTINEL = -4321.0


def test_rhs_blkwise_equals_batched_on_the_subset(cuda):
    blocks, dof = layout.uniform_grid(3, dx=1.0 / 32, outer_bflags=True)
    U = data.flat_bumps(blocks, dof, seed=11, amp=1e-2).to(cuda)
    plan = Plan(blocks, make_params())
    full = torch.zeros_like(U)
    plan.rhs(full, U)
    ids = [0, 4, 13, 26, 7]  # corner, edge, interior and boundary blocks
    sub = torch.full_like(U, SENTINEL)
    plan.rhs_blkwise(sub, U, ids)
    torch.cuda.synchronize()
    ii_all = layout.interior_index(blocks)
    S, F = sub.cpu().numpy(), full.cpu().numpy()
    picked = np.zeros(dof, bool)
    for b in ids:
        ii = layout.interior_index([blocks[b]])
        picked[ii] = True
        assert np.array_equal(S[:, ii], F[:, ii]), b
    assert np.all(S[:, ~picked] == SENTINEL)
    assert picked.sum() < ii_all.size


def test_rhs_blk_block_local_arrays(cuda):
    blocks, dof = layout.mixed_blocks(5, seed=2, dist=((13, 0.5), (19, 0.5)))
    for b in blocks:
        b["bflag"] = 0b000101
    U = data.flat_bumps(blocks, dof, seed=12, amp=1e-2).to(cuda)
    plan = Plan(blocks, make_params())
    full = torch.zeros_like(U)
    plan.rhs(full, U)
    for n, b in enumerate(blocks):
        nn = b["sz"][0] * b["sz"][1] * b["sz"][2]
        loc = U[:, b["offset"]:b["offset"] + nn].contiguous()
        out = torch.full_like(loc, SENTINEL)
        plan.rhs_blk(out, loc, n)
        torch.cuda.synchronize()
        ii = layout.interior_index([dict(b, offset=0)])
        ref = full[:, b["offset"]:b["offset"] + nn].cpu().numpy()
        got = out.cpu().numpy()
        assert np.array_equal(got[:, ii], ref[:, ii]), n
        mask = np.ones(nn, bool)
        mask[ii] = False
        assert np.all(got[:, mask] == SENTINEL)


RK3 = ([0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.25, 0.25, 0.0], [1 / 6, 1 / 6, 2 / 3], [0.0, 1.0, 0.5])


def ets_step_oracle(U, blocks, nb, dt, tableau, chi_floor=0.1):
    """ETS::evolve with BSSNCtx::rhs on unzipped uniform blocks, restated
    with the oracle's pieces."""
    a, b, _ = tableau
    ns = len(b)
    ii = layout.interior_index(blocks)
    u = U.copy()
    ks = []
    for s in range(ns):
        x = u[:, ii].copy()
        for p in range(s):
            if a[s * ns + p] != 0.0:
                x = x + (a[s * ns + p] * dt) * ks[p][:, ii]
        x, bad = po.enforce_constraints(np.ascontiguousarray(x), chi_floor)
        assert bad == 0
        tmp = u.copy()
        tmp[:, ii] = x
        po.ghost_refill_uniform(tmp, nb)
        ks.append(po.rhs_blocks(tmp, blocks, po.params(), fill=0.0))
    x = u[:, ii]
    for k in range(ns):
        x = x + (b[k] * dt) * ks[k][:, ii]
    u[:, ii] = x
    po.ghost_refill_uniform(u, nb)
    return u


@pytest.mark.parametrize("kind", ["rk3", "rk4"])
def test_ets_matches_oracle(cuda, kind):
    nb = 3
    blocks, dof = layout.uniform_grid(nb, dx=1.0 / 32, outer_bflags=True)
    U0 = data.gauge_wave(blocks, dof, amp=0.05, seed=13).numpy()
    tab = RK3 if kind == "rk3" else ([0, 0, 0, 0, .5, 0, 0, 0, 0, .5, 0, 0, 0, 0, 1, 0],
                                     [1 / 6, 1 / 3, 1 / 3, 1 / 6], [0, .5, .5, 1])
    dt = 0.25 / 32
    ref = U0
    for _ in range(2):
        ref = ets_step_oracle(ref, blocks, nb, dt, tab)
    Ud = torch.from_numpy(U0.copy()).to(cuda)
    plan = Plan(blocks, make_params())
    ets = ETS(plan, Ud, type=ETS_RK3 if kind == "rk3" else ETS_RK4)
    ets.evolve(dt, 2)
    assert ets.status_ok()
    assert abs(ets.time - 2 * dt) < 1e-15
    got = Ud.cpu().numpy()
    ets.close()
    ii = layout.interior_index(blocks)
    change = np.abs(ref[:, ii] - U0[:, ii]).max(1)
    err = np.abs(got[:, ii] - ref[:, ii]).max(1) / np.maximum(change, 1e-300)
    print("%s max rel err of the change %.2e" % (kind, err.max()), flush=True)
    assert err.max() <= 1e-11, err
    # ghosts refilled like the oracle's
    assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max()


def test_ets_tableau_rk4_agrees_with_rk4_path(cuda):
    """An explicit tableau equal to RK4 against dgr_rk4 (which also projects
    the new state): equal up to the final projection and summation order."""
    blocks, dof = layout.uniform_grid(3, dx=1.0 / 32, outer_bflags=True)
    plan = Plan(blocks, make_params())
    # a projected start: ETS projects its stage-0 input, RK4 uses the state
    Ud = data.gauge_wave(blocks, dof, amp=0.05, seed=14).to(cuda)
    assert plan.enforce_constraints(Ud)
    U0 = Ud.cpu()
    dt = 0.25 / 32
    A = U0.clone().to(cuda)
    ets = ETS(plan, A, tableau=([0, 0, 0, 0, .5, 0, 0, 0, 0, .5, 0, 0, 0, 0, 1, 0],
                                [1 / 6, 1 / 3, 1 / 3, 1 / 6], [0, .5, .5, 1]))
    ets.evolve(dt)
    ets.close()
    assert plan.enforce_constraints(A)
    B = U0.clone().to(cuda)
    rk = RK4(plan, B, use_graph=False)
    rk.step(dt)
    rk.close()
    torch.cuda.synchronize()
    ii = layout.interior_index(blocks)
    a, b, u0 = A.cpu().numpy()[:, ii], B.cpu().numpy()[:, ii], U0.numpy()[:, ii]
    change = np.abs(b - u0).max(1)
    assert (np.abs(a - b).max(1) / np.maximum(change, 1e-300)).max() <= 1e-11


def test_ets_rejects_implicit_tableau(cuda):
    blocks, dof = layout.uniform_grid(1, dx=1.0 / 32)
    U = data.flat_bumps(blocks, dof).to(cuda)
    plan = Plan(blocks, make_params())
    with pytest.raises(Exception):
        ETS(plan, U, tableau=([0.5, 0, 0, 0], [0.5, 0.5], [0.5, 1.0]))
