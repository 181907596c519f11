"""Full-size workloads on the GPU, through the C ABI: the 4096-block uniform
grid with the schedule the bench uses, and the RK4 graph on mixed block
sizes.  Checked against the oracle on sampled blocks (the whole grid is too
large for the oracle in a test) and by properties that hold at any size."""
import numpy as np
import pytest
import torch

import pyoracle as po
from pydgr import RK4, Plan, data, layout, make_params
from pydgr.native import Schedule

from helpers import TOL, rel_err_per_var

pytestmark = pytest.mark.gpu

NB = 16  # 16^3 = 4096 blocks of 13^3
# the schedule the bench's autotuner picks for this grid (Morton order cut
# into one run per XCD, one workgroup per tile) and the library default
SCHEDULES = {"sfc_grid0": {"tile_max": [7, 7, 7, 7], "order": 3, "grid_mult": 0},
             "default": {"tile_max": [7, 7, 7, 7], "order": 0, "grid_mult": 1}}
# variable permutation of the x <-> y reflection: vectors swap components 0
# and 1, symmetric tensors (00,01,02,11,12,22) swap 00<->11 and 02<->12
XY_PERM = [0, 1, 2, 4, 3, 5, 7, 6, 8, 10, 9, 11, 15, 13, 16, 12, 14, 17, 21, 19, 22, 18, 20, 23]


def sample(nb):
    """Blocks at corners, edges, faces and inside of an nb^3 block grid."""
    c = sorted({0, 1, nb // 2 - 1, nb // 2, nb - 2, nb - 1})
    return [i + nb * (j + nb * k) for k in c for j in c for i in c]


def compact(U, blocks, ids):
    """The variables of blocks `ids` packed back to back (new offsets)."""
    sub, cols, off = [], [], 0
    for b in ids:
        B = dict(blocks[b])
        n = B["sz"][0] * B["sz"][1] * B["sz"][2]
        cols.append(U[:, B["offset"]:B["offset"] + n])
        B["offset"] = off
        sub.append(B)
        off += n
    return sub, np.ascontiguousarray(np.concatenate(cols, axis=1))


@pytest.mark.parametrize("sched", sorted(SCHEDULES))
def test_uniform_4096_rhs_sampled(cuda, sched):
    blocks, dof = layout.uniform_grid(NB, dx=1.0 / 64, outer_bflags=True)
    prm = dict(ko_sigma=0.4, eta_const=2.0)
    Ud = data.gauge_wave(blocks, dof, amp=0.01, noise=1e-6, seed=0, device=cuda)
    Rd = torch.zeros_like(Ud)
    plan = Plan(blocks, make_params(**prm), schedule=Schedule.from_dict(SCHEDULES[sched]))
    plan.rhs(Rd, Ud)
    torch.cuda.synchronize()
    plan.close()
    ids = sample(NB)
    sub, Us = compact(Ud.cpu().numpy(), blocks, ids)
    _, Rs = compact(Rd.cpu().numpy(), blocks, ids)
    del Ud, Rd
    ref = po.rhs_blocks(Us, sub, po.params(**prm), fill=0.0)
    ii = layout.interior_index(sub)
    rel = rel_err_per_var(Rs, ref, ii)
    print("4096-block grid, schedule %s, %d sampled blocks: max rel per var %.3e" % (sched, len(ids), rel.max()),
          flush=True)
    assert np.all(np.isfinite(Rs[:, ii]))
    assert rel.max() <= TOL, rel


def reflect_xy(U, nb, sz=13):
    """(P U)_v(x, y, z) = U_{perm v}(y, x, z) on the centred uniform grid."""
    V = U.reshape(U.shape[0], nb, nb, nb, sz, sz, sz)  # [var, bz, by, bx, k, j, i]
    return np.ascontiguousarray(V.transpose(0, 1, 3, 2, 4, 6, 5)[XY_PERM].reshape(U.shape))


def test_uniform_4096_rk4_step_reflection_symmetric(cuda):
    """A state symmetric under x <-> y stays symmetric through one RK4 graph
    step on the whole grid (boundary faces included), up to rounding."""
    blocks, dof = layout.uniform_grid(NB, dx=1.0 / 64, outer_bflags=True)
    W = data.flat_bumps(blocks, dof, seed=3, amp=1e-3, width=0.1).numpy()
    S = 0.5 * (W + reflect_xy(W, NB))
    del W
    Ud = torch.from_numpy(S).to(cuda)
    plan = Plan(blocks, make_params(ko_sigma=0.4, eta_const=2.0),
                schedule=Schedule.from_dict(SCHEDULES["sfc_grid0"]))
    rk = RK4(plan, Ud, use_graph=True)
    rk.step(0.25 / 64)
    assert rk.status_ok()
    rk.close()
    plan.close()
    T = Ud.cpu().numpy()
    del Ud
    ii = layout.interior_index(blocks)
    assert np.all(np.isfinite(T[:, ii]))
    change = np.abs(T[:, ii] - S[:, ii]).max()
    asym = np.abs(T[:, ii] - reflect_xy(T, NB)[:, ii]).max()
    print("RK4 step on 4096 blocks: max change %.3e, max asymmetry %.3e" % (change, asym), flush=True)
    assert change > 0
    assert asym <= 1e-9 * change


def test_mixed_sizes_rk4_graph_matches_oracle(cuda):
    """RK4 graph steps on blocks of 13, 19, 25 and 37 points per side (each
    size on its own lattice, so every size has neighbours whose interiors
    the stages gather) against the oracle's RK4 with a geometric refill."""
    dx = 1.0 / 64
    blocks, dof = layout.mixed_lattices(dx=dx)
    U = data.flat_bumps(blocks, dof, seed=7, amp=1e-2, width=0.25).numpy()
    po.ghost_refill_blocks(U, blocks)  # consistent padding to start from
    dt = 0.25 * dx
    ref = U.copy()
    for _ in range(2):
        ref, bad = po.rk4_step_blocks(ref, blocks, dt, lambda A: po.ghost_refill_blocks(A, blocks),
                                      po.params(ko_sigma=0.4))
        assert bad == 0
    Ud = torch.from_numpy(U.copy()).to(cuda)
    plan = Plan(blocks, make_params(ko_sigma=0.4))
    rk = RK4(plan, Ud, use_graph=True)
    for _ in range(2):
        rk.step(dt)
    assert rk.status_ok()
    rk.close()
    got = Ud.cpu().numpy()
    ii = layout.interior_index(blocks)
    change = np.abs(ref[:, ii] - U[:, ii]).max(1)
    err = np.abs(got[:, ii] - ref[:, ii]).max(1)
    rel = err / np.where(change > 0, change, 1.0)
    print("mixed sizes %s: max error relative to the two-step change %.3e"
          % (sorted({b["sz"][0] for b in blocks}), rel.max()), flush=True)
    assert rel.max() <= 1e-11, rel
    # ghost layers agree too (the state's own refill at the end of the run)
    assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref - U).max()
