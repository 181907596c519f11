"""Ghost fill across level jumps (2:1-balanced layouts of 13^3 blocks).
No reference oracle exists for Mesh::unzip here, so the fill is checked by
what order-6 Lagrange interpolation and node copies must do: a polynomial of
degree <= 6 in each coordinate is reproduced at every ghost point inside the
domain, ghosts outside the domain are untouched, and after RK4 steps the
padding is already consistent (refilling changes nothing)."""
import numpy as np
import pytest
import torch

from pydgr import RK4, Plan, data, layout, make_params

pytestmark = pytest.mark.gpu
SENT = 1.0e30


def _poly(x, y, z):
    return (1.0 + 0.3 * x - 0.2 * y ** 2 + 0.1 * x * y * z + 0.02 * x ** 3 * y ** 2 * z - 0.001 * y ** 6
            + 0.0005 * x ** 6 * z ** 4 - 0.002 * z ** 5)


def _layout():
    return layout.adaptive_octree(half_width=4.0, base_level=1, levels=4, centers=((0.7, 0.1, -0.2),), factor=1.5)


def test_polynomial_reproduced_at_every_ghost(cuda):
    blocks, dof, lev = _layout()
    assert len(set(lev)) >= 3
    x, y, z, idx = data.coords(blocks)
    x, y, z, idx = x.numpy(), y.numpy(), z.numpy(), idx.numpy()
    assert np.array_equal(idx, np.arange(dof))
    P = np.empty(dof)
    P[idx] = _poly(x, y, z)
    ii = layout.interior_index(blocks)
    U = np.full((24, dof), SENT)
    for v in range(24):
        U[v, ii] = (1.0 + 0.01 * v) * P[ii]
    Ud = torch.from_numpy(U).to(cuda)
    plan = Plan(blocks, make_params())
    boxes, copies = plan.level_fills()
    print("level-jump boxes %d, node copies %d" % (boxes, copies), flush=True)
    assert boxes > 0 and copies > 0
    plan.ghost_refill(Ud)
    G = Ud.cpu().numpy()
    ghost = np.ones(dof, bool)
    ghost[ii] = False
    pos = np.stack([x, y, z])  # position k of the flat padded index (data.coords order)
    inside = np.all(np.abs(pos) <= 4.0 + 1e-12, axis=0)
    for v in (0, 7, 23):
        want = (1.0 + 0.01 * v) * P
        g_in = ghost & inside
        err = np.abs(G[v, g_in] - want[g_in]).max() / np.abs(want).max()
        print("var %d: max rel error at %d ghosts inside the domain %.2e" % (v, g_in.sum(), err), flush=True)
        assert err < 1e-12
        assert np.all(G[v, ghost & ~inside] == SENT)


def test_rk4_padding_consistent_after_steps(cuda):
    blocks, dof, _ = _layout()
    U = data.flat_bumps(blocks, dof, seed=2, amp=1e-3, width=0.5, box=((-4.0,) * 3, (4.0,) * 3)).to(cuda)
    plan = Plan(blocks, make_params(ko_sigma=0.4, chi_floor=1e-4))
    plan.ghost_refill(U)  # consistent padding to start from
    rk = RK4(plan, U)
    dx_min = min((b["pmax"][0] - b["pmin"][0]) / 12 for b in blocks)
    rk.step(0.25 * dx_min, 2)
    assert rk.status_ok()
    rk.close()
    torch.cuda.synchronize()
    before = U.clone()
    plan.ghost_refill(U)
    torch.cuda.synchronize()
    assert torch.isfinite(U).all()
    assert torch.equal(U, before)
