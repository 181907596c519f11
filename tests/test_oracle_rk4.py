"""The oracle's same-level ghost refill and RK4 step (test infrastructure
for the GPU RK4 path).  Ghost refill has no reference counterpart that
builds here (Mesh::unzip needs MPI); it is pinned by exactness: refilling
analytic data after wiping every ghost point restores the analytic values
bit for bit wherever a neighbour exists, and leaves the rest untouched."""
import numpy as np

import pyoracle as po
from pydgr import data, layout


def test_ghost_refill_restores_analytic_ghosts():
    nb = 3
    blocks, dof = layout.uniform_grid(nb, dx=1.0 / 32)
    U = data.flat_bumps(blocks, dof, seed=1, amp=0.05, width=0.3).numpy()
    ii = layout.interior_index(blocks)
    ghost = np.ones(dof, bool)
    ghost[ii] = False
    W = U.copy()
    W[:, ghost] = np.nan
    po.ghost_refill_uniform(W, nb)
    # ghosts with a neighbour are restored exactly; the others stay wiped
    filled = ~np.isnan(W[0])
    assert np.array_equal(W[:, filled], U[:, filled])
    x, y, z, _ = data.coords(blocks)
    x, y, z = x.numpy(), y.numpy(), z.numpy()
    lo, hi = x.min(), x.max()
    # a wiped ghost lies outside the domain interior (beyond the outermost
    # interior nodes) in at least one direction
    L = nb * 6 / 32  # extent of the block interiors
    outside = (np.abs(x) > L / 2 + 1e-12) | (np.abs(y) > L / 2 + 1e-12) | (np.abs(z) > L / 2 + 1e-12)
    assert np.array_equal(~filled, outside & ghost)
    assert lo < -L / 2 and hi > L / 2


def test_rk4_step_flat_space_is_static():
    nb = 2
    blocks, dof = layout.uniform_grid(nb, dx=1.0 / 32)
    U = data.flat_bumps(blocks, dof, amp=0.0).numpy()
    U2, bad = po.rk4_step_uniform(U, blocks, nb, dt=0.25 / 32)
    assert bad == 0
    assert np.array_equal(U2, U)


def test_geometric_refill_equals_uniform_refill():
    """The oracle's geometric same-level refill (any block list) gives the
    uniform-grid refill bitwise, and copies only between equal sizes."""
    from pydgr import data, layout
    blocks, dof = layout.uniform_grid(3, dx=1.0 / 32)
    U = data.flat_bumps(blocks, dof, seed=1, amp=0.05, width=0.3).numpy()
    ii = layout.interior_index(blocks)
    g = np.ones(dof, bool)
    g[ii] = False
    W = U.copy()
    W[:, g] = -1.0
    assert np.array_equal(po.ghost_refill_uniform(W.copy(), 3), po.ghost_refill_blocks(W.copy(), blocks))
    mb, mdof = layout.mixed_lattices(dx=1.0 / 64)
    V = np.random.default_rng(0).standard_normal((2, mdof))
    mi = layout.interior_index(mb)
    V2 = po.ghost_refill_blocks(V.copy(), mb)
    assert np.array_equal(V2[:, mi], V[:, mi])
    assert not np.array_equal(V2, V)
