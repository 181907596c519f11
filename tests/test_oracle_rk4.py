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
