"""Device initial data (dgr_init_puncture) through the C ABI against the
oracle's restatement of punctureData at every padded point: <= 1e-12 of
each variable's max |value| (libm pow / log / sqrt may differ in the last
bits between host and device), floors and zero variables exact; and the
synthetic perturbed flat space of dgr_init_bumps bitwise equal to the torch
generator it replaces on the device."""
import numpy as np
import pytest
import torch

import pyoracle as po
from pydgr import Plan, data, layout, make_params

pytestmark = pytest.mark.gpu

HOLES = [dict(mass=0.6, pos=(0.31, 0.05, -0.02), p=(0.0, 0.12, 0.01), spin=0.2, spin_theta=0.4, spin_phi=1.2),
         dict(mass=0.4, pos=(-0.29, -0.04, 0.03), p=(0.0, -0.12, 0.0), spin=-0.1, spin_theta=2.0, spin_phi=0.3)]


@pytest.mark.parametrize("kind", ["uniform", "mixed"])
def test_puncture_data_matches_oracle(cuda, kind):
    if kind == "uniform":
        blocks, dof = layout.uniform_grid(4, dx=0.03, outer_bflags=True)
    else:
        blocks, dof = layout.mixed_blocks(8, seed=3, dist=((13, 0.5), (19, 0.3), (25, 0.2)), dx=0.02)
    X, Y, Z, idx = [c.numpy() for c in data.coords(blocks)]
    ref = np.empty((24, dof))
    ref[:, idx] = po.puncture_points(np.stack([X, Y, Z]), HOLES, chi_floor=0.05)
    U = torch.full((24, dof), np.nan, dtype=torch.float64, device=cuda)
    plan = Plan(blocks, make_params(chi_floor=0.05))
    plan.init_puncture(U, HOLES)
    torch.cuda.synchronize()
    got = U.cpu().numpy()
    assert np.all(np.isfinite(got))
    for v in range(24):
        s = np.abs(ref[v]).max()
        if s == 0.0:
            assert np.all(got[v] == 0.0), v
            continue
        err = np.abs(got[v] - ref[v]).max() / s
        assert err <= 1e-12, (v, err)
    assert np.any(got[1] == 0.05) and np.all(got[1] >= 0.05)  # floor reached near the holes


def test_evolution_starts_from_device_data(cuda):
    """The RHS of device-initialised data matches the oracle's RHS of the
    oracle's initial data (the two paths share nothing)."""
    blocks, dof = layout.uniform_grid(2, dx=0.05, outer_bflags=True)
    U = torch.zeros((24, dof), dtype=torch.float64, device=cuda)
    plan = Plan(blocks, make_params(chi_floor=0.05))
    plan.init_puncture(U, HOLES)
    R = torch.zeros_like(U)
    plan.rhs(R, U)
    torch.cuda.synchronize()
    X, Y, Z, idx = [c.numpy() for c in data.coords(blocks)]
    U0 = np.empty((24, dof))
    U0[:, idx] = po.puncture_points(np.stack([X, Y, Z]), HOLES, chi_floor=0.05)
    ref = po.rhs_blocks(U0, blocks, po.params())
    ii = layout.interior_index(blocks)
    got = R.cpu().numpy()
    rel = np.abs(got[:, ii] - ref[:, ii]).max(1) / np.maximum(np.abs(ref[:, ii]).max(1), 1e-300)
    assert rel.max() <= 1e-10, rel


def test_init_bumps_equals_the_torch_generator_bitwise(cuda):
    # dgr_init_bumps (one fused kernel) and pydgr.data.flat_bumps (torch on the
    # same device) perform the same operations in the same order
    for blocks, dof in (layout.block_grid(70), layout.mixed_lattices()):
        plan = Plan(blocks, make_params())
        ref = data.flat_bumps(blocks, dof, seed=4, device=cuda)
        got = data.flat_bumps(blocks, dof, seed=4, device=cuda, out=torch.full((24, dof), float("nan"),
                                                                                dtype=torch.float64, device=cuda),
                              plan=plan)
        torch.cuda.synchronize()
        assert torch.equal(got, ref)
        plan.close()
