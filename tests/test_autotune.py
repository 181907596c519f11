"""Schedules (tile extent per size class, tile order, grid) and the
autotuner: host-side logic on the CPU; on the GPU, every schedule gives
bitwise the same RHS, and tune() picks, caches and re-reads a schedule."""
import numpy as np
import pytest
import torch

from pydgr import RK4, Plan, Schedule, data, default_schedule, layout, make_params
from pydgr import autotune
from pydgr.native import ORDER_LARGEST, ORDER_SFC, ORDER_XCD, size_class


def test_schedule_roundtrip_and_default():
    s = default_schedule()
    assert list(s.tile_max) == [7, 7, 7, 7] and s.order == 0 and s.grid_mult == 1
    d = {"tile_max": [5, 6, 7, 3], "order": 2, "grid_mult": 0}
    assert Schedule.from_dict(d).as_dict() == d


def test_size_classes():
    assert [size_class(s) for s in ([13] * 3, [19, 13, 13], [25] * 3, [37] * 3, [14, 13, 13])] == [0, 1, 2, 3, 1]


def test_relocate_and_sample_keep_blocks():
    blocks, dof = layout.mixed_blocks(50, seed=3)
    sub = autotune._sample(blocks, 20, seed=1)
    assert len(sub) == 20 and all(b in blocks for b in sub)
    rel, n = autotune._relocate(sub)
    assert rel[0]["offset"] == 0 and n == layout.padded_points(rel)
    assert all(a["sz"] == b["sz"] and a["pmin"] == b["pmin"] for a, b in zip(rel, sub))


@pytest.mark.gpu
def test_every_schedule_gives_identical_rhs(cuda):
    blocks, dof = layout.mixed_blocks(40, seed=7)
    rng = np.random.default_rng(2)
    for b in blocks:
        b["bflag"] = int(rng.integers(0, 64)) if rng.random() < 0.5 else 0
    U = data.flat_bumps(blocks, dof, seed=4, device=cuda)
    prm = make_params(ko_sigma=0.4)
    ref = torch.zeros_like(U)
    Plan(blocks, prm).rhs(ref, U)
    for d in ({"tile_max": [5, 5, 5, 5], "order": 0, "grid_mult": 1},
              {"tile_max": [7, 6, 5, 4], "order": ORDER_XCD, "grid_mult": 0},
              {"tile_max": [6, 7, 7, 6], "order": ORDER_LARGEST, "grid_mult": 2},
              {"tile_max": [7, 7, 7, 7], "order": ORDER_SFC, "grid_mult": 1}):
        R = torch.zeros_like(U)
        p = Plan(blocks, prm, schedule=Schedule.from_dict(d))
        p.rhs(R, U)
        p.rhs(R, U, per_block=True)  # the block-order table too
        torch.cuda.synchronize()
        assert torch.equal(R, ref), d


@pytest.mark.gpu
def test_tune_picks_and_caches(cuda, tmp_path, monkeypatch):
    monkeypatch.setenv("DGR_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    blocks, _ = layout.mixed_blocks(60, seed=1)
    prm = make_params(ko_sigma=0.4)
    s, rep = autotune.tune(blocks, prm, device=cuda, reps=1, probe_blocks=16, sample_blocks=40)
    assert not rep["cached"] and rep["best_ms"] > 0 and rep["default_ms"] > 0
    assert rep["best_ms"] <= min(t["ms"] for t in rep["trials"] if "order" in t) + 1e-9
    s2, rep2 = autotune.tune(blocks, prm, device=cuda)
    assert rep2["cached"] and s2.as_dict() == s.as_dict()


@pytest.mark.gpu
def test_rk4_identical_under_every_schedule_and_rk4_tuning(cuda, tmp_path, monkeypatch):
    """RK4 steps (whose RHS gathers ghosts from neighbours) give bitwise the
    same state for every tile order; the RK4-mode tuner picks and caches."""
    monkeypatch.setenv("DGR_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    blocks, dof = layout.uniform_grid(3, dx=1.0 / 32)
    U0 = data.flat_bumps(blocks, dof, seed=3, amp=1e-2, width=0.25).to(cuda)
    prm = make_params(ko_sigma=0.4)
    outs = []
    for d in ({"tile_max": [7, 7, 7, 7], "order": 0, "grid_mult": 1},
              {"tile_max": [7, 7, 7, 7], "order": ORDER_SFC, "grid_mult": 0},
              {"tile_max": [5, 7, 7, 7], "order": ORDER_XCD, "grid_mult": 2}):
        U = U0.clone()
        plan = Plan(blocks, prm, schedule=Schedule.from_dict(d))
        rk = RK4(plan, U)
        rk.step(0.25 / 32, 2)
        assert rk.status_ok()
        rk.close()
        outs.append(U)
    assert torch.equal(outs[0], outs[1]) and torch.equal(outs[0], outs[2])
    s, rep = autotune.tune(blocks, prm, device=cuda, reps=1, probe_blocks=8, mode="rk4")
    assert rep["mode"] == "rk4" and not rep["cached"] and rep["sample_blocks"] == len(blocks)
    s2, rep2 = autotune.tune(blocks, prm, device=cuda, mode="rk4")
    assert rep2["cached"] and s2.as_dict() == s.as_dict()
