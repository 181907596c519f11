"""Autotuner for the RHS kernel's tile shapes and schedule.

What is tuned (dgr_schedule, include/dgr_bssn.h):
  tile_max[c]  largest tile extent per axis for blocks of size class c
               (padded extent <= 13, <= 19, <= 25, larger); a block's
               interior is cut into the fewest near-equal runs of at most
               tile_max points per axis;
  order        tile order of the persistent launch: block order, XCD-aware
               (a block's tiles on one XCD so its halo stays in one L2),
               largest blocks first, or Morton order cut into one run per
               XCD (neighbouring blocks share an L2, which the RK4 stages'
               ghost gathers read);
  grid_mult    persistent grid of grid_mult x CUs, or one workgroup per tile.

Every schedule gives bitwise the same result; only the speed differs.  The
search is measured, not modelled: per size class present, the candidate
tile extents are timed on (up to `probe_blocks`) blocks of that class alone;
then, with the winning extents, every (order, grid) pair is timed on the
whole block list (or a class-stratified sample of it).  Times are HIP-event
means of `reps` launches after one warm-up on seeded flat-space-plus-bumps
data (the kernel's time does not depend on the values beyond the sign of
the shift).  Results are cached in a JSON file keyed by device name and
the workload's size-class histogram, so a workload shape is tuned once.
"""
import json
import os
import time

import numpy as np
import torch

from . import data, layout
from .native import (ORDER_BLOCK, ORDER_LARGEST, ORDER_SFC, ORDER_XCD, RK4, SIZE_CLASSES, Plan, Schedule,
                     default_schedule, size_class)

TILE_CANDIDATES = (7, 6, 5)
ORDER_CANDIDATES = (ORDER_BLOCK, ORDER_XCD, ORDER_LARGEST, ORDER_SFC)
GRID_CANDIDATES = (1, 2, 0)


def cache_path():
    p = os.environ.get("DGR_AUTOTUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "dgr", "autotune.json")


def _load_cache(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save_cache(path, cache):
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp%d" % os.getpid()
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # an unwritable cache only costs a re-tune


def workload_key(blocks, device):
    """Device name + per-class block counts (rounded to powers of two) +
    whether any block has boundary faces."""
    counts = [0] * SIZE_CLASSES
    for b in blocks:
        counts[size_class(b["sz"])] += 1
    hist = ",".join("%d:%d" % (c, 1 << int(np.log2(n))) for c, n in enumerate(counts) if n)
    faces = int(any(b["bflag"] for b in blocks))
    return "%s|%s|faces=%d" % (torch.cuda.get_device_name(device), hist, faces)


def _relocate(blocks):
    """Copy of a block list with contiguous offsets from 0."""
    out, off = [], 0
    for b in blocks:
        c = dict(b)
        c["offset"] = off
        off += b["sz"][0] * b["sz"][1] * b["sz"][2]
        out.append(c)
    return out, off


def _time_rhs(blocks, dof, params, sched, device, reps, mode="rhs"):
    """Mean ms of one RHS launch, or (mode "rk4") of one hipGraph RK4 step."""
    U = data.flat_bumps(blocks, dof, seed=1, amp=1e-3, width=0.1, device=device)
    R = torch.zeros_like(U) if mode == "rhs" else None
    plan = Plan(blocks, params, schedule=sched)
    rk = None
    try:
        stream = torch.cuda.current_stream(device)
        if mode == "rk4":
            dx = min((b["pmax"][0] - b["pmin"][0]) / (b["sz"][0] - 1) for b in blocks)
            rk = RK4(plan, U, use_graph=True)
            rk.step(0.25 * dx, 1, stream)
            run = lambda: rk.run(1, stream)  # noqa: E731
        else:
            plan.rhs(R, U, stream)
            run = lambda: plan.rhs(R, U, stream)  # noqa: E731
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            run()
        e1.record(stream)
        torch.cuda.synchronize(device)
        return e0.elapsed_time(e1) / reps
    finally:
        if rk is not None:
            rk.close()
        plan.close()
        del U, R


def _sample(blocks, limit, seed=0):
    """Up to `limit` blocks, keeping the size-class mix."""
    if len(blocks) <= limit:
        return list(blocks)
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(len(blocks), size=limit, replace=False))
    return [blocks[i] for i in idx]


def tune(blocks, params, device="cuda:0", reps=3, probe_blocks=1024, sample_blocks=4096, cache=True, log=None,
         mode="rhs"):
    """Best Schedule for `blocks` (cached).  mode "rhs" times RHS launches;
    "rk4" times the order / grid candidates as hipGraph RK4 steps (whose RHS
    gathers ghosts from neighbours) on the whole workload, which is then not
    sampled.  Returns (schedule, report)."""
    device = torch.device(device)
    path = cache_path()
    key = workload_key(blocks, device) + "|" + mode
    store = _load_cache(path) if cache else {}
    if key in store:
        return Schedule.from_dict(store[key]["schedule"]), dict(store[key], cached=True)

    t0 = time.perf_counter()
    best = default_schedule()
    trials = []
    by_class = {}
    for b in blocks:
        by_class.setdefault(size_class(b["sz"]), []).append(b)
    # 1. tile extent per size class, each class alone
    for c, members in sorted(by_class.items()):
        probe, dof = _relocate(_sample(members, probe_blocks, seed=c))
        pts = layout.interior_points(probe)
        res = {}
        for tm in TILE_CANDIDATES:
            s = default_schedule()
            s.tile_max[c] = tm
            ms = _time_rhs(probe, dof, params, s, device, reps)
            res[tm] = ms
            trials.append({"class": c, "tile_max": tm, "ms": ms, "points_per_s": pts / (ms * 1e-3)})
            if log:
                log("autotune: class %d tile_max %d: %.3f ms (%d blocks)" % (c, tm, ms, len(probe)))
        best.tile_max[c] = min(res, key=res.get)
    # 2. order and grid on the whole (sampled) workload; RK4 steps need the
    # blocks' neighbours, so that mode keeps every block
    full, dof = _relocate(blocks if mode == "rk4" else _sample(blocks, sample_blocks))
    pts = layout.interior_points(full)
    res = {}
    for order in ORDER_CANDIDATES:
        for gm in GRID_CANDIDATES:
            s = Schedule.from_dict(dict(best.as_dict(), order=order, grid_mult=gm))
            ms = _time_rhs(full, dof, params, s, device, reps, mode)
            res[(order, gm)] = ms
            trials.append({"order": order, "grid_mult": gm, "ms": ms, "points_per_s": pts / (ms * 1e-3)})
            if log:
                log("autotune: order %d grid_mult %d: %.3f ms (%d blocks)" % (order, gm, ms, len(full)))
    order, gm = min(res, key=res.get)
    best.order = order
    best.grid_mult = gm
    default_ms = _time_rhs(full, dof, params, default_schedule(), device, reps, mode)
    report = {"schedule": best.as_dict(), "key": key, "mode": mode, "tuning_s": time.perf_counter() - t0,
              "best_ms": res[(order, gm)], "default_ms": default_ms, "sample_blocks": len(full),
              "trials": trials}
    if cache:
        store = _load_cache(path)
        store[key] = report
        _save_cache(path, store)
    return best, dict(report, cached=False)
