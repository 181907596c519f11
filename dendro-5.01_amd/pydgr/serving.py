"""Continuous-batching server for independent short BSSN evolutions.

A job is a uniform grid of 13^3 blocks with its own initial data, time step
and number of RK4 steps.  Jobs are admitted, first come first served, into
one device pool (24 state arrays + the RK4 work arrays) when both a free
range of the pool and the HBM budget allow; their state stays on the device
until they finish.  Every server iteration advances every resident job by one
RK4 step with single batched launches over the union of their blocks (each
job is one block group, so ghost refill never crosses jobs and each job keeps
its own dt).  Jobs arriving while others run are admitted at the next
iteration boundary as memory frees up, so their work overlaps on the device.

Admission also keeps the resident interior points under `max_resident_points`
(enough to fill the GPU): past that, adding jobs only stretches every resident
job's latency without raising throughput.

Because every kernel computes each point from its own block (and its group's
neighbours) only, a job's result is bitwise the same alone and in any mix.
"""
import bisect
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import data, layout
from .native import RK4, Plan, _block_dtype, make_params

BYTES_PER_POINT = 4 * 24 * 8  # state + 3 RK4 work arrays, FP64, 24 variables


@dataclass
class Job:
    """One evolution request."""
    job_id: int
    arrival: float                 # seconds after the server starts
    nb: int                        # nb^3 blocks of 13^3
    steps: int = 50
    kind: str = "flat"             # "flat" (perturbed flat space) or "bl" (Brill-Lindquist)
    seed: int = 0
    mass_ratio: float = 1.0
    dx: float = 1.0 / 64
    cfl: float = 0.25
    # filled by the server
    base: int = -1
    admitted: float = None
    finished: float = None
    result: torch.Tensor = None
    blocks: list = field(default=None, repr=False)
    dof: int = 0

    @property
    def dt(self):
        return self.cfl * self.dx

    @property
    def points(self):
        return self.nb ** 3 * 343

    def build(self):
        """Block table of the job (offsets from 0); Brill-Lindquist grids are
        centred 1.5 from the puncture at x = +4."""
        if self.blocks is None:
            blocks, self.dof = layout.uniform_grid(self.nb, dx=self.dx, outer_bflags=True)
            if self.kind == "bl":
                for b in blocks:
                    b["pmin"][0] += 2.5
                    b["pmax"][0] += 2.5
            self.blocks = blocks
        return self.blocks

    def initial_state(self, device):
        blocks = self.build()
        if self.kind == "bl":
            q = self.mass_ratio
            m1, m2 = q / (1.0 + q), 1.0 / (1.0 + q)
            return data.brill_lindquist(blocks, self.dof, m1=m1, m2=m2, sep=8.0, device=device)
        return data.flat_bumps(blocks, self.dof, seed=self.seed, amp=1e-3, width=0.1, device=device)


class PoolAllocator:
    """First-fit allocator of point ranges in the device pool."""

    def __init__(self, capacity, align=256):
        self.capacity = int(capacity)
        self.align = align
        self.free = [(0, self.capacity)]  # sorted (start, length)

    def alloc(self, n):
        n = -(-int(n) // self.align) * self.align
        for i, (s, l) in enumerate(self.free):
            if l >= n:
                if l == n:
                    self.free.pop(i)
                else:
                    self.free[i] = (s + n, l - n)
                return s, n
        return None

    def release(self, start, n):
        i = bisect.bisect(self.free, (start, 0))
        self.free.insert(i, (start, n))
        merged = []
        for s, l in self.free:
            if merged and merged[-1][0] + merged[-1][1] == s:
                merged[-1] = (merged[-1][0], merged[-1][1] + l)
            else:
                merged.append((s, l))
        self.free = merged


class Server:
    """Runs jobs to completion; returns per-job latency and throughput."""

    def __init__(self, hbm_budget_bytes, device="cuda:0", use_graph=True, params=None, keep_results=True,
                 max_resident_points=None):
        self.device = torch.device(device)
        self.max_resident_points = max_resident_points
        self.capacity = int(hbm_budget_bytes // BYTES_PER_POINT)
        self.U = torch.zeros((24, self.capacity), dtype=torch.float64, device=self.device)
        self.work = torch.empty(3 * 24 * self.capacity, dtype=torch.float64, device=self.device)
        self.alloc = PoolAllocator(self.capacity)
        self.use_graph = use_graph
        # chi floor of the reference's puncture parameter files (1e-4)
        self.params = params or make_params(ko_sigma=0.4, chi_floor=1e-4)
        self.keep_results = keep_results
        self.rebuilds = 0

    def _rebuild(self, resident):
        """Plan + evolver over the union of the resident jobs' blocks."""
        dt = _block_dtype()
        parts = []
        for g, j in enumerate(resident):
            a = np.zeros(len(j.blocks), dtype=dt)
            a["offset"] = [b["offset"] + j.base for b in j.blocks]
            a["sz"] = [b["sz"] for b in j.blocks]
            a["bflag"] = [b["bflag"] for b in j.blocks]
            a["pmin"] = [b["pmin"] for b in j.blocks]
            a["pmax"] = [b["pmax"] for b in j.blocks]
            a["group"] = g
            parts.append(a)
        arr = np.concatenate(parts)
        plan = Plan(arr, self.params)
        rk = RK4(plan, self.U, use_graph=self.use_graph, work=self.work)
        rk.set_group_dt([j.dt for j in resident])
        self.rebuilds += 1
        return plan, rk

    def run(self, jobs, realtime=True):
        pending = sorted(jobs, key=lambda j: (j.arrival, j.job_id))
        for j in pending:
            j.build()
        resident = []
        done = []
        plan = rk = None
        host_s = 0.0
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731
        while pending or resident:
            changed = False
            # admit in arrival order while the head of the queue fits
            th = time.perf_counter()
            while pending and (pending[0].arrival <= now() or not realtime):
                j = pending[0]
                if (self.max_resident_points and resident
                        and sum(r.points for r in resident) + j.points > self.max_resident_points):
                    break
                got = self.alloc.alloc(j.dof)
                if got is None:
                    if not resident:
                        raise RuntimeError("job %d (%d points) exceeds the pool" % (j.job_id, j.dof))
                    break
                pending.pop(0)
                j.base, j._len = got
                j.admitted = now()
                j._done = 0
                self.U[:, j.base:j.base + j.dof] = j.initial_state(self.device)
                resident.append(j)
                changed = True
            if not resident:
                time.sleep(max(0.0, min(0.01, pending[0].arrival - now())))
                continue
            if changed or rk is None:
                if rk is not None:
                    rk.close()
                    plan.close()
                plan, rk = self._rebuild(resident)
            host_s += time.perf_counter() - th
            rk.run(1)
            torch.cuda.current_stream().synchronize()
            finished = []
            for j in resident:
                j._done += 1
                if j._done == j.steps:
                    finished.append(j)
            if finished:
                t = now()
                for j in finished:
                    j.finished = t
                    if self.keep_results:
                        j.result = self.U[:, j.base:j.base + j.dof].clone()
                    self.alloc.release(j.base, j._len)
                    resident.remove(j)
                    done.append(j)
                if rk is not None:
                    rk.close()
                    plan.close()
                    plan = rk = None
        if rk is not None:
            rk.close()
            plan.close()
        makespan = now()
        lat = np.array([j.finished - j.arrival for j in done])
        rhs_points = sum(4 * j.steps * j.points for j in done)
        return {
            "jobs": len(done),
            "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else None,
            "makespan_s": makespan,
            "rhs_points_per_s": rhs_points / makespan if makespan > 0 else None,
            "plan_rebuilds": self.rebuilds,
            "admission_and_rebuild_s": host_s,
            "max_resident_points": self.max_resident_points,
        }


def poisson_jobs(n, rate=20.0, seed=0, min_blocks=64, max_blocks=4096, steps=50):
    """Poisson arrivals (mean `rate` jobs/s) of jobs with a seeded block count
    in [min_blocks, max_blocks] (rounded to a cube nb^3, nb >= 4) and initial
    data drawn from perturbed flat space or Brill-Lindquist with mass ratio
    in [1, 3]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        nblk = int(rng.integers(min_blocks, max_blocks + 1))
        nb = max(4, int(round(nblk ** (1.0 / 3.0))))
        kind = "bl" if rng.random() < 0.5 else "flat"
        jobs.append(Job(job_id=i, arrival=t, nb=nb, steps=steps, kind=kind, seed=int(rng.integers(1 << 30)),
                        mass_ratio=float(rng.uniform(1.0, 3.0))))
    return jobs
