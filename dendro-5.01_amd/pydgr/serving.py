"""Continuous-batching server for independent short BSSN evolutions.

A job is a grid of 13^3 blocks (an exact block count: full layers of a cube,
then a partial layer) with its own initial data (perturbed flat space or
Brill-Lindquist), time step and number of RK4 steps.

The server keeps *slots*.  A slot is everything one job geometry needs on
the device: the plan of its block table, the state (24 arrays) and RK4 work
memory (a range of one device pool that holds the whole HBM budget, so no
device memory is allocated or freed while jobs run), the RK4 evolver with its
captured hipGraph, and a stream.  Admission,
first come first served, takes a free slot of the job's geometry (a job of a
geometry seen before reuses its plan and graph: no plan build, no capture),
or builds one when the HBM budget allows (evicting free slots of other
geometries if needed).  The job's initial data is written into the slot and
its RK4 steps are enqueued on its stream a few at a time, an event marking
the end of each chunk.  The host only polls events: it never waits for a
step, and
admitting or finishing a job never touches the plans or graphs of the jobs
already running.  Jobs on different slots run on different streams, so a
job that arrives while others are on the GPU starts at once and its kernels
overlap theirs.

Admission also keeps the running interior points under
`max_resident_points` (enough to fill the GPU): past that, adding jobs only
stretches every running job's latency without raising throughput.

Each job is evolved by its own plan from its own state, so its result is
bitwise the same alone and in any mix.
"""
import bisect
import threading
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import data, layout
from .native import RK4, Plan, lib, make_params


@dataclass
class Job:
    """One evolution request.  `nblocks` blocks of 13^3 (default nb^3)."""
    job_id: int
    arrival: float                 # seconds after the server starts
    nb: int = 4                    # nb^3 blocks unless nblocks is given
    steps: int = 50
    kind: str = "flat"             # "flat" (perturbed flat space) or "bl" (Brill-Lindquist)
    seed: int = 0
    mass_ratio: float = 1.0
    dx: float = 1.0 / 64
    cfl: float = 0.25
    nblocks: int = None
    # filled by the server
    admitted: float = None
    finished: float = None
    result: torch.Tensor = None
    blocks: list = field(default=None, repr=False)
    dof: int = 0

    def __post_init__(self):
        if self.nblocks is None:
            self.nblocks = self.nb ** 3

    @property
    def dt(self):
        return self.cfl * self.dx

    @property
    def points(self):
        return self.nblocks * 343

    @property
    def geometry(self):
        """Jobs with equal geometry share a plan: block count, grid shift, dx."""
        return (self.nblocks, self.kind == "bl", self.dx)

    def build(self):
        """Block table of the job (offsets from 0); Brill-Lindquist grids are
        centred 1.5 from the puncture at x = +4."""
        if self.blocks is None:
            blocks, self.dof = layout.block_grid(self.nblocks, dx=self.dx)
            if self.kind == "bl":
                for b in blocks:
                    b["pmin"][0] += 2.5
                    b["pmax"][0] += 2.5
            self.blocks = blocks
        return self.blocks

    def initial_state(self, device, out=None, plan=None):
        """The job's 24 initial arrays (written into `out` when given; flat
        jobs with the fused kernel when a plan of the job's blocks is
        given)."""
        blocks = self.build()
        if self.kind == "bl":
            q = self.mass_ratio
            m1, m2 = q / (1.0 + q), 1.0 / (1.0 + q)
            return data.brill_lindquist(blocks, self.dof, m1=m1, m2=m2, sep=8.0, device=device, out=out)
        return data.flat_bumps(blocks, self.dof, seed=self.seed, amp=1e-3, width=0.1, device=device, out=out,
                               plan=plan)


class PoolAllocator:
    """First-fit allocator of ranges of the server's device pool."""

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


def slot_doubles(job):
    """Pool doubles of a slot for `job`: state (24 x dof) + RK4 work
    (dgr_rk4_work_doubles = 48 x (dof + interior points) + 96 x 343 per
    tile; a 13^3 block is one tile of 343 interior points; Slot checks it
    against the library's figure)."""
    job.build()
    return 24 * job.dof + 48 * (job.dof + 343 * job.nblocks) + 96 * 343 * job.nblocks


class Slot:
    """Plan, state and RK4 work memory (a range of the pool), evolver
    (hipGraph) and stream of one job geometry."""

    def __init__(self, job, params, pool, rng, use_graph, stream, graph_lock):
        self.geometry = job.geometry
        self.range = rng
        self.plan = Plan(job.build(), params)
        self.dof = job.dof
        start = rng[0]
        self.U = pool[start:start + 24 * job.dof].view(24, job.dof)
        need = int(lib().dgr_rk4_work_doubles(self.plan.handle))
        assert 24 * job.dof + need <= rng[1]
        self.work = pool[start + 24 * job.dof:start + 24 * job.dof + need]
        self.rk = RK4(self.plan, self.U, use_graph=use_graph, work=self.work)
        self.dt = job.dt
        self.rk.set_group_dt([job.dt])
        # capture the step graph now (run(0) captures and launches nothing);
        # captures, graph launches and graph destruction are serialised
        # across threads (graph_lock)
        with graph_lock:
            self.rk.run(0, stream=stream)
        self.stream = None
        self.done = torch.cuda.Event()
        self.job = None
        self.last_used = 0

    def prepare(self, job, stream):
        """Initial data of `job` written into the slot's state, enqueued on
        `stream`, which the job's steps then use (the temporaries hold a few
        variables at a time: torch caches freed memory per stream)."""
        self.job = job
        self.stream = stream
        self.steps_left = job.steps
        if job.dt != self.dt:
            self.rk.set_group_dt([job.dt])  # stream-ordered on the evolver's stream
            self.dt = job.dt
        with torch.cuda.stream(self.stream):
            job.initial_state(self.U.device, out=self.U)
        # ghosts no neighbour fills (concave corners of a partial cube) keep
        # this job's initial values in the stage buffer too
        self.rk.reset_stage(stream=self.stream)

    def launch(self, nsteps, keep_result):
        """The job's next `nsteps` steps, then the chunk's end event (after
        the last chunk, first a copy of the result if asked), enqueued on the
        slot's stream; returns at once."""
        self.rk.run(nsteps, stream=self.stream)
        self.steps_left -= nsteps
        with torch.cuda.stream(self.stream):
            if keep_result and self.steps_left == 0:
                self.job.result = self.U.clone()
            self.done.record(self.stream)

    def retire(self):
        """Stop using the slot's pool range (the caller frees it; the slot is
        free, so its last job has ended); the plan and evolver stay until
        close()."""
        self.U = self.work = self.rk.u = None

    def close(self):
        self.done.synchronize()
        self.rk.close()
        self.plan.close()
        self.U = self.work = None


class Server:
    """Runs jobs to completion; returns per-job latency and throughput.

    Preparer threads take jobs in arrival order as they arrive and ready a
    slot for each (pool memory, plan and evolver when the geometry has no
    free slot, initial data): the calls that can wait on device queues
    (plan set-up, allocation) happen there, several at a time and ahead of
    time, while the device runs other jobs.  The main loop launches prepared
    jobs in arrival order while the running points allow, and polls end
    events; it makes no blocking device call."""

    RETIRED_BATCH = 16   # retired slots destroyed together (one device sync)
    PREPARERS = 4        # preparer threads
    PREP_AHEAD = 6       # jobs claimed by preparers and not yet launched, at most
    STREAMS = 8          # streams the jobs take turns on
    CHUNK = 5            # RK4 steps enqueued at a time per job: a set-up call
                         # that waits behind queued work waits for one chunk

    def __init__(self, hbm_budget_bytes, device="cuda:0", use_graph=True, params=None, keep_results=True,
                 max_resident_points=None):
        self.device = torch.device(device)
        self.budget = int(hbm_budget_bytes)
        self.max_resident_points = max_resident_points
        self.use_graph = use_graph
        # chi floor of the reference's puncture parameter files (1e-4)
        self.params = params or make_params(ko_sigma=0.4, chi_floor=1e-4)
        self.keep_results = keep_results
        # the whole budget as one device pool: slots take ranges of it, so no
        # device memory is allocated or freed while jobs run
        self.pool = torch.empty(self.budget // 8, dtype=torch.float64, device=self.device)
        self.alloc = PoolAllocator(self.pool.numel())
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(self.STREAMS)]
        self.slots = []          # slots holding a pool range (busy, ready or free)
        self.retired = []        # evicted slots: range returned, plan and evolver not yet destroyed
        self.slots_built = 0
        self.slots_evicted = 0
        self.clock = 0
        self.lock = threading.Condition()
        self.graph_lock = threading.Lock()
        # host seconds by activity: preparing jobs (preparer thread: building
        # slots, initial data, destroying retired slots), and the main loop's
        # launches and polls
        self.host = {"build": 0.0, "prepare": 0.0, "destroy": 0.0, "launch": 0.0, "poll": 0.0}

    def _reserve(self, job):
        """Under the lock: a free slot of the job's geometry (marked taken),
        or a pool range for a new one (evicting free slots, least recently
        used first), or None (wait for memory)."""
        for s in self.slots:
            if s.job is None and s.geometry == job.geometry:
                s.job = job
                return s
        need = slot_doubles(job)
        if need > self.pool.numel():
            raise RuntimeError("job %d (%d bytes) exceeds the HBM budget" % (job.job_id, 8 * need))
        rng = self.alloc.alloc(need)
        free = sorted((s for s in self.slots if s.job is None), key=lambda s: s.last_used)
        while rng is None and free:
            victim = free.pop(0)  # least recently used free slot
            victim.retire()
            self.alloc.release(*victim.range)
            self.slots.remove(victim)
            self.retired.append(victim)
            self.slots_evicted += 1
            rng = self.alloc.alloc(need)
        return rng

    def _stream(self):
        """Under the lock: the stream fewest running or prepared jobs use."""
        use = [0] * self.STREAMS
        for s in self.slots:
            if s.job is not None:
                for k, st in enumerate(self.streams):
                    if s.stream is st:
                        use[k] += 1
        return self.streams[use.index(min(use))]

    def _preparer(self, jobs, prepared, state, now, realtime, stop):
        """Prepares jobs in arrival order (several of these threads run
        side by side, each on the next unclaimed job)."""
        try:
            torch.cuda.set_device(self.device)
            while not stop[0]:
                with self.lock:
                    # at most PREP_AHEAD jobs claimed but not yet launched
                    while (not stop[0] and state["next"] < len(jobs)
                           and state["next"] - state["launched"] >= self.PREP_AHEAD):
                        self.lock.wait(0.01)
                    if stop[0] or state["next"] >= len(jobs):
                        return
                    i = state["next"]
                    state["next"] += 1
                j = jobs[i]
                while realtime and j.arrival > now() and not stop[0]:
                    time.sleep(min(j.arrival - now(), 0.002))
                while not stop[0]:
                    with self.lock:
                        # memory is reserved in arrival order: a later job
                        # never holds what an earlier one waits for
                        if state["resv"] != i:
                            self.lock.wait(0.01)
                            continue
                        got = self._reserve(j)
                        if got is None:
                            self.waits += 1
                            self.lock.wait(0.01)  # memory frees when a job ends
                            continue
                        state["resv"] += 1
                        self.lock.notify_all()
                        stream = self._stream()
                        old = []
                        if len(self.retired) >= self.RETIRED_BATCH:
                            old, self.retired = self.retired, []
                    break
                if stop[0]:
                    return
                # outside the lock: the calls that may wait on device queues
                if old:
                    t = time.perf_counter()
                    for r in old:
                        r.done.synchronize()
                        with self.graph_lock:
                            r.close()
                    self._add("destroy", time.perf_counter() - t)
                if isinstance(got, Slot):
                    s = got
                else:
                    t = time.perf_counter()
                    s = Slot(j, self.params, self.pool, got, self.use_graph, stream, self.graph_lock)
                    s.job = j
                    self._add("build", time.perf_counter() - t)
                    with self.lock:
                        self.slots.append(s)
                        self.slots_built += 1
                t = time.perf_counter()
                s.prepare(j, stream)
                self._add("prepare", time.perf_counter() - t)
                with self.lock:
                    prepared[i] = s
                    self.lock.notify_all()
        except BaseException as e:  # surfaced by run()
            self.error = e

    def _add(self, key, dt):
        with self.lock:
            self.host[key] += dt

    def run(self, jobs, realtime=True):
        jobs = sorted(jobs, key=lambda j: (j.arrival, j.job_id))
        for j in jobs:
            j.build()
        n = len(jobs)
        prepared, running, done = {}, [], []
        state = {"next": 0, "resv": 0, "launched": 0}
        self.running = running
        self.waits = 0
        self.error = None
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731
        stop = [False]
        preps = [threading.Thread(target=self._preparer, args=(jobs, prepared, state, now, realtime, stop),
                                  daemon=True) for _ in range(self.PREPARERS)]
        for p in preps:
            p.start()
        try:
            while len(done) < n:
                if self.error is not None:
                    raise self.error
                th = time.perf_counter()
                with self.lock:
                    # a finished chunk: the next one, or the job is done and
                    # its slot free (nothing is destroyed)
                    for s in list(running):
                        if s.done.query():
                            if s.steps_left > 0:
                                with self.graph_lock:
                                    s.launch(min(self.CHUNK, s.steps_left), self.keep_results)
                                continue
                            s.job.finished = now()
                            done.append(s.job)
                            s.job = None
                            running.remove(s)
                            self.lock.notify_all()
                    t1 = time.perf_counter()
                    self.host["poll"] += t1 - th
                    # launch prepared jobs in arrival order while the running points allow
                    while state["launched"] in prepared:
                        s = prepared[state["launched"]]
                        if (self.max_resident_points and running and
                                sum(r.job.points for r in running) + s.job.points > self.max_resident_points):
                            break
                        del prepared[state["launched"]]
                        state["launched"] += 1
                        s.job.admitted = now()
                        self.clock += 1
                        s.last_used = self.clock
                        with self.graph_lock:
                            s.launch(min(self.CHUNK, s.steps_left), self.keep_results)
                        running.append(s)
                        self.lock.notify_all()
                    self.host["launch"] += time.perf_counter() - t1
                time.sleep(0.0005)
        finally:
            stop[0] = True
            with self.lock:
                self.lock.notify_all()
            for p in preps:
                p.join()
        makespan = now()
        torch.cuda.synchronize(self.device)
        lat = np.array([j.finished - j.arrival for j in done])
        rhs_points = sum(4 * j.steps * j.points for j in done)
        return {
            "jobs": len(done),
            "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else None,
            "makespan_s": makespan,
            "rhs_points_per_s": rhs_points / makespan if makespan > 0 else None,
            "slots_built": self.slots_built,
            "slots_evicted": self.slots_evicted,
            "admission_waits": self.waits,
            "main_loop_host_s": self.host["launch"] + self.host["poll"],
            "host_s_by_activity": dict(self.host),
            "max_resident_points": self.max_resident_points,
        }

    def close(self):
        for s in self.slots + self.retired:
            s.close()
        self.slots, self.retired = [], []
        self.pool = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def poisson_jobs(n, rate=20.0, seed=0, min_blocks=64, max_blocks=4096, steps=50):
    """Poisson arrivals (mean `rate` jobs/s) of jobs with a seeded block count
    in [min_blocks, max_blocks] (exact: full cube layers, then a partial
    layer) and initial data drawn from perturbed flat space or
    Brill-Lindquist with mass ratio in [1, 3]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        nblk = int(rng.integers(min_blocks, max_blocks + 1))
        nb = 1
        while nb ** 3 < nblk:
            nb += 1
        kind = "bl" if rng.random() < 0.5 else "flat"
        jobs.append(Job(job_id=i, arrival=t, nb=nb, nblocks=nblk, steps=steps, kind=kind,
                        seed=int(rng.integers(1 << 30)), mass_ratio=float(rng.uniform(1.0, 3.0))))
    return jobs
