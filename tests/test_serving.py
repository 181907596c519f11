"""Serving loop: allocator logic on the CPU; on the GPU, every job's result
is bitwise the same alone and in a mix (with admission waits forced by a
small HBM budget), and matches the oracle's RK4 evolution."""
import numpy as np
import pytest
import torch

from pydgr import layout
from pydgr.serving import Job, PoolAllocator, Server, poisson_jobs


def test_pool_allocator_first_fit_and_merge():
    a = PoolAllocator(1024, align=64)
    x = a.alloc(100)
    y = a.alloc(200)
    z = a.alloc(300)
    assert x == (0, 128) and y == (128, 256) and z == (384, 320)
    assert a.alloc(400) is None
    a.release(*y)
    assert a.alloc(250) == (128, 256)
    a.release(128, 256)
    a.release(*x)
    a.release(*z)
    assert a.free == [(0, 1024)]


def test_poisson_jobs_are_seeded():
    j1 = poisson_jobs(10, seed=3)
    j2 = poisson_jobs(10, seed=3)
    assert [(j.arrival, j.nb, j.kind) for j in j1] == [(j.arrival, j.nb, j.kind) for j in j2]
    assert all(4 <= j.nb <= 16 for j in j1)
    assert all(np.diff([j.arrival for j in j1]) > 0)


@pytest.mark.gpu
def test_job_results_identical_alone_and_in_mix(cuda):
    def mk():
        return [Job(job_id=0, arrival=0.0, nb=2, steps=3, kind="flat", seed=11),
                Job(job_id=1, arrival=0.0, nb=3, steps=2, kind="bl", mass_ratio=2.0),
                Job(job_id=2, arrival=0.0, nb=2, steps=4, kind="bl", mass_ratio=1.3),
                Job(job_id=3, arrival=0.0, nb=2, steps=1, kind="flat", seed=5)]
    # budget for about two small jobs at a time: later jobs wait for memory
    budget = 80 * 2197 * 768
    mix = mk()
    stats = Server(budget, device=cuda).run(mix, realtime=False)
    assert stats["jobs"] == 4 and stats["plan_rebuilds"] >= 3
    for j in mk():
        Server(budget, device=cuda).run([j], realtime=False)
        m = [x for x in mix if x.job_id == j.job_id][0]
        assert torch.equal(j.result, m.result), j.job_id


@pytest.mark.gpu
def test_served_job_matches_oracle(cuda):
    import pyoracle as po
    j = Job(job_id=0, arrival=0.0, nb=2, steps=2, kind="flat", seed=11)
    Server(64 * 2197 * 768, device=cuda).run([j], realtime=False)
    U = j.initial_state("cpu").numpy()
    ref = U.copy()
    for _ in range(2):
        ref, bad = po.rk4_step_uniform(ref, j.blocks, 2, j.dt, po.params(ko_sigma=0.4))
    ii = layout.interior_index(j.blocks)
    got = j.result.cpu().numpy()
    change = np.abs(ref[:, ii] - U[:, ii]).max(1)
    rel = np.abs(got[:, ii] - ref[:, ii]).max(1) / np.where(change > 0, change, 1.0)
    assert rel.max() <= 1e-11, rel
