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
    assert all(64 <= j.nblocks <= 4096 and (j.nb - 1) ** 3 < j.nblocks <= j.nb ** 3 for j in j1)
    for j in j1[:3]:
        blocks, _ = layout.block_grid(j.nblocks)
        assert len(blocks) == j.nblocks
    assert all(np.diff([j.arrival for j in j1]) > 0)


def test_initial_state_written_in_place_is_the_same():
    # the server writes a job's initial data straight into its pool range
    for kind in ("flat", "bl"):
        j = Job(0, 0.0, nblocks=70, kind=kind, seed=5, mass_ratio=2.0)
        ref = j.initial_state("cpu")
        out = torch.full((24, j.dof), 3.0, dtype=torch.float64)
        assert j.initial_state("cpu", out=out) is out
        assert torch.equal(out, ref)
    with pytest.raises(ValueError):
        j.initial_state("cpu", out=torch.zeros((24, j.dof + 1), dtype=torch.float64))


@pytest.mark.gpu
def test_job_results_identical_alone_and_in_mix(cuda):
    def mk():
        return [Job(job_id=0, arrival=0.0, nb=2, steps=3, kind="flat", seed=11),
                Job(job_id=1, arrival=0.0, nb=3, steps=2, kind="bl", mass_ratio=2.0),
                Job(job_id=2, arrival=0.0, nb=2, steps=4, kind="bl", mass_ratio=1.3),
                Job(job_id=3, arrival=0.0, nb=2, steps=1, kind="flat", seed=5)]
    # budget for the 27-block job and one 8-block job at a time
    budget = 60 * 10 ** 6
    mix = mk()
    srv = Server(budget, device=cuda)
    stats = srv.run(mix, realtime=False)
    srv.close()
    # the three geometries (71.4 MB) never fit together: a job waited for
    # memory or a free slot of another geometry was evicted
    assert stats["jobs"] == 4 and stats["admission_waits"] + stats["slots_evicted"] >= 1, stats
    for j in mk():
        srv = Server(budget, device=cuda)
        srv.run([j], realtime=False)
        srv.close()
        m = [x for x in mix if x.job_id == j.job_id][0]
        assert torch.equal(j.result, m.result), j.job_id


@pytest.mark.gpu
def test_same_geometry_reuses_its_slot_and_partial_grids(cuda):
    """Jobs of one geometry share a slot (plan and graph built once); a
    block count that is not a cube runs as a partial cube, whose concave
    corners keep the job's own initial values, so a job gives the same bits
    in a reused slot as in a fresh one."""
    jobs = [Job(job_id=i, arrival=0.0, nblocks=10, steps=2, kind="flat", seed=i) for i in range(3)]
    # a budget for one slot of this geometry (16.6 MB): the jobs take turns in it
    srv = Server(20 * 10 ** 6, device=cuda)
    stats = srv.run(jobs, realtime=False)
    srv.close()
    assert stats["jobs"] == 3 and stats["slots_built"] == 1 and stats["admission_waits"] >= 1, stats
    for j in jobs:
        assert len(j.blocks) == 10 and torch.isfinite(j.result).all()
    again = Job(job_id=9, arrival=0.0, nblocks=10, steps=2, kind="flat", seed=1)
    srv = Server(10 ** 9, device=cuda)
    srv.run([again], realtime=False)
    srv.close()
    assert torch.equal(again.result, jobs[1].result)


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
