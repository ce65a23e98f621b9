"""Batched AO jobs in one persistent launch (HIP graph) against the same
jobs run one by one through each solver's run_all: per-job results must be
identical bit for bit, on every graph replay and after autotuning."""
import numpy as np
import pytest

import fornos_amd as fa
from fornos_amd import batch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def jobs(gpu):
    return [batch.Job(s) for s in batch.job_specs(10, scale=0.02)]


def test_batch_equals_sequential(jobs):
    seq = []
    for j in jobs:
        j.solver.reset()
        j.solver.run_all()
        seq.append((j.solver.getResults().view(np.uint32).copy(), j.solver.counts().copy()))
    assert sum(j.rays for j in jobs) > 0
    b = batch.Batch([j.solver for j in jobs])
    for rep in range(3):
        b.run()
        for j, (r, c) in zip(jobs, seq):
            assert np.array_equal(j.solver.counts(), c), (rep, j.spec)
            assert np.array_equal(j.solver.getResults().view(np.uint32), r), (rep, j.spec)
        # a sequential run in between must not disturb the next replay
        jobs[0].solver.reset()
        jobs[0].solver.run_all()
    b.close()


def test_batch_after_autotune_and_launch_changes(jobs, tmp_path):
    seq = []
    for j in jobs:
        j.solver.reset()
        j.solver.run_all()
        seq.append(j.solver.getResults().view(np.uint32).copy())
    b = batch.Batch([j.solver for j in jobs])
    t = batch.tune(b, jobs, cache_path=str(tmp_path / "tune.json"), fraction=0.1)
    assert not t["cached"] and b.get_launch() == (t["blocks_per_cu"], t["refill"], t["stack"])
    assert b.get_tile_shift() == t["tile_shift"] in (-1, 0, 1)
    for k, shape in enumerate([None, (1, 1, 16), (3, 64, 32), (16, 8, 24), (2, 5, 16)]):
        if shape:
            b.set_launch(*shape)
        b.set_tile_shift((0, -1, 1, 1, -1)[k])
        base = [j.solver.get_tile() for j in jobs]
        shifted = [max(1, t0 // 2) if b.get_tile_shift() < 0 else min(32, 2 * t0) if b.get_tile_shift() > 0 else t0
                   for t0 in base]
        assert [b.get_tile(j.solver) for j in jobs] == shifted
        b.run()
        for j, r in zip(jobs, seq):
            assert np.array_equal(j.solver.getResults().view(np.uint32), r), (shape, j.spec)
    b2 = batch.Batch([j.solver for j in jobs])
    c = batch.tune(b2, jobs, cache_path=str(tmp_path / "tune.json"))
    assert c["cached"] and b2.get_tile_shift() == t["tile_shift"]
    with pytest.raises(fa.FornosError):
        b2.set_tile_shift(2)
