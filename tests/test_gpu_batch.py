"""Batched AO jobs in one persistent launch (HIP graph) against the same
jobs run one by one through each solver's run_all: per-job results must be
identical bit for bit, on every graph replay and after autotuning."""
import numpy as np
import pytest

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
    for shape in [None, (1, 1, 16), (3, 64, 32), (16, 8, 24), (2, 5, 16)]:
        if shape:
            b.set_launch(*shape)
        b.run()
        for j, r in zip(jobs, seq):
            assert np.array_equal(j.solver.getResults().view(np.uint32), r), (shape, j.spec)
    b2 = batch.Batch([j.solver for j in jobs])
    assert batch.tune(b2, jobs, cache_path=str(tmp_path / "tune.json"))["cached"]
