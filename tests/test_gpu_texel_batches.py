"""Passes of fewer than 256 samples hand their texels out in batches (several
texels per atomic of the launch's counter, a variant of the hemisphere
kernel).  Checked against the oracle on every texel of a map where most
texels miss (a low-poly quad twice as wide as the heightfield under it), so
batches mix texels that cast rays with texels that cast none, for AO, bent
normals, thickness, the paired AO + bent-normal pass and a texel list; and
the batch kernel, which takes texel batches when one of its jobs has fewer
than 256 samples, against each job's own pass."""
import numpy as np
import pytest

import fornos_amd as fa
from bakes import MISS, Bake, same_bits
from fornos_amd import _abi
from meshes import soup_mesh

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def holes(gpu):
    hi = fa.Mesh.generate(_abi.GEN_HEIGHTFIELD, 3, 0.05, 24, 5)
    p = hi.arrays()["positions"]
    lo_xy, hi_xy = p[:, :2].min(0), p[:, :2].max(0)
    c, e = (lo_xy + hi_xy) / 2, (hi_xy - lo_xy).max()
    z = float(p[:, 2].max()) + 0.05
    x0, x1, y0, y1 = c[0] - e, c[0] + e, c[1] - e, c[1] + e
    corners = [[(x0, y0, z), (x1, y0, z), (x1, y1, z)], [(x0, y0, z), (x1, y1, z), (x0, y1, z)]]
    uvs = [[(0.01, 0.01), (0.99, 0.01), (0.99, 0.99)], [(0.01, 0.01), (0.99, 0.99), (0.01, 0.99)]]
    m = soup_mesh(corners, uvs=uvs)
    lo = fa.Mesh.from_arrays(m["positions"], m["vertices"], m["triangles"], m["normals"], m["texcoords"])
    b = Bake.from_meshes(lo, hi, 64, cull=True)
    miss = int(np.sum(b.tidx == MISS))
    assert 0.5 * len(b.tidx) < miss < len(b.tidx)  # most texels miss, some hit
    return b


@pytest.mark.parametrize("S", [16, 48])
@pytest.mark.parametrize("kind", ["ao", "bentnormals", "thickness"])
def test_small_sample_passes_match_oracle(holes, kind, S):
    s = holes.solver(kind, sampleCount=S, minDistance=0.001, maxDistance=0.3)
    s.run_all()
    assert holes.check_texels(kind, s, np.arange(len(holes.tidx))) == len(holes.tidx)
    # a pass over the texels listed in reverse order: batches of other
    # texels, the same results
    ref = s.getResults().view(np.uint32).copy()
    s.reset()
    s.run_texels(np.arange(len(holes.tidx))[::-1].copy())
    assert np.array_equal(s.getResults().view(np.uint32), ref)


def test_small_sample_paired_pass_matches_separate_passes(holes):
    S = 16
    ao0 = holes.solver("ao", sampleCount=S, minDistance=0.001, maxDistance=0.3)
    bn0 = holes.solver("bentnormals", sampleCount=S, minDistance=0.001, maxDistance=0.3)
    ao0.run_all()
    bn0.run_all()
    ao1 = holes.solver("ao", sampleCount=S, minDistance=0.001, maxDistance=0.3)
    bn1 = holes.solver("bentnormals", sampleCount=S, minDistance=0.001, maxDistance=0.3)
    ao1.pair(bn1)
    ao1.run_all()
    assert same_bits(ao0.getResults(), ao1.getResults()) and np.array_equal(ao0.counts(), ao1.counts())
    assert same_bits(bn0.getResults(), bn1.getResults()) and np.array_equal(bn0.counts(), bn1.counts())
    holes.check_texels("bentnormals", bn1, np.arange(len(holes.tidx)))


def test_batch_kernel_with_small_sample_jobs(holes):
    # the batch kernel hands out texel batches when a job's tile is above 1
    # (each job keeps its solver's tile); jobs of 16, 256 (tile 4) and 48
    # samples in one launch equal their passes.  Each job's tile must reach
    # the job table: a tile of 0 there once kept the kernel from advancing
    # its counter.
    jobs = [holes.solver("ao", sampleCount=S, minDistance=0.001, maxDistance=0.3) for S in (16, 256, 48)]
    jobs[1].set_tile(4)
    refs = []
    for j in jobs:
        j.run_all()
        refs.append(j.getResults().view(np.uint32).copy())
    b = fa.Batch(jobs)
    for _ in range(2):
        b.run()
        for j, r in zip(jobs, refs):
            assert np.array_equal(j.getResults().view(np.uint32), r)
    b.close()
