"""Oracle parity on the workloads as benchmarked and on every traversal path.

- heightfield2048 exactly as the AO rays/s figure is measured on it (2048^2
  map, 2,097,152-triangle heightfield): the culling mesh mapping of every
  texel, and AO-256, the paired AO + bent normals-256 pass the bake runs and
  thickness-128 on windows spread over the map, plus every texel whose
  thickness rays were redone on the reference tree.
- the torus-knot workload (culling mapping, AO-512 and thickness-256 at
  maxD 10) at 512^2 with a 256 x 256 tube: the mapping of every texel, and the
  texels whose rays overflowed the 16-entry traversal stack and were redone
  on the stackless culling tree.
- a flat, axis-aligned scene whose ray directions have exactly-zero
  components: the rays traced on the reference tree from the start.

The counting run of each pass (fornos_solver_count_work) reports which
texels took which path; the test asserts the path fired and compares those
texels with the oracle (Shaders/ao_step1.comp:111-165,
Shaders/thick_step1.comp:111-158, Shaders/bentnormals_step1.comp:123-171,
Shaders/meshmapping_nobackfaces.comp:102-199).
"""
import numpy as np
import pytest

import fornos_amd as fa
from bakes import MISS, Bake, same_bits, windows
from meshes import grid, merge, quad, soup_mesh

pytestmark = pytest.mark.gpu

FB_OVERFLOW, FB_REDO, FB_SLOW = fa.Solver.FB_OVERFLOW, fa.Solver.FB_REDO, fa.Solver.FB_SLOW


def run_and_count(s):
    """Runs a pass, then its counting run; the counting run must give the
    same results.  Returns (results, counts, work counters, fallback flags)."""
    s.run_all()
    res, cnt = s.getResults().copy(), s.counts().copy()
    w = s.count_work()
    assert same_bits(s.getResults(), res) and np.array_equal(s.counts(), cnt)
    return res, cnt, w, s.fallback_flags()


# ---- heightfield2048 as benchmarked ----------------------------------------

@pytest.fixture(scope="module")
def hf(gpu):
    return Bake("heightfield2048")


def test_heightfield2048_mapping_every_texel(hf):
    n = len(hf.tidx)
    assert n == hf.prep["texels"].count and n > 4_000_000
    coords, tidx = hf.oracle_map()
    assert np.array_equal(hf.tidx, tidx)
    assert same_bits(hf.coords, coords)
    assert np.sum(tidx == MISS) < n // 100


def test_heightfield2048_ao256_windows(hf):
    s = hf.solver("ao")
    assert s.params.sample_count == 256
    s.run_all()
    n = len(hf.tidx)
    assert hf.check_texels("ao", s, windows(n, 32, 256)) == 32 * 256


def test_heightfield2048_paired_ao_bn_windows(hf):
    # the bake traces AO and bent normals with one traversal
    ao, bn = hf.solver("ao"), hf.solver("bentnormals")
    ao.pair(bn)
    ao.run_all()
    n = len(hf.tidx)
    w = windows(n, 16, 256)
    hf.check_texels("ao", ao, w)
    hf.check_texels("bentnormals", bn, w)
    ao.pair(None)


def test_heightfield2048_thickness_windows_and_redos(hf):
    s = hf.solver("thickness")
    assert s.params.sample_count == 128
    res, cnt, w, flags = run_and_count(s)
    redo = np.flatnonzero(flags & FB_REDO)
    print("heightfield2048 thickness-128: %d rays redone on the reference tree, %d texels; %d overflows"
          % (w["redo_rays"], len(redo), w["stack_overflows"]))
    assert w["redo_rays"] > 0 and len(redo) > 0
    assert len(redo) <= w["redo_rays"]
    n = len(hf.tidx)
    hf.check_texels("thickness", s, np.concatenate([windows(n, 16, 256), redo]), res, cnt)


# ---- knot: the cfg's generator and pass parameters at 512^2 ---------------

@pytest.fixture(scope="module")
def knot(gpu):
    return Bake(fa.configs.scaled("knot4096", size=512, high_quads=256))


def test_knot_culling_mapping_every_texel(knot):
    assert knot.cull
    coords, tidx = knot.oracle_map()
    assert np.array_equal(knot.tidx, tidx)
    assert same_bits(knot.coords, coords)
    assert np.sum(tidx == MISS) < len(tidx) // 10


@pytest.mark.parametrize("kind", ["ao", "thickness", "bentnormals"])
def test_knot_stack_overflows_match_oracle(knot, kind):
    over = {"sampleCount": 128, "minDistance": 0.01, "maxDistance": 10.0} if kind == "bentnormals" else {}
    s = knot.solver(kind, **over)
    assert s.get_launch()[2] == 16
    res, cnt, w, flags = run_and_count(s)
    ovf = np.flatnonzero(flags & FB_OVERFLOW)
    print("knot %s: %d stack overflows (%d texels), %d redone on the reference tree, %d slow"
          % (kind, w["stack_overflows"], len(ovf), w["redo_rays"], w["slow_rays"]))
    assert w["stack_overflows"] > 0 and len(ovf) > 0
    n = len(knot.tidx)
    checked = knot.check_texels(kind, s, np.concatenate([np.flatnonzero(flags), windows(n, 8, 128)]), res, cnt)
    assert checked >= len(ovf)
    # a deeper stack takes none of them off the tree, same results
    s.set_launch(8, 48, 32)
    s.reset()
    s.run_all()
    assert same_bits(s.getResults(), res) and np.array_equal(s.counts(), cnt)
    assert s.count_work()["stack_overflows"] < w["stack_overflows"]
    # rays handed out in map-row order rather than by the origins' Morton code
    # (the automatic choice for maxD 10 on this scene): same results
    s.set_launch(8, 48, 16)
    s.set_ray_order(s.RAY_ORDER_TEXELS)
    s.reset()
    s.run_all()
    assert same_bits(s.getResults(), res) and np.array_equal(s.counts(), cnt)


# ---- slow directions: exact zero components ---------------------------------

# Sample directions (tangent-frame x, y, z) with exact zeros: on a plane whose
# normals are exactly +z the frame is (-x, y, z), so these rays have x = y = 0
# or one zero component; a few generic ones go through the fast path.
AXIS_SAMPLES = np.float32([(0, 0, 1), (0.6, 0, 0.8), (0, 0.6, 0.8), (-0.6, 0, 0.8), (0.36, 0.48, 0.8),
                           (0, -0.8, 0.6), (0.8, 0, 0.6), (-0.48, 0.36, 0.8)])


@pytest.fixture(scope="module")
def axis(gpu):
    ground = grid(16, z=0.0)
    under = grid(8, z=-0.05)
    roof = soup_mesh(quad(0.1, size=0.4, flip=True, x0=0.3, y0=0.3),
                     normals=np.tile(np.float32([0, 0, -1]), (6, 1)))
    h = merge(ground, under, roof)
    hi = fa.Mesh.from_arrays(h["positions"], h["vertices"], h["triangles"], h["normals"], h["texcoords"])
    g = grid(4, z=0.05)
    lo = fa.Mesh.from_arrays(g["positions"], g["vertices"], g["triangles"], g["normals"], g["texcoords"])
    return Bake.from_meshes(lo, hi, 48, cull=True)


@pytest.mark.parametrize("kind", ["ao", "bentnormals", "thickness"])
def test_slow_directions_match_oracle(axis, kind):
    table = np.concatenate([np.roll(AXIS_SAMPLES, k, axis=0) for k in range(4)])
    maxd = 0.2 if kind == "thickness" else 0.5
    s = axis.solver(kind, sampleCount=8, minDistance=0.01, maxDistance=maxd, table=table)
    res, cnt, w, flags = run_and_count(s)
    slow = np.flatnonzero(flags & FB_SLOW)
    live = int(np.sum(axis.tidx != MISS))
    print("axis scene %s: %d texels, %d slow rays, %d texels with one" % (kind, live, w["slow_rays"], len(slow)))
    assert live > 1000
    assert w["slow_rays"] > live  # most texels have several
    assert w["slow_rays"] < w["rays"]  # and the generic samples take the fast path
    # every texel, against the oracle
    axis.check_texels(kind, s, np.arange(len(axis.tidx)), res, cnt)
    if kind == "ao":
        assert np.any(cnt > 0) and np.any(cnt == 0)  # the roof occludes some texels, not all
