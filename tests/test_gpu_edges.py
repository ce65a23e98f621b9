"""Edge cases of the device path: an empty texel set (no texel centre inside
any UV triangle) through the mapping, every solver, the exports and the
runner; and a one-texel map."""
import os

import numpy as np
import pytest

import fornos_amd as fa
from fornos_amd import _abi
from fornos_amd import export as fx
from meshes import soup_mesh, write_obj

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scene(gpu):
    return fa.Scene(fa.Mesh.generate(_abi.GEN_HEIGHTFIELD, 3, 0.05, 16, 8))


def empty_low():
    # a sliver in UV space between texel centres of a 4 x 4 map
    e = 0.01
    m = soup_mesh([[(0, 0, 0), (1, 0, 0), (0, 1, 0)]], uvs=[[(0.30, 0.30), (0.30 + e, 0.30), (0.30, 0.30 + e)]])
    return m, fa.Mesh.from_arrays(m["positions"], m["vertices"], m["triangles"], m["normals"], m["texcoords"])


SOLVERS = [(fa.HeightSolver, {}), (fa.PositionSolver, {}), (fa.NormalsSolver, {"tangentSpace": False}),
           (fa.AmbientOcclusionSolver, {"sampleCount": 8}), (fa.BentNormalsSolver, {"sampleCount": 8, "tangentSpace": False}),
           (fa.ThicknessSolver, {"sampleCount": 8})]


def test_empty_texel_set(scene):
    _, lo = empty_low()
    mm = fa.MeshMapping().init_from_mesh(lo, None, 4, 4, scene, True)
    assert mm.count == 0
    mm.run_all()
    assert mm.runStep()  # nothing left after run_all
    for cls, kw in SOLVERS:
        s = cls(**kw).init(None, mm)
        s.run_all()
        assert s.runStep()
        assert s.getResults().size == 0
    ao = fa.AmbientOcclusionSolver(sampleCount=8).init(None, mm)
    ao.run_all()
    rgb, rng = fx.float_rgb8(ao, (0.0, 0.0), True, 16)
    assert not rgb.any()
    assert rng[0] == np.float32(3.4028234663852886e38) and rng[1] == -np.float32(3.4028234663852886e38)


def test_one_texel_map(scene):
    # two triangles covering most of the UV square: a 1 x 1 map has one
    # texel (UVs at exactly 0 would round to -1 and wrap: the reference
    # rasterises nothing then, and so do both rasters here)
    a, b = 0.01, 0.99
    m = soup_mesh([[(0, 0, 0.5), (1, 0, 0.5), (1, 1, 0.5)], [(0, 0, 0.5), (1, 1, 0.5), (0, 1, 0.5)]],
                  uvs=[[(a, a), (b, a), (b, b)], [(a, a), (b, b), (a, b)]])
    lo = fa.Mesh.from_arrays(m["positions"], m["vertices"], m["triangles"], m["normals"], m["texcoords"])
    host = fa.CompressedMapUV(lo, None, 1, 1).arrays()
    mm = fa.MeshMapping().init_from_mesh(lo, None, 1, 1, scene, True)
    assert mm.count == len(host["indices"]) == 1
    mm.run_all()
    ao = fa.AmbientOcclusionSolver(sampleCount=16, minDistance=0.0, maxDistance=1.0).init(None, mm)
    ao.run_all()
    assert 0.0 <= ao.getResults()[0] <= 1.0


def test_runner_with_no_texels(gpu, tmp_path):
    m, _ = empty_low()
    write_obj(tmp_path / "low.obj", m)
    from fornos_amd.runner import FornosParameters, FornosRunner
    p = FornosParameters()
    p.shared.loPolyMeshPath = str(tmp_path / "low.obj")
    p.shared.hiPolyMeshPath = ""
    p.shared.texWidth = p.shared.texHeight = 4
    p.ao.enabled, p.ao.sampleCount, p.ao.outputPath = True, 8, str(tmp_path / "ao.png")
    p.normals.enabled, p.normals.outputPath = True, str(tmp_path / "n.tga")
    r = FornosRunner()
    ok, err = r.start(p)
    assert ok, err
    r.run_all()
    assert os.path.getsize(tmp_path / "ao.png") > 0 and os.path.getsize(tmp_path / "n.tga") == 18 + 4 * 4 * 3
