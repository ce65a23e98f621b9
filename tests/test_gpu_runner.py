"""The native runner (FornosRunner, Src/fornos.cpp:53-251) against the Python
assembly of the same bake (fornos_amd.configs.prepare + bake), file for
file, on mesh files; its task order, progress and error reporting."""
import os

import numpy as np
import pytest

import fornos_amd as fa
from fornos_amd import _abi, bake as fb
from fornos_amd.runner import FornosParameters, FornosRunner, MeshMappingMethod, NormalImport
from meshes import write_obj

pytestmark = pytest.mark.gpu

MAPS = {"height": {"maxDistance": 0.05}, "position": {}, "normals": {"tangentSpace": True},
        "ao": {"sampleCount": 32, "minDistance": 0.01, "maxDistance": 0.3},
        "bentnormals": {"sampleCount": 16, "minDistance": 0.01, "maxDistance": 0.3, "tangentSpace": True},
        "thickness": {"sampleCount": 16, "minDistance": 0.0, "maxDistance": 0.1}}
ORDER = ["Mesh mapping", "Height", "Position", "Normals", "Ambient Occlusion", "Bent normals", "Thickness"]
MODES = {"import": NormalImport.Import, "face": NormalImport.ComputePerFace, "vertex": NormalImport.ComputePerVertex}
METHODS = {"smooth": MeshMappingMethod.Smooth, "lowpoly": MeshMappingMethod.LowPolyNormals,
           "hybrid": MeshMappingMethod.Hybrid}


@pytest.fixture(scope="module")
def meshes(tmp_path_factory):
    d = tmp_path_factory.mktemp("meshes")
    lo = fa.Mesh.generate(_abi.GEN_HEIGHTFIELD_LOW, 5, 0.05, 16, 0)
    hi = fa.Mesh.generate(_abi.GEN_HEIGHTFIELD, 5, 0.05, 64, 40)
    write_obj(d / "low.obj", lo.arrays())
    write_obj(d / "high.obj", hi.arrays())
    return str(d / "low.obj"), str(d / "high.obj")


def runner_params(lo, hi, out, size, lo_mode, mapping, edge=0.05):
    p = FornosParameters()
    sh = p.shared
    sh.loPolyMeshPath, sh.hiPolyMeshPath = lo, hi
    sh.loPolyMeshNormal = MODES[lo_mode]
    sh.texWidth = sh.texHeight = size
    sh.mapping, sh.mappingEdge = METHODS[mapping], edge
    p.height.enabled, p.height.outputPath, p.height.maxDistance = True, os.path.join(out, "height.png"), 0.05
    p.positions.enabled, p.positions.outputPath = True, os.path.join(out, "position.exr")
    p.normals.enabled, p.normals.outputPath = True, os.path.join(out, "normals.png")
    for dst, name in ((p.ao, "ao"), (p.bentNormals, "bentnormals"), (p.thickness, "thickness")):
        kw = MAPS[name]
        dst.enabled, dst.outputPath = True, os.path.join(out, name + ".png")
        dst.sampleCount, dst.minDistance, dst.maxDistance = kw["sampleCount"], kw["minDistance"], kw["maxDistance"]
    return p


@pytest.mark.parametrize("lo_mode,mapping", [("import", "smooth"), ("face", "hybrid"), ("vertex", "lowpoly")])
def test_runner_equals_python_assembly(gpu, meshes, tmp_path, lo_mode, mapping):
    lo, hi = meshes
    size = 192
    a, b = tmp_path / "runner", tmp_path / "python"
    os.makedirs(a)
    r = FornosRunner()
    ok, err = r.start(runner_params(lo, hi, str(a), size, lo_mode, mapping))
    assert ok, err
    names, last = [], {}
    while r.pending():
        t = r.currentTask()
        if not names or names[-1] != t.name():
            names.append(t.name())
        p = t.progress()
        assert 0.0 <= p <= 1.0 and p >= last.get(t.name(), 0.0)
        last[t.name()] = p
        r.run()
    assert names == ORDER
    cfg = {"low": lo, "high": hi, "size": size, "lowNormals": lo_mode, "mapping": mapping, "mappingEdge": 0.05,
           "cull": True, "maps": MAPS}
    rep = fb.bake(cfg, str(b), "png")
    for f in sorted(os.listdir(b)):
        assert open(a / f, "rb").read() == open(b / f, "rb").read(), f
    for kind, name in ((_abi.SOLVER_AO, "ao"), (_abi.SOLVER_THICKNESS, "thickness"), (_abi.SOLVER_HEIGHT, "height")):
        assert np.allclose(r.exported_range(kind), rep[name + "_range"], rtol=0, atol=0), name


def test_runner_without_hi_poly_and_without_outputs(gpu, meshes):
    lo, _ = meshes
    p = FornosParameters()
    p.shared.loPolyMeshPath = lo
    p.shared.texWidth, p.shared.texHeight = 128, 96
    p.ao.enabled, p.ao.sampleCount, p.ao.maxDistance = True, 8, 0.2
    r = FornosRunner()
    ok, err = r.start(p)
    assert ok, err
    r.run_all()
    assert not r.pending() and r.exported_range(_abi.SOLVER_AO) is None


def test_runner_reports_missing_texcoords(gpu, tmp_path):
    # MapUV::fromMesh returns nullptr (Src/fornos.cpp:150-153)
    m = fa.Mesh.generate(_abi.GEN_ICOSPHERE_WAVES, 1, 0.02, 1, 0)
    write_obj(tmp_path / "m.obj", m.arrays())
    p = FornosParameters()
    p.shared.loPolyMeshPath = str(tmp_path / "m.obj")
    p.shared.texWidth = p.shared.texHeight = 32
    ok, err = FornosRunner().start(p)
    assert not ok and err == "Low poly mesh is missing texture coordinates or normals information"


def test_submit_and_poll_equal_run_all(gpu):
    # the runner's pacing (submit in parts + poll) gives the results of run_all
    cfg = fa.configs.scaled("heightfield2048", size=256, high_quads=48)
    prep = fa.prepare(cfg)
    mm0 = fa.MeshMapping().init_from_mesh(prep["low"], prep["dirs"], 256, 256, prep["scene"], True)
    mm0.run_all()
    mm1 = fa.MeshMapping().init_from_mesh(prep["low"], prep["dirs"], 256, 256, prep["scene"], True)
    import ctypes as C
    from fornos_amd._abi import check, lib
    check(lib().fornos_meshmap_submit(mm1.handle, 7))
    prog, done = C.c_float(), C.c_int()
    while True:
        check(lib().fornos_meshmap_poll(mm1.handle, C.byref(prog), C.byref(done)))
        if done.value:
            break
    assert prog.value == 1.0
    c0, t0 = mm0.read()
    c1, t1 = mm1.read()
    assert np.array_equal(t0, t1) and np.array_equal(c0.view(np.uint32), c1.view(np.uint32))
    s0 = fa.AmbientOcclusionSolver(sampleCount=32, minDistance=0.01, maxDistance=0.2).init(None, mm0)
    s0.run_all()
    s1 = fa.AmbientOcclusionSolver(sampleCount=32, minDistance=0.01, maxDistance=0.2).init(None, mm1)
    check(lib().fornos_solver_submit(s1.handle, 5))
    while True:
        check(lib().fornos_solver_poll(s1.handle, C.byref(prog), C.byref(done)))
        if done.value:
            break
    assert np.array_equal(s0.getResults().view(np.uint32), s1.getResults().view(np.uint32))
