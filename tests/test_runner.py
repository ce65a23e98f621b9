"""The runner's parameters and start() checks (Src/fornos.h:35-149,
Src/fornos.cpp:53-153) on the CPU: defaults, the error strings start()
reports before any device work, and the seeded workload keys of prepare()."""
import ctypes as C

import numpy as np
import pytest

import fornos_amd as fa
from fornos_amd import _abi, configs
from fornos_amd._abi import lib
from fornos_amd.runner import FornosParameters, FornosRunner, MeshMappingMethod, NormalImport
from meshes import write_obj


def test_python_defaults_equal_native_defaults():
    c = _abi.Parameters()
    lib().fornos_parameters_defaults(C.byref(c))
    p = FornosParameters().to_c()
    for f, _ in _abi.ParamsShared._fields_:
        if not f.endswith("_path"):
            assert getattr(c.shared, f) == getattr(p.shared, f), f
    for m, _ in _abi.Parameters._fields_[1:]:
        for f, _ in _abi.ParamsMap._fields_:
            if f != "output_path":
                assert getattr(getattr(c, m), f) == getattr(getattr(p, m), f), (m, f)
    # the reference's member initialisers (Src/fornos.h:38-112)
    assert (c.shared.tex_width, c.shared.tex_height, c.shared.tex_dilation, c.shared.bvh_tris_per_node) == (2048, 2048, 16, 8)
    assert c.shared.ignore_backfaces == 1 and c.shared.mapping == MeshMappingMethod.Smooth
    assert np.float32(c.shared.mapping_edge) == np.float32(0.05)
    assert (c.ao.sample_count, np.float32(c.ao.min_distance), c.ao.max_distance) == (256, np.float32(0.01), 10.0)
    assert c.height.max_distance == 0.0 and c.height.normalize_output == 1
    assert c.normals.tangent_space == 1 and c.bent_normals.tangent_space == 1
    assert not any(getattr(c, m).enabled for m, _ in _abi.Parameters._fields_[1:])


@pytest.fixture(scope="module")
def lowpoly(tmp_path_factory):
    d = tmp_path_factory.mktemp("runner")
    lo = fa.Mesh.generate(_abi.GEN_HEIGHTFIELD_LOW, 3, 0.05, 8, 0)
    path = d / "low.obj"
    write_obj(path, lo.arrays())
    return str(path)


def start(p):
    return FornosRunner().start(p)


def test_missing_low_poly_mesh(tmp_path):
    p = FornosParameters()
    ok, err = start(p)
    assert not ok and err.startswith("Missing low poly mesh")
    p.shared.loPolyMeshPath = str(tmp_path / "nothing.obj")
    ok, err = start(p)
    assert not ok and err.startswith("Missing low poly mesh")


def test_missing_high_poly_mesh(lowpoly, tmp_path):
    p = FornosParameters()
    p.shared.loPolyMeshPath = lowpoly
    p.shared.hiPolyMeshPath = str(tmp_path / "nothing.ply")
    ok, err = start(p)
    assert not ok and "Missing high poly mesh" in err


@pytest.mark.parametrize("field,value,text", [("mapping", 7, "mapping method"), ("mappingEdge", 0.0, "mappingEdge"),
                                              ("texWidth", 0, "texture size"), ("bvhTrisPerNode", 0, "bvhTrisPerNode"),
                                              ("loPolyMeshNormal", 9, "normal import mode")])
def test_bad_parameters_are_reported(lowpoly, field, value, text):
    p = FornosParameters()
    p.shared.loPolyMeshPath = lowpoly
    if field == "mappingEdge":
        p.shared.mapping = MeshMappingMethod.Hybrid
    setattr(p.shared, field, value)
    ok, err = start(p)
    assert not ok and text in err


def test_idle_runner():
    r = FornosRunner()
    assert not r.pending() and r.currentTask() is None
    r.run()
    r.run_all()


def test_prepare_file_inputs_and_modes(lowpoly):
    # prepare() with a mesh file, hi = lo, per-face normals and each mapping
    cfg = {"low": lowpoly, "high": None, "size": 64, "lowNormals": "face", "maps": {"ao": {}}}
    for mapping, dirs_kind in (("smooth", "copy"), ("lowpoly", None), ("hybrid", "copy")):
        c = dict(cfg, mapping=mapping)
        prep = configs.prepare(c, with_scene=False)
        assert prep["high"] is prep["low"]
        lo = prep["low"].arrays()
        # per-face normals: one per triangle
        assert len(lo["normals"]) == len(lo["triangles"])
        if dirs_kind is None:
            assert prep["dirs"] is None
        else:
            assert prep["dirs"] is not prep["low"]
        assert prep["edge"] == (0.05 if mapping == "hybrid" else 0.0)
    # per-vertex normals: the low-poly mesh itself gives the directions
    prep = configs.prepare(dict(cfg, lowNormals="vertex"), with_scene=False)
    assert prep["dirs"] is prep["low"]
    with pytest.raises(ValueError):
        configs.prepare(dict(cfg, mapping="hybrid", mappingEdge=0.0), with_scene=False)
    with pytest.raises(ValueError):
        configs.prepare(dict(cfg, lowNormals="smooth"), with_scene=False)
