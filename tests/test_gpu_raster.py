"""GPU texel set (MapUV raster + CompressedMapUV on the device) against the
host restatement, which tests/test_host.py holds bit-identical to the oracle.
"""
import numpy as np
import pytest

import fornos_amd as fa
import oracle as orc
from meshes import soup_mesh

pytestmark = pytest.mark.gpu


def eq_nan(a, b):
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    n = np.isnan(a) & np.isnan(b)
    return np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(a.view(np.uint32)[~n], b.view(np.uint32)[~n])


@pytest.mark.parametrize("name,size,quads", [("sphere512", None, None), ("heightfield2048", 512, 64),
                                              ("knot4096", 384, 48)])
def test_device_texel_set_equals_host(gpu, name, size, quads):
    cfg = fa.configs.scaled(name, size=size, high_quads=quads)
    prep = fa.prepare(cfg)
    host = prep["texels"].arrays()
    mm = fa.MeshMapping().init_from_mesh(prep["low"], prep["dirs"], cfg["size"], cfg["size"], prep["scene"],
                                         cfg["cull"])
    dev = mm.read_texels()
    assert mm.count == len(host["indices"])
    assert np.array_equal(dev["indices"], host["indices"])
    for k in ("positions", "directions", "normals", "tangents", "bitangents"):
        assert eq_nan(dev[k], host[k]), k
    # and the mapping run from it equals the one from the host texel set
    ref = fa.MeshMapping().init(prep["texels"], prep["scene"], cfg["cull"])
    ref.run_all()
    mm.run_all()
    c0, t0 = ref.read()
    c1, t1 = mm.read()
    assert np.array_equal(t0, t1) and np.array_equal(c0.view(np.uint32), c1.view(np.uint32))


def test_device_raster_overwrite_order_and_bound_column(gpu):
    # Overlapping triangles (the later one wins) and UVs reaching u = 1: the
    # bound column x == W passes the -0.001 tolerance at this size and writes
    # onto the next row's first texel (Src/compute.cpp:136-152).
    W = H = 2048
    e = 1.0 / 4096
    tris = [[(0, 0, 0), (1, 0, 0), (1, 1, 0)], [(0, 0, 1), (1, 0, 1), (1, 1, 1)]]
    uvs = [[(e, e), (1 - e, e), (1 - e, 1 - e)], [(e, e), (1.0, e), (1.0, 1 - e)]]
    m = soup_mesh(tris, uvs=uvs)
    mp = orc.mapuv(m, W, H)
    idx = orc.compress(mp["directions"])
    # above its first rows, column 0 is written only through triangle 1's
    # bound column (its x lies past the quad's edge)
    col0 = mp["positions"].reshape(H, W, 3)[:, 0]
    assert np.all(col0[:, 2] == 1) and np.all(col0[8:, 0] > 1)
    lo = fa.Mesh.from_arrays(m["positions"], m["vertices"], m["triangles"], m["normals"], m["texcoords"])
    hi = fa.Mesh.generate(fa._abi.GEN_HEIGHTFIELD, 1, 0.05, 8, 0)
    scene = fa.Scene(hi)
    mm = fa.MeshMapping().init_from_mesh(lo, None, W, H, scene, True)
    dev = mm.read_texels()
    assert np.array_equal(dev["indices"], idx)
    assert eq_nan(dev["positions"], mp["positions"][idx])
    assert eq_nan(dev["tangents"], mp["tangents"][idx])


def _faceted(mesh):
    """The mesh with one vertex per triangle corner, as the OBJ loader builds
    it, and per-face normals (Mesh::computeFaceNormals)."""
    a = mesh.arrays()
    verts = a["vertices"][a["triangles"].reshape(-1)]
    tris = np.arange(len(verts), dtype=np.uint32).reshape(-1, 3)
    f = fa.Mesh.from_arrays(a["positions"], verts, tris, a["normals"], a["texcoords"])
    f.computeFaceNormals()
    return f


@pytest.mark.parametrize("faceted", [False, True])
@pytest.mark.parametrize("edge", [0.05, 0.3])
def test_device_hybrid_texel_set_equals_host(gpu, edge, faceted):
    # MapUV::fromMeshes_Hybrid on the device against the host raster, with a
    # tangent frame (read by normal index in this method) and, faceted, with
    # face normals so the blend between the smooth and the surface direction
    # matters everywhere
    cfg = fa.configs.scaled("heightfield2048", size=384, high_quads=48)
    prep = fa.prepare(cfg)
    lo = _faceted(prep["low"]) if faceted else prep["low"]
    if faceted:
        lo.computeTangentSpace()
    dirs = lo.copy()
    dirs.computeVertexNormalsAggressive()
    W = cfg["size"]
    host = fa.CompressedMapUV(lo, dirs, W, W, edge=edge).arrays()
    mm = fa.MeshMapping().init_from_mesh(lo, dirs, W, W, prep["scene"], cfg["cull"], edge=edge)
    dev = mm.read_texels()
    assert np.array_equal(dev["indices"], host["indices"])
    for k in ("positions", "directions", "normals", "tangents", "bitangents"):
        assert eq_nan(dev[k], host[k]), k


def test_device_hybrid_wraps_like_the_oracle(gpu):
    # texture coordinates past u = 1 (the Hybrid box is not clamped to the
    # map, so columns x >= W land on following rows) and an overlapping
    # second triangle
    from meshes import soup_mesh
    e = 1.0 / 128
    tris = [[(0, 0, 0), (1, 0, 0.2), (1, 1, 0)], [(0, 0, 1), (1, 0, 1), (0, 1, 1.3)]]
    nrm = [[(0, 0, 1), (0.3, 0, 0.95), (0, 0.3, 0.95)], [(0.1, 0, 1), (0, 0, 1), (0, 0.2, 1)]]
    uvs = [[(e, e), (1.5, e), (1.5, 1 - e)], [(e, e), (1.25, e), (e, 1.0)]]
    m = soup_mesh(tris, normals=nrm, uvs=uvs)
    W, H = 96, 64
    ref = orc.mapuv_edge(m, W, H, 0.05)
    idx = orc.compress(ref["directions"])
    lo = fa.Mesh.from_arrays(m["positions"], m["vertices"], m["triangles"], m["normals"], m["texcoords"])
    scene = fa.Scene(fa.Mesh.generate(fa._abi.GEN_HEIGHTFIELD, 1, 0.05, 8, 0))
    mm = fa.MeshMapping().init_from_mesh(lo, None, W, H, scene, True, edge=0.05)
    dev = mm.read_texels()
    assert np.array_equal(dev["indices"], idx)
    for k in ("positions", "directions", "normals"):
        assert eq_nan(dev[k], ref[k][idx]), k
