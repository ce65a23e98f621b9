"""Host-side input producers of the product against the oracle (CPU only).

Normals, tangent frames, the MapUV raster + CompressedMapUV, the sample table
and the reference-layout BVH + flatten must be bit-identical to the oracle's
restatement of the reference functions.
"""
import ctypes as C

import numpy as np
import pytest

import fornos_amd as fa
import oracle as orc
from fornos_amd._abi import check, lib, ptr


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def eq_nan(a, b):
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    n = np.isnan(a) & np.isnan(b)
    return np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(bits(a)[~n], bits(b)[~n])


@pytest.fixture(scope="module", params=["sphere512", "heightfield_small", "knot_small"])
def prep(request):
    if request.param == "sphere512":
        cfg = fa.CONFIGS["sphere512"]
    elif request.param == "heightfield_small":
        cfg = fa.configs.scaled("heightfield2048", size=384, high_quads=96)
    else:
        cfg = fa.configs.scaled("knot4096", size=256, high_quads=64)
    return fa.prepare(cfg, with_scene=False)


def test_vertex_normals(prep):
    h = prep["high"].arrays()
    n, vn = orc.vertex_normals(h["positions"], h["vertices"], h["triangles"])
    assert np.array_equal(bits(n), bits(h["normals"]))
    assert np.array_equal(vn, h["vertices"][:, 2])


def test_aggressive_normals(prep):
    lo, d = prep["low"].arrays(), prep["dirs"].arrays()
    n, vn = orc.vertex_normals_aggressive(lo["positions"], lo["vertices"], lo["triangles"])
    assert np.array_equal(bits(n), bits(d["normals"]))
    assert np.array_equal(vn, d["vertices"][:, 2])


def test_tangent_space(prep):
    lo = prep["low"].arrays()
    if "tangents" not in lo:
        pytest.skip("workload bakes no tangent-space map")
    t, b = orc.tangent_space(lo["positions"], lo["normals"], lo["texcoords"], lo["vertices"], lo["triangles"])
    assert eq_nan(t, lo["tangents"]) and eq_nan(b, lo["bitangents"])


def test_mapuv_and_compress(prep):
    lo, d, tex = prep["low"].arrays(), prep["dirs"].arrays(), prep["texels"]
    W = tex.width
    m = orc.mapuv(lo, W, W, d["normals"], d["vertices"][:, 2], lo.get("tangents"), lo.get("bitangents"))
    pm = tex.mapuv()
    for k in m:
        assert eq_nan(m[k], pm[k]), k
    idx = orc.compress(m["directions"])
    ta = tex.arrays()
    assert np.array_equal(idx, ta["indices"])
    assert eq_nan(m["positions"][idx], ta["positions"])
    assert eq_nan(m["directions"][idx], ta["directions"])


@pytest.mark.parametrize("edge", [0.05, 0.3])
def test_hybrid_mapuv_matches_oracle(prep, edge):
    # MapUV::fromMeshes_Hybrid: the Hybrid mapping method
    lo, d = prep["low"].arrays(), prep["dirs"].arrays()
    W = prep["texels"].width // 2
    m = orc.mapuv_edge(lo, W, W, edge, d["normals"], d["vertices"][:, 2], lo.get("tangents"), lo.get("bitangents"))
    tex = fa.CompressedMapUV(prep["low"], prep["dirs"], W, W, edge=edge)
    pm = tex.mapuv()
    for k in m:
        assert eq_nan(m[k], pm[k]), k


def test_hybrid_oracle_limits(prep):
    # pins of the oracle's blend d = normalize(dsmooth (1 - t) + n t) on a
    # faceted low-poly mesh (n = face normal, dsmooth = aggressive vertex
    # normal): a vanishing edge width gives the face normal everywhere
    # (t = 1), a huge one the Smooth mapping's direction (t ~ 0); positions
    # and normals do not depend on the width
    lo, d = dict(prep["low"].arrays()), prep["dirs"].arrays()
    fn, fv = orc.face_normals(lo["positions"], lo["vertices"], lo["triangles"])
    lo["normals"] = fn
    lo["vertices"] = lo["vertices"].copy()
    lo["vertices"][:, 2] = fv
    W = 64
    args = (d["normals"], d["vertices"][:, 2])
    sharp = orc.mapuv_edge(lo, W, W, 1e-30, *args)
    soft = orc.mapuv_edge(lo, W, W, 1e30, *args)
    smooth = orc.mapuv(lo, W, W, *args)
    hit = np.abs(sharp["normals"]).sum(1) > 0
    assert hit.sum() > W * W // 4
    # texels exactly on an edge have distance 0, so t = 0 there at any width
    near_n = np.abs(sharp["directions"] - sharp["normals"]).max(1) <= 1e-6
    near_s = np.abs(sharp["directions"] - smooth["directions"]).max(1) <= 1e-6
    assert np.all(near_n[hit] | near_s[hit]) and near_n[hit].mean() > 0.98
    assert np.allclose(soft["directions"], smooth["directions"], atol=1e-6)
    assert np.array_equal(sharp["positions"], soft["positions"]) and np.array_equal(sharp["normals"], soft["normals"])
    # the two limits differ on this mesh, so neither check above is vacuous
    assert np.abs(soft["directions"][hit] - sharp["directions"][hit]).max() > 1e-2


def test_reference_bvh_matches_oracle(prep):
    hi = prep["high"]
    n = C.c_uint32()
    check(lib().fornos_bvh_reference_export(hi.handle, 8, 8192, C.byref(n), None, None))
    nodes = np.empty((n.value, 9), np.uint32)
    ids = np.empty(hi.n_triangles, np.uint32)
    check(lib().fornos_bvh_reference_export(hi.handle, 8, 8192, C.byref(n), ptr(nodes), ptr(ids)))
    on, sp, sn, st = orc.bvh_build(hi.arrays())
    assert np.array_equal(on, nodes)
    assert np.array_equal(st, ids)


@pytest.mark.parametrize("S", [1, 32, 256, 1000])
def test_sample_table(S):
    tab, rnd = fa.sample_table(S, 64, seed=99)
    assert np.array_equal(bits(orc.samples(S, rnd)), bits(tab))


def test_sample_table_seeded_and_deterministic():
    a, ra = fa.sample_table(32, 16, seed=5)
    b, rb = fa.sample_table(32, 16, seed=5)
    c, rc = fa.sample_table(32, 16, seed=6)
    assert np.array_equal(a, b) and np.array_equal(ra, rb)
    assert not np.array_equal(ra, rc)


def test_generator_sizes():
    assert fa.Mesh.generate(0, 0, 1.0, 32, 16).n_triangles == 960
    assert fa.Mesh.generate(1, 7, 0.02, 6, 0).n_triangles == 81920
    assert fa.Mesh.generate(3, 1, 0.05, 64, 0).n_triangles == 8192
    assert fa.Mesh.generate(2, 1, 0.05, 32, 10).n_triangles == 2 * 32 * 32


def test_generators_deterministic():
    a = fa.Mesh.generate(2, 3, 0.05, 48, 20).arrays()
    b = fa.Mesh.generate(2, 3, 0.05, 48, 20).arrays()
    c = fa.Mesh.generate(2, 4, 0.05, 48, 20).arrays()
    assert np.array_equal(a["positions"], b["positions"])
    assert not np.array_equal(a["positions"], c["positions"])


def _wrap_mesh():
    # texture coordinates past u = 1: the Hybrid raster does not clamp its
    # box to the map, so columns x >= W land on following rows; a second,
    # overlapping triangle overwrites part of the first; tilted normals make
    # the blend visible
    from meshes import soup_mesh
    e = 1.0 / 128
    tris = [[(0, 0, 0), (1, 0, 0.2), (1, 1, 0)], [(0, 0, 1), (1, 0, 1), (0, 1, 1.3)]]
    nrm = [[(0, 0, 1), (0.3, 0, 0.95), (0, 0.3, 0.95)], [(0.1, 0, 1), (0, 0, 1), (0, 0.2, 1)]]
    uvs = [[(e, e), (1.5, e), (1.5, 1 - e)], [(e, e), (1.25, e), (e, 1.0)]]
    m = soup_mesh(tris, normals=nrm, uvs=uvs)
    return m, fa.Mesh.from_arrays(m["positions"], m["vertices"], m["triangles"], m["normals"], m["texcoords"])


@pytest.mark.parametrize("edge", [0.02, 0.25])
def test_hybrid_uv_past_the_edge_matches_oracle(edge):
    m, lo = _wrap_mesh()
    W, H = 96, 64
    ref = orc.mapuv_edge(m, W, H, edge)
    pm = fa.CompressedMapUV(lo, None, W, H, edge=edge).mapuv()
    for k in ref:
        assert eq_nan(ref[k], pm[k]), k
    # columns past the edge did wrap onto the next rows: u in [1, 1.5] is
    # x in [0.66, 1] on the first triangle, found at the start of rows
    pos = ref["positions"].reshape(H, W, 3)
    assert np.any(pos[1:, 0, 0] > 0.6)


def test_hybrid_rejects_runaway_uv_boxes():
    from meshes import soup_mesh
    m = soup_mesh([[(0, 0, 0), (1, 0, 0), (1, 1, 0)]], uvs=[[(0, 0), (1e6, 0), (1e6, 1)]])
    lo = fa.Mesh.from_arrays(m["positions"], m["vertices"], m["triangles"], m["normals"], m["texcoords"])
    with pytest.raises(Exception, match="too far outside"):
        fa.CompressedMapUV(lo, None, 2048, 2048, edge=0.05)
    # the Smooth raster clamps its box to the map and takes the same mesh
    assert fa.CompressedMapUV(lo, None, 64, 64).count >= 0
