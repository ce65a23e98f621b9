"""CPU oracle of the Fornos baking hot path -- TEST INFRASTRUCTURE ONLY.

A plain-C restatement of the reference shaders (Shaders/*.comp), of its
binary BVH build and flatten (Src/bvh.cpp, Src/meshmapping.cpp:62-125), of the
texel raster (Src/compute.cpp:75-193,372-416), of the mesh normal/tangent
functions (Src/mesh.cpp:650-768) and of the sample table (Src/math.h:324-382),
with numpy wrappers.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it, as the checker; the product never does.

How each part is pinned (golden fixtures, closed forms, brute force,
invariants) is listed in tests/test_oracle.py.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liborc.so")
MISS = 0xFFFFFFFF
FLT_MAX = np.float32(3.4028234663852886e38)

_lib = None
P = C.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _load(LIB_PATH)
    return _lib


_mutants = {}


def mutant(name):
    """The oracle library with one planted mistake ("bn_maxd": bent normals
    count only hits below maxD; "thick_clamp": thickness clamps hits to maxD),
    for tests that show the pins reject it."""
    if name not in _mutants:
        path = os.path.join(HERE, "_build", "liborc_mut_%s.so" % name)
        if not os.path.exists(path):
            build()
        _mutants[name] = _load(path)
    return _mutants[name]


def _load(path):
    L = C.CDLL(path)
    u32, i32, f32 = C.c_uint32, C.c_int, C.c_float
    sigs = {
        "orc_vertex_normals": (None, [P, u32, P, u32, P, u32, P, P]),
        "orc_vertex_normals_aggressive": (u32, [P, P, u32, P, u32, P, P]),
        "orc_face_normals": (None, [P, P, u32, P, u32, P, P]),
        "orc_tangent_space": (None, [P, P, P, P, u32, P, u32, P, P]),
        "orc_mapuv": (i32, [P, P, P, P, P, u32, P, P, P, P, u32, u32, P, P, P, P, P]),
        "orc_compress": (u32, [P, u32, P]),
        "orc_mapuv_edge": (i32, [P, P, P, P, P, u32, P, P, P, P, u32, u32, f32, P, P, P, P, P]),
        "orc_samples": (None, [u32, u32, P, P]),
        "orc_bvh_build": (u32, [P, u32, P, P, P, u32, u32, u32, P, P, P, P, P]),
        "orc_meshmapping": (None, [P, P, u32, u32, P, P, u32, i32, P, P, i32]),
        "orc_hemisphere": (None, [i32, u32, u32, f32, f32, P, P, P, u32, u32, P, P, P, u32, P, P, P, i32]),
        "orc_tangentspace": (None, [P, P, P, u32, u32, P]),
        "orc_height": (None, [P, u32, P]),
        "orc_texel_attr": (None, [i32, P, P, u32, P, P]),
        "orc_map_bruteforce": (None, [P, P, u32, P, u32, i32, P, P]),
        "orc_hemi_bruteforce": (f32, [P, P, f32, P, u32]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def _c(a, dt):
    return None if a is None else np.ascontiguousarray(a, dtype=dt)


def vertex_normals(pos, verts, tris):
    pos, verts, tris = _c(pos, np.float32), _c(verts, np.uint32), _c(tris, np.uint32)
    out = np.empty_like(pos)
    vn = np.empty(len(verts), np.uint32)
    lib().orc_vertex_normals(_p(pos), len(pos), _p(verts), len(verts), _p(tris), len(tris), _p(out), _p(vn))
    return out, vn


def face_normals(pos, verts, tris):
    """(normals [ntri, 3], per-vertex normal index) of Mesh::computeFaceNormals."""
    pos, verts, tris = _c(pos, np.float32), _c(verts, np.uint32), _c(tris, np.uint32)
    nt = len(tris)
    out = np.empty((nt, 3), np.float32)
    vn = np.empty(len(verts), np.uint32)
    lib().orc_face_normals(_p(pos), _p(verts), len(verts), _p(tris), nt, _p(out), _p(vn))
    return out, vn


def vertex_normals_aggressive(pos, verts, tris):
    pos, verts, tris = _c(pos, np.float32), _c(verts, np.uint32), _c(tris, np.uint32)
    out = np.empty((3 * len(tris) + 1, 3), np.float32)
    vn = verts[:, 2].copy()
    n = lib().orc_vertex_normals_aggressive(_p(pos), _p(verts), len(verts), _p(tris), len(tris), _p(out), _p(vn))
    return out[:n].copy(), vn


def tangent_space(pos, nrm, uv, verts, tris):
    a = [_c(pos, np.float32), _c(nrm, np.float32), _c(uv, np.float32), _c(verts, np.uint32), _c(tris, np.uint32)]
    t = np.empty((len(a[3]), 3), np.float32)
    b = np.empty((len(a[3]), 3), np.float32)
    lib().orc_tangent_space(_p(a[0]), _p(a[1]), _p(a[2]), _p(a[3]), len(a[3]), _p(a[4]), len(a[4]), _p(t), _p(b))
    return t, b


def mapuv(mesh, width, height, dir_normals=None, dir_vert_nrm=None, tangents=None, bitangents=None):
    """MapUV arrays (W*H, 3) x 5, or None when texcoords/normals are missing."""
    pos, nrm, uv = _c(mesh["positions"], np.float32), _c(mesh["normals"], np.float32), _c(mesh["texcoords"], np.float32)
    verts, tris = _c(mesh["vertices"], np.uint32), _c(mesh["triangles"], np.uint32)
    dn, dv = _c(dir_normals, np.float32), _c(dir_vert_nrm, np.uint32)
    tg, bt = _c(tangents, np.float32), _c(bitangents, np.float32)
    npx = width * height
    outs = [np.empty((npx, 3), np.float32) for _ in range(5)]
    ok = lib().orc_mapuv(_p(pos), _p(nrm), _p(uv), _p(verts), _p(tris), len(tris), _p(dn), _p(dv),
                         _p(tg), _p(bt), width, height, *[_p(o) for o in outs])
    if not ok:
        return None
    return dict(zip(["positions", "directions", "normals", "tangents", "bitangents"], outs))


def mapuv_edge(mesh, width, height, edge, dir_normals=None, dir_vert_nrm=None, tangents=None, bitangents=None):
    """MapUV::fromMeshes_Hybrid arrays (W*H, 3) x 5, or None."""
    pos, nrm, uv = _c(mesh["positions"], np.float32), _c(mesh["normals"], np.float32), _c(mesh["texcoords"], np.float32)
    verts, tris = _c(mesh["vertices"], np.uint32), _c(mesh["triangles"], np.uint32)
    dn, dv = _c(dir_normals, np.float32), _c(dir_vert_nrm, np.uint32)
    tg, bt = _c(tangents, np.float32), _c(bitangents, np.float32)
    npx = width * height
    outs = [np.empty((npx, 3), np.float32) for _ in range(5)]
    ok = lib().orc_mapuv_edge(_p(pos), _p(nrm), _p(uv), _p(verts), _p(tris), len(tris), _p(dn), _p(dv),
                              _p(tg), _p(bt), width, height, float(edge), *[_p(o) for o in outs])
    if not ok:
        return None
    return dict(zip(["positions", "directions", "normals", "tangents", "bitangents"], outs))


def compress(directions):
    d = _c(directions, np.float32)
    idx = np.empty(len(d), np.uint32)
    n = lib().orc_compress(_p(d), len(d), _p(idx))
    return idx[:n].copy()


def samples(S, rand):
    r = _c(rand, np.uint32)
    P_ = len(r)
    out = np.empty((P_ * S, 3), np.float32)
    lib().orc_samples(S, P_, _p(r), _p(out))
    return out


def bvh_build(mesh, max_tris=8, max_depth=8192):
    """Reference BVH + flatten: (nodes[n,9] uint32, soup_pos[3t,4], soup_nrm[3t,4], soup_tri[t])."""
    pos, nrm = _c(mesh["positions"], np.float32), _c(mesh["normals"], np.float32)
    verts, tris = _c(mesh["vertices"], np.uint32), _c(mesh["triangles"], np.uint32)
    nt = len(tris)
    nodes = np.zeros((2 * nt + 1, 9), np.uint32)
    sp = np.zeros((3 * nt, 4), np.float32)
    sn = np.zeros((3 * nt, 4), np.float32)
    st = np.zeros(nt, np.uint32)
    nsoup = C.c_uint32()
    nn = lib().orc_bvh_build(_p(pos), len(pos), _p(nrm), _p(verts), _p(tris), nt, max_tris, max_depth,
                             _p(nodes), _p(sp), _p(sn), _p(st), C.byref(nsoup))
    assert nsoup.value == 3 * nt
    return nodes[:nn].copy(), sp, sn, st


def meshmapping(pix_p, pix_d, spos, nodes, cull, begin=0, end=None, nthreads=0):
    pp, pd = _c(pix_p, np.float32), _c(pix_d, np.float32)
    n = len(pp)
    end = n if end is None else end
    coords = np.zeros((n, 4), np.float32)
    tidx = np.full(n, MISS, np.uint32)
    nodes = _c(nodes, np.uint32)
    lib().orc_meshmapping(_p(pp), _p(pd), begin, end, _p(_c(spos, np.float32)), _p(nodes), len(nodes),
                          int(cull), _p(coords), _p(tidx), nthreads)
    return coords, tidx


KIND_AO, KIND_BN, KIND_THICK = 0, 1, 2


def hemisphere(kind, S, mind, maxd, table, coords, tidx, spos, snrm, nodes, begin=0, end=None,
               nthreads=0, stats=False, library=None):
    """Per-texel AO / bent normal (pre tangent space) / thickness over texels
    [begin, end).  library: another build of the oracle (a mutant)."""
    coords, tidx = _c(coords, np.float32), _c(tidx, np.uint32)
    table = _c(table, np.float32)
    end = len(tidx) if end is None else end
    m = end - begin
    out = np.zeros((m, 3) if kind == KIND_BN else (m,), np.float32)
    counts = np.zeros(m, np.uint32)
    st = np.zeros(2, np.uint64)
    nodes = _c(nodes, np.uint32)
    Pn = len(table) // S
    (library or lib()).orc_hemisphere(kind, S, Pn, float(mind), float(maxd), _p(table), _p(coords), _p(tidx), begin, end,
                         _p(_c(spos, np.float32)), _p(_c(snrm, np.float32)), _p(nodes), len(nodes),
                         _p(out), _p(counts), _p(st) if stats else None, nthreads)
    if stats:
        return out, counts, st
    return out, counts


def tangentspace(n, t, b, vals, begin=0):
    n, t, b = _c(n, np.float32), _c(t, np.float32), _c(b, np.float32)
    v = np.array(vals, dtype=np.float32, copy=True, order="C")
    lib().orc_tangentspace(_p(n), _p(t), _p(b), begin, begin + len(v), _p(v))
    return v


def height(coords):
    c = _c(coords, np.float32)
    out = np.empty(len(c), np.float32)
    lib().orc_height(_p(c), len(c), _p(out))
    return out


def texel_attr(kind, coords, tidx, soup):
    c, t, s = _c(coords, np.float32), _c(tidx, np.uint32), _c(soup, np.float32)
    out = np.empty((len(c), 3), np.float32)
    lib().orc_texel_attr(kind, _p(c), _p(t), len(c), _p(s), _p(out))
    return out


def map_bruteforce(pix_p, pix_d, spos, cull):
    pp, pd, sp = _c(pix_p, np.float32), _c(pix_d, np.float32), _c(spos, np.float32)
    coords = np.zeros((len(pp), 4), np.float32)
    tidx = np.full(len(pp), MISS, np.uint32)
    lib().orc_map_bruteforce(_p(pp), _p(pd), len(pp), _p(sp), len(sp), int(cull), _p(coords), _p(tidx))
    return coords, tidx


def hemi_bruteforce(o, d, mind, spos):
    o, d, sp = _c(o, np.float32), _c(d, np.float32), _c(spos, np.float32)
    return float(lib().orc_hemi_bruteforce(_p(o), _p(d), float(mind), _p(sp), len(sp)))
