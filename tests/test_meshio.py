"""Mesh files (Mesh::loadFile, Src/mesh.cpp:235-617): the library's OBJ and
PLY readers against the oracle's restatement (oracle/meshio.py, numbers read
with the C library's strtod / strtol as the reference reads them), error
cases, face normals against the C oracle, and a generator mesh round trip."""
import struct

import numpy as np
import pytest

import fornos_amd as fa
import oracle as orc
from fornos_amd import _abi
from oracle import meshio as mo


def same(mesh, ref):
    a = mesh.arrays()
    for k in ("positions", "texcoords", "normals", "vertices", "triangles"):
        x = np.asarray(a[k]).reshape(ref[k].shape) if a[k].size else np.zeros(ref[k].shape, ref[k].dtype)
        assert np.array_equal(x.view(np.uint32) if x.dtype == np.float32 else x,
                              ref[k].view(np.uint32) if ref[k].dtype == np.float32 else ref[k]), k


OBJ_TEXT = """# a quad, a triangle and a pentagon
v 0 0 0
v 1 0 0 1.0
   v 1 1 0
v 0 1 0\r
v 0.5 2e0 -0.25
v 2 2 2
vt 0 0
vt 1 0 0
vt 1 1
vt 0 1
vn 0 0 1
vp 0.5
f 1/1/1 2/2/1 3/3/1 4/4/1
f 4/4/1 3/3/1 5/1/1
f 1/1/1 2/2/1 3/3/1 5/4/1 6/1/1\r
l 1 2
o object_name
"""


def test_obj_matches_oracle(tmp_path):
    p = tmp_path / "m.obj"
    p.write_text(OBJ_TEXT)
    ref = mo.load_obj(str(p))
    assert len(ref["triangles"]) == 2 + 1 + 3
    same(fa.Mesh.load(p), ref)


def test_obj_oracle_known_answer(tmp_path):
    # OBJ_TEXT by hand: one mesh vertex per face corner (0-based position,
    # texcoord, normal indices), polygons fanned from their first corner
    p = tmp_path / "m.obj"
    p.write_text(OBJ_TEXT)
    ref = mo.load_obj(str(p))
    assert np.array_equal(ref["positions"], np.float32([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0.5, 2, -0.25],
                                                        [2, 2, 2]]))
    assert np.array_equal(ref["texcoords"], np.float32([[0, 0], [1, 0], [1, 1], [0, 1]]))
    assert np.array_equal(ref["normals"], np.float32([[0, 0, 1]]))
    corners = [(0, 0), (1, 1), (2, 2), (3, 3), (3, 3), (2, 2), (4, 0), (0, 0), (1, 1), (2, 2), (4, 3), (5, 0)]
    assert np.array_equal(ref["vertices"], np.uint32([(v, t, 0) for v, t in corners]))
    assert np.array_equal(ref["triangles"], np.uint32([[0, 1, 2], [0, 2, 3], [4, 5, 6], [7, 8, 9], [7, 9, 10],
                                                       [7, 10, 11]]))


@pytest.mark.parametrize("text", ["v 1 2 3\nv 1 2 3\nv 1 2 3\nf 1//1 2//1 3//1\nvn 0 0 1\n",
                                  "v 1 2 3\nv 4 5 6\nv 7 8 9\nf 1 2 3\nf 3 2 1\n"])
def test_obj_face_forms(tmp_path, text):
    p = tmp_path / "m.obj"
    p.write_text(text)
    same(fa.Mesh.load(p), mo.load_obj(str(p)))


@pytest.mark.parametrize("text", ["v 1 2 3 x\n", "v 1 2\n", "vt 1\n", "v 1 1 1\nv 2 2 2\nv 3 3 3\nvt 0 0\nf 1/1 2 3\n",
                                  "v 1e999 0 0\n", "v 1 1 1\nf 1 2 3\n", "v 1 1 1\nv 2 2 2\nv 3 3 3\nf 1 2 -1\n",
                                  "v 1 1 1\nf 1 1 1 # c\n"])
def test_obj_errors(tmp_path, text):
    p = tmp_path / "m.obj"
    p.write_text(text)
    with pytest.raises(fa.FornosError):
        fa.Mesh.load(p)
    # the oracle rejects the same files, except index ranges the reference never checks
    if "f 1 2 3\n" not in text and "-1" not in text:
        with pytest.raises(mo.ParseError):
            mo.load_obj(str(p))


def write_ply(path, fmt, pos, faces, st=None, vn=None, fn=None, coord="float", cnt="uchar", idx="int"):
    props = [("x", coord), ("y", coord), ("z", coord)]
    if st is not None:
        props += [("s", "float"), ("t", "float")]
    if vn is not None:
        props += [("nx", "float"), ("ny", "float"), ("nz", "float")]
    head = ["ply", "format %s 1.0" % fmt, "comment test", "element vertex %d" % len(pos)]
    head += ["property %s %s" % (t, n) for n, t in props]
    head += ["element face %d" % len(faces), "property list %s %s vertex_indices" % (cnt, idx)]
    if fn is not None:
        head += ["property float nx", "property float ny", "property float nz"]
    head += ["end_header"]
    rows = []
    for i, p in enumerate(pos):
        r = list(p) + (list(st[i]) if st is not None else []) + (list(vn[i]) if vn is not None else [])
        rows.append(r)
    with open(path, "wb") as f:
        f.write(("\n".join(head) + "\n").encode())
        if fmt == "ascii":
            for r in rows:
                f.write((" ".join("%.9g" % v for v in r) + "\n").encode())
            for i, fc in enumerate(faces):
                extra = list(fn[i]) if fn is not None else []
                f.write((" ".join([str(len(fc))] + [str(v) for v in fc] + ["%.9g" % v for v in extra]) + "\n").encode())
        else:
            bo = "<" if fmt == "binary_little_endian" else ">"
            code = {"float": "f", "double": "d", "uchar": "B", "int": "i", "uint": "I", "ushort": "H"}
            for r in rows:
                for (n, t), v in zip(props, r):
                    f.write(struct.pack(bo + code[t], v))
            for i, fc in enumerate(faces):
                f.write(struct.pack(bo + code[cnt], len(fc)))
                f.write(struct.pack(bo + code[idx] * len(fc), *fc))
                if fn is not None:
                    f.write(struct.pack(bo + "fff", *fn[i]))


@pytest.mark.parametrize("fmt,coord,idx,normals", [("ascii", "float", "int", "vertex"),
                                                   ("binary_little_endian", "double", "uint", "face"),
                                                   ("binary_big_endian", "float", "ushort", "none"),
                                                   ("binary_little_endian", "float", "int", "vertex")])
def test_ply_matches_oracle(tmp_path, fmt, coord, idx, normals):
    rng = np.random.default_rng(4)
    pos = rng.standard_normal((9, 3)).astype(np.float32)
    faces = [[0, 1, 2], [2, 3, 4, 5], [5, 6, 7], [8, 7, 6, 5, 4]]
    st = rng.random((9, 2)).astype(np.float32)
    vn = rng.standard_normal((9, 3)).astype(np.float32) if normals == "vertex" else None
    fn = rng.standard_normal((4, 3)).astype(np.float32) if normals == "face" else None
    p = tmp_path / "m.ply"
    write_ply(p, fmt, pos, faces, st=st, vn=vn, fn=fn, coord=coord, idx=idx)
    ref = mo.load_ply(str(p))
    # known answer: the arrays written, one mesh vertex per triangle corner,
    # polygons fanned from their first corner; a face normal is indexed by
    # its face
    corners = [0, 1, 2, 2, 3, 4, 2, 4, 5, 5, 6, 7, 8, 7, 6, 8, 6, 5, 8, 5, 4]
    face_of = [0, 0, 0, 1, 1, 1, 1, 1, 1, 2, 2, 2, 3, 3, 3, 3, 3, 3, 3, 3, 3]
    none = 0xFFFFFFFF
    assert np.array_equal(ref["positions"], pos)
    assert np.array_equal(ref["texcoords"], st)
    assert np.array_equal(ref["normals"], vn if vn is not None else fn if fn is not None else np.zeros((0, 3)))
    nidx = corners if vn is not None else face_of if fn is not None else [none] * len(corners)
    assert np.array_equal(ref["vertices"], np.uint32([(v, v, k) for v, k in zip(corners, nidx)]))
    assert np.array_equal(ref["triangles"], np.arange(len(corners), dtype=np.uint32).reshape(-1, 3))
    same(fa.Mesh.load(p), ref)


def test_face_normals_oracle_closed_forms():
    # Mesh::computeFaceNormals (Src/mesh.cpp:627-646): normalize(cross(p1 - p0, p2 - p0))
    # per triangle, and each corner's normal index set to its triangle, the
    # later triangle winning on a shared vertex
    from meshes import grid
    # (normalize is v * (1/sqrt(v.v)), Src/math.h, so a unit result can be 1 ulp short)
    g = grid(3)
    n, vn = orc.face_normals(g["positions"], g["vertices"], g["triangles"])
    assert np.all(n[:, :2] == 0) and np.all(np.abs(n[:, 2] - 1) <= 6e-8)
    assert np.array_equal(vn[g["triangles"]], np.repeat(np.arange(len(n), dtype=np.uint32)[:, None], 3, 1))
    n, _ = orc.face_normals(g["positions"], g["vertices"], g["triangles"][:, ::-1].copy())
    assert np.all(n[:, :2] == 0) and np.all(np.abs(n[:, 2] + 1) <= 6e-8)
    # a tilted triangle: cross((1,0,0), (0,1,1)) = (0,-1,1)
    pos = np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 1], [5, 5, 5]])
    verts = np.uint32([[0, 0, 7], [1, 0, 7], [2, 0, 7], [3, 0, 9]])
    n, vn = orc.face_normals(pos, verts, np.uint32([[0, 1, 2]]))
    assert np.allclose(n[0], [0, -2 ** -0.5, 2 ** -0.5], atol=1e-7, rtol=0) and n[0, 0] == 0
    assert list(vn) == [0, 0, 0, 9]  # a vertex outside every triangle keeps its index
    # shared vertices 1 and 2: the second triangle's index wins
    pos = np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]])
    verts = np.uint32([[i, 0, 0] for i in range(4)])
    n, vn = orc.face_normals(pos, verts, np.uint32([[0, 1, 2], [2, 1, 3]]))
    assert list(vn) == [0, 1, 1, 1]
    assert np.array_equal(n, np.float32([[0, 0, 1], [0, 0, 1]]))


def test_face_normals_match_oracle():
    m = fa.Mesh.generate(_abi.GEN_ICOSPHERE_WAVES, 3, 0.02, 3, 0)
    a = m.arrays()
    m.computeFaceNormals()
    b = m.arrays()
    n, vn = orc.face_normals(a["positions"], a["vertices"], a["triangles"])
    assert np.array_equal(b["normals"].view(np.uint32), n.view(np.uint32))
    assert np.array_equal(b["vertices"][:, 2], vn)


def test_generator_mesh_round_trips_through_obj(tmp_path):
    m = fa.Mesh.generate(_abi.GEN_TORUS_KNOT, 0, 0.0, 32, 8)
    from meshes import write_obj
    a = m.arrays()
    p = tmp_path / "knot.obj"
    write_obj(p, a)
    b = fa.Mesh.load(p).arrays()
    # per-triangle corners carry the same positions, texcoords and normals
    def corners(x):
        v = x["vertices"][x["triangles"].reshape(-1)]
        return x["positions"][v[:, 0]], x["texcoords"][v[:, 1]], x["normals"][v[:, 2]]
    for u, w in zip(corners(a), corners(b)):
        assert np.array_equal(u.view(np.uint32), w.view(np.uint32))


def test_unknown_extension_is_an_error(tmp_path):
    p = tmp_path / "m.stl"
    p.write_text("solid")
    with pytest.raises(fa.FornosError):
        fa.Mesh.load(p)


@pytest.mark.parametrize("count", ["4294967295", "1e300", "nan", "2.5", "-1"])
def test_ply_corrupt_list_count_is_an_error(tmp_path, count):
    # a face list whose count is huge, not finite, fractional or negative:
    # rejected with a status before any allocation sized by it
    p = tmp_path / "m.ply"
    p.write_text("ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nproperty float y\n"
                 "property float z\nelement face 1\nproperty list uchar int vertex_indices\nend_header\n"
                 "0 0 0\n1 0 0\n0 1 0\n%s 0 1 2\n" % count)
    with pytest.raises(fa.FornosError) as e:
        fa.Mesh.load(p)
    assert e.value.status == _abi.INVALID_ARG


def test_ply_binary_oversized_count_is_an_error(tmp_path):
    # binary little-endian: a uint count of 0xFFFFFFFF in a short file
    head = ("ply\nformat binary_little_endian 1.0\nelement vertex 3\nproperty float x\nproperty float y\n"
            "property float z\nelement face 1\nproperty list uint int vertex_indices\nend_header\n").encode()
    body = struct.pack("<9f", 0, 0, 0, 1, 0, 0, 0, 1, 0) + struct.pack("<I", 0xFFFFFFFF) + struct.pack("<3i", 0, 1, 2)
    p = tmp_path / "m.ply"
    p.write_bytes(head + body)
    with pytest.raises(fa.FornosError):
        fa.Mesh.load(p)
    p.write_bytes(head + body[:40])  # truncated inside the face element
    with pytest.raises(fa.FornosError):
        fa.Mesh.load(p)
