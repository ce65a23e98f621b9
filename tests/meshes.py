"""Small hand-built meshes (the reference Mesh model as numpy arrays) for
closed-form and known-answer tests."""
import numpy as np


def soup_mesh(tris, normals=None, uvs=None):
    """Unshared vertices, one per corner, as the OBJ loader builds them.
    tris: (T,3,3) corner positions; normals: (T,3,3) or None (then +z);
    uvs: (T,3,2) or None."""
    tris = np.asarray(tris, np.float32).reshape(-1, 3, 3)
    T = len(tris)
    pos = tris.reshape(-1, 3).copy()
    nrm = (np.asarray(normals, np.float32).reshape(-1, 3) if normals is not None
           else np.tile(np.float32([0, 0, 1]), (3 * T, 1)))
    idx = np.arange(3 * T, dtype=np.uint32)
    uv = None
    tex_idx = np.full(3 * T, 0xFFFFFFFF, np.uint32)
    if uvs is not None:
        uv = np.asarray(uvs, np.float32).reshape(-1, 2).copy()
        tex_idx = idx.copy()
    verts = np.stack([idx, tex_idx, idx], axis=1).astype(np.uint32)
    tri_idx = idx.reshape(-1, 3).copy()
    m = {"positions": pos, "normals": nrm, "vertices": verts, "triangles": tri_idx}
    m["texcoords"] = uv if uv is not None else np.zeros((0, 2), np.float32)
    return m


def quad(z, size=1.0, flip=False, x0=None, y0=None):
    """Two triangles covering [x0, x0+size] x [y0, y0+size] at height z,
    counter-clockwise seen from +z (normal +z) unless flipped."""
    x0 = -size / 2 if x0 is None else x0
    y0 = -size / 2 if y0 is None else y0
    a, b, c, d = (x0, y0, z), (x0 + size, y0, z), (x0 + size, y0 + size, z), (x0, y0 + size, z)
    t = [[a, b, c], [a, c, d]]
    if flip:
        t = [[a, c, b], [a, d, c]]
    return t


def wall_x(x, size=1.0):
    """Vertical wall in the plane x = const covering y, z in [-size/2, size/2]."""
    h = size / 2
    a, b, c, d = (x, -h, -h), (x, h, -h), (x, h, h), (x, -h, h)
    return [[a, b, c], [a, c, d]]


def octahedron(shared=True):
    """Unit octahedron with outward winding; positions shared (indexed) or
    duplicated per corner."""
    P = np.float32([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]])
    F = [(0, 2, 4), (2, 1, 4), (1, 3, 4), (3, 0, 4), (2, 0, 5), (1, 2, 5), (3, 1, 5), (0, 3, 5)]
    if shared:
        verts = []
        tris = []
        for f in F:
            base = len(verts)
            for p in f:
                verts.append((p, 0xFFFFFFFF, p))
            tris.append((base, base + 1, base + 2))
        return {"positions": P.copy(), "normals": P.copy(), "vertices": np.array(verts, np.uint32),
                "triangles": np.array(tris, np.uint32), "texcoords": np.zeros((0, 2), np.float32)}
    return soup_mesh(np.array([[P[i] for i in f] for f in F], np.float32))


def grid(n, z=0.0):
    """n x n quads over [0,1]^2 at height z, shared positions, planar UVs."""
    xs = np.linspace(0, 1, n + 1, dtype=np.float32)
    X, Y = np.meshgrid(xs, xs)
    pos = np.stack([X.ravel(), Y.ravel(), np.full(X.size, z, np.float32)], 1).astype(np.float32)
    uv = pos[:, :2].copy()
    verts, tris = [], []
    for j in range(n):
        for i in range(n):
            a = j * (n + 1) + i
            b, c, d = a + 1, a + n + 2, a + n + 1
            for tri in ((a, b, c), (a, c, d)):
                base = len(verts)
                for p in tri:
                    verts.append((p, p, p))
                tris.append((base, base + 1, base + 2))
    nrm = np.tile(np.float32([0, 0, 1]), (len(pos), 1))
    return {"positions": pos, "normals": nrm, "texcoords": uv, "vertices": np.array(verts, np.uint32),
            "triangles": np.array(tris, np.uint32)}


def icosphere(radius, levels, inward=False):
    """Icosphere with shared positions and radial normals."""
    t = (1 + 5 ** 0.5) / 2
    P = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    P = [np.array(p) / np.linalg.norm(p) for p in P]
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
         (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(levels):
        mid = {}

        def m(a, b):
            k = (min(a, b), max(a, b))
            if k not in mid:
                v = P[a] + P[b]
                P.append(v / np.linalg.norm(v))
                mid[k] = len(P) - 1
            return mid[k]
        F = [g for a, b, c in F for g in ((a, m(a, b), m(c, a)), (b, m(b, c), m(a, b)),
                                           (c, m(c, a), m(b, c)), (m(a, b), m(b, c), m(c, a)))]
    pos = (np.array(P) * radius).astype(np.float32)
    nrm = np.array(P, np.float32) * (-1 if inward else 1)
    verts, tris = [], []
    for f in F:
        f = f[::-1] if inward else f
        base = len(verts)
        for p in f:
            verts.append((p, 0xFFFFFFFF, p))
        tris.append((base, base + 1, base + 2))
    return {"positions": pos, "normals": nrm.astype(np.float32), "vertices": np.array(verts, np.uint32),
            "triangles": np.array(tris, np.uint32), "texcoords": np.zeros((0, 2), np.float32)}


def merge(*meshes):
    """Concatenation of meshes (indices offset)."""
    out = {"positions": [], "normals": [], "texcoords": [], "vertices": [], "triangles": []}
    po = no = to = vo = 0
    for m in meshes:
        v = m["vertices"].astype(np.int64).copy()
        v[:, 0] += po
        v[:, 2] += no
        v[:, 1] = np.where(v[:, 1] == 0xFFFFFFFF, 0xFFFFFFFF, v[:, 1] + to)
        out["positions"].append(m["positions"])
        out["normals"].append(m["normals"])
        out["texcoords"].append(m["texcoords"])
        out["vertices"].append(v.astype(np.uint32))
        out["triangles"].append(m["triangles"] + vo)
        po += len(m["positions"]); no += len(m["normals"]); to += len(m["texcoords"]); vo += len(m["vertices"])
    return {k: np.concatenate(v).astype(v[0].dtype) for k, v in out.items()}


def write_obj(path, a):
    """Writes mesh arrays (Mesh.arrays() layout) as an OBJ file whose floats
    read back bit for bit (9 significant digits); faces v/vt/vn, 1-based."""
    lines = ["v %.9g %.9g %.9g" % tuple(p) for p in a["positions"]]
    lines += ["vt %.9g %.9g" % tuple(t) for t in a["texcoords"]]
    lines += ["vn %.9g %.9g %.9g" % tuple(n) for n in a["normals"]]
    V = a["vertices"]
    has_uv = len(a["texcoords"]) > 0
    for t in a["triangles"]:
        if has_uv:
            lines.append("f " + " ".join("%d/%d/%d" % (V[v, 0] + 1, V[v, 1] + 1, V[v, 2] + 1) for v in t))
        else:
            lines.append("f " + " ".join("%d//%d" % (V[v, 0] + 1, V[v, 2] + 1) for v in t))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
