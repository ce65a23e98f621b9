"""CPU oracle of Mesh::loadWavefrontObj and Mesh::loadPly (Src/mesh.cpp:235-604)
-- TEST INFRASTRUCTURE ONLY.

OBJ numbers are read with the C library's strtod / strtol (through ctypes),
as the reference reads them.  Returns dicts of numpy arrays in the library's
Mesh layout: positions, texcoords, normals, vertices [n, 3] (position,
texcoord, normal index; 0xFFFFFFFF = none), triangles [m, 3]."""
import ctypes as C
import struct

import numpy as np

NONE = 0xFFFFFFFF
_libc = C.CDLL(None, use_errno=True)
_libc.strtod.restype = C.c_double
_libc.strtod.argtypes = [C.c_char_p, C.POINTER(C.c_char_p)]
_libc.strtol.restype = C.c_long
_libc.strtol.argtypes = [C.c_char_p, C.POINTER(C.c_char_p), C.c_int]


class ParseError(ValueError):
    pass


def _num(buf, pos, integer, required):
    """(value, new position) of strtod / strtol at buf[pos:], failing like
    the reference's readFloat / readInt (nothing parsed when required, or
    errno set)."""
    cs = C.c_char_p(buf[pos:])
    end = C.c_char_p()
    C.set_errno(0)
    if integer:
        v = C.c_int(_libc.strtol(cs, C.byref(end), 10)).value  # static_cast<int>
    else:
        v = float(np.float32(_libc.strtod(cs, C.byref(end))))  # static_cast<float>
    err = C.get_errno()
    used = C.cast(end, C.c_void_p).value - C.cast(cs, C.c_void_p).value
    if required and used == 0:
        raise ParseError("number expected")
    if err != 0:
        raise ParseError("range error")
    return v, pos + used


def _eol(buf, pos):
    return buf[pos:].strip() == b""


def load_obj(path):
    pos, tex, nrm, corners, faces = [], [], [], [], []
    opts = None
    for raw in open(path, "rb").read().split(b"\n"):
        line = raw
        stripped = line.lstrip()
        if not stripped:
            continue
        b = len(line) - len(stripped)
        tok = stripped.split()[0]
        p = b + len(tok)
        if tok[:1] == b"#":
            continue
        if tok in (b"v", b"vt", b"vn"):
            n = {b"v": 3, b"vt": 2, b"vn": 3}[tok]
            vals = []
            for _ in range(n):
                v, p = _num(line, p, False, True)
                vals.append(v)
            if tok != b"vn":
                _, p = _num(line, p, False, False)
            if not _eol(line, p):
                raise ParseError("trailing text")
            {b"v": pos, b"vt": tex, b"vn": nrm}[tok].append(vals)
        elif tok == b"f":
            first = len(corners)
            while not _eol(line, p):
                vi, p = _num(line, p, True, True)
                ti = ni = 0
                if line[p:p + 1] == b"/":
                    ti, p = _num(line, p + 1, True, False)
                if line[p:p + 1] == b"/":
                    ni, p = _num(line, p + 1, True, True)
                o = (ti != 0, ni != 0)
                if opts is None:
                    opts = o
                elif o != opts:
                    raise ParseError("mixed face forms")
                corners.append((vi, ti, ni))
                while p < len(line) and line[p:p + 1].isspace():
                    p += 1
            faces.append((first, len(corners) - first))
    verts, tris = [], []
    for first, cnt in faces:
        v0 = prev = None
        for i in range(cnt):
            vi, ti, ni = corners[first + i]
            k = len(verts)
            verts.append((vi - 1, ti - 1 if ti else NONE, ni - 1 if ni else NONE))
            if v0 is None:
                v0 = k
            elif prev is None:
                prev = k
            else:
                tris.append((v0, prev, k))
                prev = k
    return {"positions": np.array(pos, np.float32).reshape(-1, 3), "texcoords": np.array(tex, np.float32).reshape(-1, 2),
            "normals": np.array(nrm, np.float32).reshape(-1, 3), "vertices": np.array(verts, np.uint32).reshape(-1, 3),
            "triangles": np.array(tris, np.uint32).reshape(-1, 3)}


_FMT = {"char": "b", "int8": "b", "uchar": "B", "uint8": "B", "short": "h", "int16": "h", "ushort": "H",
        "uint16": "H", "int": "i", "int32": "i", "uint": "I", "uint32": "I", "float": "f", "float32": "f",
        "double": "d", "float64": "d"}


def load_ply(path):
    """Vertices x, y, z (+ s, t, + vertex or face nx, ny, nz), faces fanned;
    one mesh vertex per triangle corner (the reference's loadPly)."""
    data = open(path, "rb").read()
    end = data.index(b"end_header") + len(b"end_header")
    end = data.index(b"\n", end) + 1
    header = data[:end].decode().replace("\r", "").split("\n")
    fmt, elems = None, []
    for l in header:
        w = l.split()
        if not w:
            continue
        if w[0] == "format":
            fmt = w[1]
        elif w[0] == "element":
            elems.append((w[1], int(w[2]), []))
        elif w[0] == "property":
            elems[-1][2].append((w[-1], w[2], w[3]) if w[1] == "list" else (w[2], None, w[1]))
    body = data[end:]
    toks = body.split() if fmt == "ascii" else None
    ti = 0
    off = 0
    bo = "<" if fmt == "binary_little_endian" else ">"
    rows = {}
    for name, count, props in elems:
        out = []
        for _ in range(count):
            row = {}
            for pname, ctype, ptype in props:
                if ctype is not None:
                    if toks is not None:
                        n = int(float(toks[ti])); ti += 1
                        row[pname] = [float(t) for t in toks[ti:ti + n]]; ti += n
                    else:
                        (n,) = struct.unpack_from(bo + _FMT[ctype], body, off); off += struct.calcsize(_FMT[ctype])
                        sz = struct.calcsize(_FMT[ptype])
                        row[pname] = list(struct.unpack_from(bo + _FMT[ptype] * int(n), body, off)); off += sz * int(n)
                else:
                    if toks is not None:
                        row[pname] = float(toks[ti]); ti += 1
                    else:
                        (row[pname],) = struct.unpack_from(bo + _FMT[ptype], body, off); off += struct.calcsize(_FMT[ptype])
            out.append(row)
        rows[name] = (out, [p[0] for p in props])
    vrows, vprops = rows["vertex"]
    frows, fprops = rows.get("face", ([], []))
    f32 = lambda v: float(np.float32(v))
    pos = [[f32(r["x"]), f32(r["y"]), f32(r["z"])] for r in vrows]
    has_st = "s" in vprops and "t" in vprops
    tex = [[f32(r["s"]), f32(r["t"])] for r in vrows] if has_st else []
    vn = all(k in vprops for k in ("nx", "ny", "nz"))
    fn = not vn and all(k in fprops for k in ("nx", "ny", "nz"))
    nrm = ([[f32(r["nx"]), f32(r["ny"]), f32(r["nz"])] for r in vrows] if vn else
           [[f32(r["nx"]), f32(r["ny"]), f32(r["nz"])] for r in frows] if fn else [])
    verts, tris = [], []
    for fi, r in enumerate(frows):
        idx = [int(v) for v in r.get("vertex_indices", r.get("vertex_index", []))]
        for k in range(1, len(idx) - 1):
            s = len(verts)
            for v in (idx[0], idx[k], idx[k + 1]):
                verts.append((v, v if has_st else NONE, v if vn else (fi if fn else NONE)))
            tris.append((s, s + 1, s + 2))
    return {"positions": np.array(pos, np.float32).reshape(-1, 3), "texcoords": np.array(tex, np.float32).reshape(-1, 2),
            "normals": np.array(nrm, np.float32).reshape(-1, 3), "vertices": np.array(verts, np.uint32).reshape(-1, 3),
            "triangles": np.array(tris, np.uint32).reshape(-1, 3)}
