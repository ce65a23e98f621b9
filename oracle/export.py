"""CPU oracle of the image export (Src/image.cpp) -- TEST INFRASTRUCTURE ONLY.

numpy restatement, in float32 arithmetic with the reference's operation
order, of getMinMax / computeScaleBias, the 8-bit and float branches of
exportFloatImage, exportNormalImage, exportVectorImage, the valid-pixel tables
and dilateRGB, plus tinyexr's float -> half conversion.  Pinned by hand-made
cases and by numpy's own float16 conversion in tests/test_export_oracle.py.

Orientation: texel index idx = y * w + x goes to image row h - y - 1.
"""
import numpy as np

FLT_MAX = np.float32(3.4028234663852886e38)
OFFSETS = [(1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (1, -1), (-1, 1), (-1, -1)]  # Src/image.cpp:131


def minmax(values):
    """getMinMax (Src/image.cpp:37-47): fminf / fmaxf, so NaN is ignored."""
    v = np.asarray(values, np.float32)
    v = v[~np.isnan(v)]
    if len(v) == 0:
        return FLT_MAX, -FLT_MAX
    return np.float32(v.min()), np.float32(v.max())


def scale_bias(mn, mx, normalize):
    """computeScaleBias (Src/image.cpp:72-80)."""
    if normalize and mn != FLT_MAX and mx != -FLT_MAX:
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            scale = np.float32(1.0) / np.float32(mx - mn)
            return scale, np.float32(-mn * scale)
    return np.float32(1.0), np.float32(0.0)


def filtered_range(values, lo, hi):
    """The range exportFloatImage uses (Src/image.cpp:199-204)."""
    tmn, tmx = minmax(values)
    lo, hi = np.float32(lo), np.float32(hi)
    if hi > lo:
        return np.fmax(tmn, lo), np.fmin(tmx, hi)
    return tmn, tmx


def to_u8(x):
    """(uint8_t)(float) on x86: truncating int32 conversion (NaN and out of
    range -> 0x80000000), low byte."""
    x = np.asarray(x, np.float32)
    bad = np.isnan(x) | (x >= 2147483648.0) | (x < -2147483648.0)
    i = np.where(bad, np.int64(-2147483648), np.trunc(np.where(bad, 0, x)).astype(np.int64))
    return (i & 0xFF).astype(np.uint8)


def _pix(indices, w, h):
    idx = np.asarray(indices, np.int64)
    x, y = idx % w, idx // w
    return (h - y - 1) * w + x


def float_rgb8(values, indices, w, h, lo=0.0, hi=0.0, normalize=False, dilate=0):
    """exportFloatImage, PNG / TGA branch (Src/image.cpp:209-240) -> (rgb [h*w*3], (min, max))."""
    v = np.asarray(values, np.float32)
    mn, mx = filtered_range(v, lo, hi)
    sc, bi = scale_bias(mn, mx, normalize)
    lo32, hi32 = np.float32(lo), np.float32(hi)
    filt = hi32 > lo32
    with np.errstate(invalid="ignore", over="ignore"):
        df = np.where(filt & ((v < lo32) | (v > hi32)), np.float32(0), v).astype(np.float32)
        t = (df * sc).astype(np.float32) + bi
        t = t.astype(np.float32)
        m = np.where(t < 0, np.float32(0), t)
        m = np.where(np.float32(1) < m, np.float32(1), m).astype(np.float32)
        c = to_u8((m * np.float32(255)).astype(np.float32))
    rgb = np.zeros((w * h, 3), np.uint8)
    rgb[_pix(indices, w, h)] = c[:, None]
    if dilate > 0:
        valid = np.zeros(w * h, bool)
        valid[np.asarray(indices, np.int64)] = True
        rgb = dilate_rgb(rgb, valid, w, h, dilate)
    return rgb.reshape(-1), (mn, mx)


def float_f32(values, indices, w, h, lo=0.0, hi=0.0, normalize=False):
    """exportFloatImage, EXR branch (Src/image.cpp:241-260) -> (plane [h*w], (min, max))."""
    v = np.asarray(values, np.float32)
    mn, mx = filtered_range(v, lo, hi)
    sc, bi = scale_bias(mn, mx, normalize)
    lo32, hi32 = np.float32(lo), np.float32(hi)
    with np.errstate(invalid="ignore", over="ignore"):
        df = np.fmin(np.fmax(v, lo32), hi32) if hi32 > lo32 else v
        t = ((df * sc).astype(np.float32) + bi).astype(np.float32) if normalize else df
    img = np.zeros(w * h, np.float32)
    img[_pix(indices, w, h)] = t
    return img, (mn, mx)


def normal_rgb8(values, indices, w, h, dilate=0):
    """exportNormalImage, PNG / TGA branch (Src/image.cpp:364-386); the valid
    table takes texels whose colour is not (0, 0, 0), row idx // w."""
    v = np.asarray(values, np.float32).reshape(-1, 3)
    with np.errstate(invalid="ignore", over="ignore"):
        n = ((v * np.float32(0.5)).astype(np.float32) + np.float32(0.5)).astype(np.float32)
        c = to_u8((n * np.float32(255)).astype(np.float32))
    rgb = np.zeros((w * h, 3), np.uint8)
    pix = _pix(indices, w, h)
    rgb[pix] = c
    if dilate > 0:
        valid = np.zeros(w * h, bool)
        valid[np.asarray(indices, np.int64)] = np.any(rgb[pix] != 0, axis=1)
        rgb = dilate_rgb(rgb, valid, w, h, dilate)
    return rgb.reshape(-1)


def vector_planes(values, indices, w, h):
    """exportVectorImage (Src/image.cpp:299-306): planes B, G, R = z, y, x."""
    v = np.asarray(values, np.float32).reshape(-1, 3)
    planes = np.zeros((3, w * h), np.float32)
    pix = _pix(indices, w, h)
    planes[0, pix], planes[1, pix], planes[2, pix] = v[:, 2], v[:, 1], v[:, 0]
    return planes


def dilate_rgb(rgb, valid, w, h, max_dist):
    """dilateRGB (Src/image.cpp:127-183) on rgb [w*h, 3] (image rows) and
    valid [w*h] (map order), as its loops run on one thread: the pixels in
    map row-major order (y, then x), each invalid one taking, in place, the
    colour the image holds at the first centre pos + offset * dist
    (dist < max_dist) strictly inside the border whose eight neighbours are
    valid -- so a centre the loop has already filled gives its new colour."""
    out = rgb.copy()
    vm = valid.reshape(h, w)
    # centres whose eight neighbours are valid
    good = np.zeros((h, w), bool)
    inner = np.ones((h - 2, w - 2), bool) if h > 2 and w > 2 else np.zeros((0, 0), bool)
    for ox, oy in OFFSETS:
        if inner.size:
            inner &= vm[1 + oy:h - 1 + oy, 1 + ox:w - 1 + ox]
    if inner.size:
        good[1:h - 1, 1:w - 1] = inner
    ys, xs = np.nonzero(~vm)
    for y, x in zip(ys.tolist(), xs.tolist()):
        done = False
        for dist in range(1, max_dist):
            for ox, oy in OFFSETS:
                cx, cy = x + ox * dist, y + oy * dist
                if 0 < cx < w - 1 and 0 < cy < h - 1 and good[cy, cx]:
                    out[(h - y - 1) * w + x] = out[(h - cy - 1) * w + cx]
                    done = True
                    break
            if done:
                break
    return out


def float_to_half(values):
    """tinyexr's float_to_half_full (3rdParty/tinyexr/tinyexr.h:7044-7080)."""
    f = np.ascontiguousarray(values, np.float32).view(np.uint32).astype(np.int64)
    sign = (f >> 16) & 0x8000
    e = (f >> 23) & 0xFF
    m = f & 0x7FFFFF
    ne = e - 127 + 15
    h = np.zeros_like(f)
    special = e == 255
    h = np.where(special, (31 << 10) | np.where(m != 0, 0x200, 0), h)
    normal = (e != 0) & ~special
    over = normal & (ne >= 31)
    h = np.where(over, 31 << 10, h)
    mid = normal & (ne > 0) & (ne < 31)
    hm = (np.clip(ne, 0, 31) << 10) | (m >> 13)
    hm = hm + ((m & 0x1000) != 0)
    h = np.where(mid, hm, h)
    under = normal & (ne <= 0) & ((14 - ne) <= 24)
    sh = np.clip(14 - ne, 0, 63)
    mant = m | 0x800000
    hu = (mant >> sh) + ((mant >> np.clip(sh - 1, 0, 63)) & 1)
    h = np.where(under, hu, h)
    return ((h & 0x7FFF) | sign).astype(np.uint16)


def read_exr(path):
    """Minimal reader of the library's EXR files (uncompressed scanlines,
    half channels): {name: float32 [h, w]}."""
    import struct
    data = open(path, "rb").read()
    assert data[:4] == b"\x76\x2f\x31\x01"
    p = 8
    attrs = {}
    while data[p] != 0:
        e = data.index(b"\0", p)
        name = data[p:e].decode()
        e2 = data.index(b"\0", e + 1)
        size = struct.unpack_from("<i", data, e2 + 1)[0]
        attrs[name] = data[e2 + 5:e2 + 5 + size]
        p = e2 + 5 + size
    p += 1
    chans = []
    cl = attrs["channels"]
    q = 0
    while cl[q] != 0:
        e = cl.index(b"\0", q)
        chans.append((cl[q:e].decode(), struct.unpack_from("<i", cl, e + 1)[0]))
        q = e + 17
    assert attrs["compression"] == b"\0"
    x0, y0, x1, y1 = struct.unpack("<4i", attrs["dataWindow"])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    offs = struct.unpack_from("<%dQ" % h, data, p)
    out = {c: np.zeros((h, w), np.float32) for c, _ in chans}
    for y in range(h):
        o = offs[y]
        yy, size = struct.unpack_from("<ii", data, o)
        row = np.frombuffer(data, np.float16, count=w * len(chans), offset=o + 8).astype(np.float32)
        for k, (c, t) in enumerate(chans):
            assert t == 1
            out[c][yy] = row[k * w:(k + 1) * w]
    return out
