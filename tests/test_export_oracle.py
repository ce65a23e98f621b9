"""Image export: the oracle (oracle/export.py) pinned on hand-made cases,
tinyexr's half rounding against numpy's float16, and the library's host
writers (PNG, TGA, EXR) and half conversion checked by decoding their files
with Pillow and with a minimal EXR reader.  No GPU needed."""
import ctypes as C

import numpy as np
import pytest

from fornos_amd._abi import lib
from oracle import export as ox


def test_quantisation_known_answer():
    # 2 x 2 map, texels in map order; normalised over [0, 1]
    rgb, (mn, mx) = ox.float_rgb8([0.0, 0.5, 1.0, 0.25], [0, 1, 2, 3], 2, 2, normalize=True)
    img = rgb.reshape(2, 2, 3)[:, :, 0]
    # texel (x, y) -> row h - y - 1: texels 2, 3 (y = 1) in row 0
    assert img.tolist() == [[255, 63], [0, 127]]  # 0.5*255 = 127.5 and 0.25*255 = 63.75 truncate
    assert (mn, mx) == (0.0, 1.0)
    # not normalised: scale 1, bias 0 (the reference warns but does not force it)
    rgb, _ = ox.float_rgb8([0.5, 2.0, -1.0, 0.1], [0, 1, 2, 3], 2, 2)
    assert rgb.reshape(2, 2, 3)[:, :, 0].tolist() == [[0, 25], [127, 255]]


def test_filter_zeroes_in_8_bit_and_clamps_in_float():
    v = [0.0, 0.05, 0.2, 0.08]
    rgb, (mn, mx) = ox.float_rgb8(v, [0, 1, 2, 3], 2, 2, lo=0.0, hi=0.1, normalize=True)
    assert np.isclose(mn, 0.0) and np.isclose(mx, 0.1)
    img = rgb.reshape(2, 2, 3)[:, :, 0]
    assert img[0, 0] == 0            # 0.2 is outside the range: 0, not 255
    f, _ = ox.float_f32(v, [0, 1, 2, 3], 2, 2, lo=0.0, hi=0.1, normalize=True)
    assert f.reshape(2, 2)[0, 0] == 1.0  # clamped to 0.1, normalised


def test_dilation_takes_the_first_fully_surrounded_centre():
    w = h = 8
    idx = [y * w + x for y in range(2, 5) for x in range(2, 5)]  # a valid 3 x 3 block
    vals = np.linspace(0.1, 0.9, len(idx)).astype(np.float32)
    rgb, _ = ox.float_rgb8(vals, idx, w, h, normalize=True, dilate=5)
    img = rgb.reshape(h, w, 3)[::-1, :, 0]  # back to map rows
    centre = img[3, 3]
    # only (3, 3) has eight valid neighbours; (6, 3) reaches it at distance 3
    assert img[3, 6] == centre and img[6, 6] == centre and img[0, 3] == centre
    # with max distance 3 (distances 1 and 2) (6, 3) stays black
    rgb, _ = ox.float_rgb8(vals, idx, w, h, normalize=True, dilate=3)
    assert rgb.reshape(h, w, 3)[::-1, :, 0][3, 6] == 0
    # valid pixels never change
    rgb0, _ = ox.float_rgb8(vals, idx, w, h, normalize=True)
    a, b = rgb0.reshape(h, w, 3)[::-1], rgb.reshape(h, w, 3)[::-1]
    assert np.array_equal(a[2:5, 2:5], b[2:5, 2:5])


def test_dilation_fills_in_place_in_row_major_order():
    # dilateRGB fills the image in place as its loops visit the map's pixels
    # (y, then x): a centre that is itself a hole the loop has already filled
    # gives its new colour.  Map rows top to bottom ('#' = texel):
    mask = ["###.####",
            ".##..###",
            ".####.##",
            "###..###",
            "########",
            "########",
            "#.#.#..#",
            "###...##"]
    w = h = 8
    valid = np.array([[c == "#" for c in row] for row in mask])
    idx = np.flatnonzero(valid.ravel())
    vals = (0.2 + 0.7 * np.arange(len(idx)) / len(idx)).astype(np.float32)
    rgb, _ = ox.float_rgb8(vals, idx, w, h, normalize=False, dilate=4)
    img = rgb.reshape(h, w, 3)[::-1, :, 0]  # back to map rows
    # (6, 1) is a hole with eight valid neighbours; its first centre is the
    # texel (4, 1), two rows up, so it takes that colour
    assert not valid[6, 1] and valid[5:8, 0:3].sum() == 8
    assert img[6, 1] == img[4, 1] != 0
    # (6, 3), later on the same row, finds (6, 1) first (offset (-1, 0) at
    # distance 2, nothing earlier qualifies): it copies (6, 1)'s new colour,
    # where a read before any fill would give black
    assert img[6, 3] == img[4, 1]


def test_normal_quantisation_and_rgb_validity():
    n = np.array([[0, 0, 1], [1, 0, 0], [-1, -1, -1], [np.nan, 0, 0]], np.float32)
    rgb = ox.normal_rgb8(n, [0, 1, 2, 3], 2, 2, dilate=0).reshape(2, 2, 3)
    assert rgb[1, 0].tolist() == [127, 127, 255]
    assert rgb[1, 1].tolist() == [255, 127, 127]
    assert rgb[0, 0].tolist() == [0, 0, 0]       # (-1,-1,-1) -> black, invalid for dilation
    assert rgb[0, 1].tolist() == [0, 127, 127]   # NaN -> 0


def test_half_conversion_matches_numpy_away_from_ties_and_rounds_ties_up():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(20000) * 10 ** rng.uniform(-8, 5, 20000),
                        [0.0, -0.0, 1.0, 65504.0, 65519.0, 1e-5, -3.25, np.inf, -np.inf]]).astype(np.float32)
    h = ox.float_to_half(x)
    with np.errstate(over="ignore"):
        ref = x.astype(np.float16).view(np.uint16)
    bits = x.view(np.uint32)
    e = (bits >> 23) & 0xFF
    tie = np.where(e >= 113, (bits & 0x1FFF) == 0x1000, False)  # normal halves: exactly half an ulp
    tiny = np.abs(x) < 6.2e-5  # half denormal range: numpy keeps float denormals, tinyexr flushes them
    ok = ~tie & ~tiny
    assert np.array_equal(h[ok], ref[ok])
    t = np.array([1 + 2 ** -11, 2049.0, -(1 + 3 * 2 ** -11)], np.float32)  # ties
    assert ox.float_to_half(t).tolist() == [0x3C01, 0x6801, 0xBC02]
    assert ox.float_to_half(np.array([1e-40, np.nan], np.float32)).tolist() == [0, 0x7E00]
    # the library's conversion is the same function
    out = np.empty(len(x), np.uint16)
    lib().fornos_float_to_half(x.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p), len(x))
    assert np.array_equal(out, h)


def test_png_and_tga_round_trip(tmp_path):
    from PIL import Image
    rng = np.random.default_rng(1)
    w, h = 37, 19
    rgb = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    for ext, fn in (("png", lib().fornos_write_png), ("tga", lib().fornos_write_tga)):
        p = str(tmp_path / ("x." + ext))
        assert fn(p.encode(), w, h, rgb.ctypes.data_as(C.c_void_p)) == 0
        back = np.asarray(Image.open(p).convert("RGB"))
        assert np.array_equal(back, rgb), ext


def test_exr_round_trip(tmp_path):
    rng = np.random.default_rng(2)
    w, h = 23, 11
    planes = [rng.standard_normal(w * h).astype(np.float32) for _ in range(3)]
    names = (C.c_char_p * 3)(b"B", b"G", b"R")
    ptrs = (C.c_void_p * 3)(*[p.ctypes.data for p in planes])
    p = str(tmp_path / "x.exr")
    assert lib().fornos_write_exr(p.encode(), w, h, 3, names, ptrs) == 0
    img = ox.read_exr(p)
    for name, plane in zip("BGR", planes):
        expect = ox.float_to_half(plane).view(np.float16).astype(np.float32).reshape(h, w)
        assert np.array_equal(img[name], expect)
    bad = (C.c_char_p * 2)(b"R", b"B")
    assert lib().fornos_write_exr(p.encode(), w, h, 2, bad, ptrs) != 0  # names must be sorted


@pytest.mark.parametrize("w,h", [(1200, 1000), (1, 1), (700, 3)])
def test_png_segments_form_one_zlib_stream(tmp_path, w, h):
    # the writer deflates ~256 KiB row segments on several threads; the IDAT
    # must still be one zlib stream with the right Adler-32 (zlib checks it)
    # and Sub-filtered rows that decode to the input
    import struct
    import zlib
    from PIL import Image
    rng = np.random.default_rng(3)
    yy, xx = np.mgrid[0:h, 0:w]
    smooth = (128 + 100 * np.sin(xx / 37.0) * np.cos(yy / 53.0)).astype(np.uint8)
    rgb = np.stack([smooth, rng.integers(0, 256, (h, w), dtype=np.uint8), smooth[::-1]], axis=2).copy()
    p = str(tmp_path / "x.png")
    assert lib().fornos_write_png(p.encode(), w, h, rgb.ctypes.data_as(C.c_void_p)) == 0
    data = open(p, "rb").read()
    pos, idat = 8, b""
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        typ = data[pos + 4:pos + 8]
        if typ == b"IDAT":
            idat += data[pos + 8:pos + 8 + n]
        pos += 12 + n
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 3 * w + 1)
    assert np.all(raw[:, 0] == 1)
    rows = raw[:, 1:].astype(np.int64)
    out = rows.copy()
    for k in range(3, 3 * w):
        out[:, k] = (out[:, k] + out[:, k - 3]) & 0xFF
    assert np.array_equal(out.astype(np.uint8).reshape(h, w, 3), rgb)
    assert np.array_equal(np.asarray(Image.open(p).convert("RGB")), rgb)
