"""CPU oracle of the U-Net denoiser (oracle/denoise_oracle.c) -- TEST
INFRASTRUCTURE ONLY.  Double precision; with round=True it rounds to bf16
exactly where the library stores bf16.  Pinned in tests/test_denoise_oracle.py
by finite differences and by torch's CPU conv2d."""
import ctypes as C
import os

import numpy as np

from . import HERE, build

LIB_PATH = os.path.join(HERE, "_build", "libdno.so")
IN, OUT, MASK = 12, 4, 11
_lib = None
P = C.c_void_p


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        u32, i32 = C.c_uint32, C.c_int
        sigs = {
            "dno_conv3x3": (None, [P, u32, i32, P, u32, u32, u32, u32, P, P, u32, i32, P]),
            "dno_conv3x3_wgrad": (None, [P, u32, i32, P, u32, u32, u32, u32, P, u32, P, P]),
            "dno_param_count": (C.c_uint64, [u32, u32]),
            "dno_forward": (i32, [u32, u32, P, P, u32, u32, u32, i32, P]),
            "dno_backward": (i32, [u32, u32, P, P, P, u32, u32, u32, i32, P, P, P]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return np.ascontiguousarray(a, np.float64)


def _p(a):
    return None if a is None else a.ctypes.data_as(P)


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def conv3x3(a, c1, up, b, c2, n, h, w, weights, bias, cout, relu):
    a, weights = _d(a), _d(weights)
    b = None if b is None else _d(b)
    bias = None if bias is None else _d(bias)
    out = np.empty((n, h, w, cout), np.float64)
    lib().dno_conv3x3(_p(a), c1, int(up), _p(b), c2, n, h, w, _p(weights), _p(bias), cout, int(relu), _p(out))
    return out


def conv3x3_wgrad(a, c1, up, b, c2, n, h, w, g, cout):
    a, g = _d(a), _d(g)
    b = None if b is None else _d(b)
    dw = np.empty((cout, 9, c1 + c2), np.float64)
    db = np.empty(cout, np.float64)
    lib().dno_conv3x3_wgrad(_p(a), c1, int(up), _p(b), c2, n, h, w, _p(g), cout, _p(dw), _p(db))
    return dw, db


def param_count(levels, base):
    return int(lib().dno_param_count(levels, base))


def forward(levels, base, params, inp, round=True):
    params, inp = _d(params), _d(inp)
    n, h, w, _ = inp.shape
    out = np.empty((n, h, w, OUT), np.float64)
    assert lib().dno_forward(levels, base, _p(params), _p(inp), n, h, w, int(round), _p(out)) == 0
    return out


def backward(levels, base, params, inp, target, round=True):
    """(loss, gradient, output)"""
    params, inp, target = _d(params), _d(inp), _d(target)
    n, h, w, _ = inp.shape
    g = np.empty_like(params)
    out = np.empty((n, h, w, OUT), np.float64)
    loss = C.c_double()
    assert lib().dno_backward(levels, base, _p(params), _p(inp), _p(target), n, h, w, int(round), _p(g),
                              C.byref(loss), _p(out)) == 0
    return loss.value, g, out
