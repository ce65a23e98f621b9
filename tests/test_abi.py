"""The C-ABI library: it loads, exports every symbol include/*.h declares, and
rejects bad arguments with a status and a message (no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import fornos_amd as fa
from fornos_amd import _abi
from fornos_amd._abi import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            text = open(os.path.join(ROOT, "include", h)).read()
            names |= set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(fornos_[a-z_0-9]+)\s*\(", text, re.M))
    return names


def test_every_declared_symbol_is_exported():
    names = declared_symbols()
    assert len(names) >= 30
    L = lib()
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_declarations():
    assert declared_symbols() == set(_abi.SYMBOLS)


def test_version_and_device_count():
    assert b"gfx950" in lib().fornos_version()
    assert lib().fornos_device_count() >= 0


def test_invalid_arguments_return_status():
    L = lib()
    out = C.c_void_p()
    st = L.fornos_mesh_generate(99, 0, C.byref(_abi.GenParams(1.0, 4, 4)), C.byref(out))
    assert st == _abi.INVALID_ARG
    assert b"unknown generator" in L.fornos_last_error()
    assert L.fornos_scene_create(None, None, C.byref(out)) == _abi.INVALID_ARG
    assert L.fornos_solver_create(1, None, None, None, 0, C.byref(out)) == _abi.INVALID_ARG
    assert L.fornos_sample_table(0, 1, 0, None, None) == _abi.INVALID_ARG


def test_missing_texcoords_is_reported():
    # a hi-poly mesh has no texcoords: MapUV::fromMesh returns nullptr and
    # FornosRunner::start reports it (Src/fornos.cpp:150-153)
    hi = fa.Mesh.generate(_abi.GEN_ICOSPHERE_WAVES, 1, 0.02, 1, 0)
    with pytest.raises(fa.FornosError) as e:
        fa.CompressedMapUV(hi, None, 16, 16)
    assert e.value.status == _abi.INVALID_ARG
    assert "texture coordinates" in str(e.value)


def test_mesh_from_view_validates_indices():
    pos = np.zeros((3, 3), np.float32)
    verts = np.array([[0, 0, 0], [1, 0, 0], [5, 0, 0]], np.uint32)
    tris = np.array([[0, 1, 2]], np.uint32)
    with pytest.raises(fa.FornosError):
        fa.Mesh.from_arrays(pos, verts, tris, normals=np.zeros((1, 3), np.float32))


def test_scene_without_device_is_unsupported():
    if lib().fornos_device_count() > 0:
        pytest.skip("a device is visible")
    hi = fa.Mesh.generate(_abi.GEN_ICOSPHERE_WAVES, 1, 0.02, 1, 0)
    with pytest.raises(fa.FornosError) as e:
        fa.Scene(hi)
    assert e.value.status == _abi.UNSUPPORTED
