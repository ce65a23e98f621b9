"""Image export of baked maps (include/fornos_export.h): the reference's
exportFloatImage / exportNormalImage / exportVectorImage (Src/image.cpp) on the
solver's device results, and the per-task export each solver's Task::finish
performs (Src/solver_*.cpp)."""
import ctypes as C

import numpy as np

from . import _abi
from ._abi import check, lib

DEFAULT_DILATION = 16  # FornosParameters::shared.texDilation (Src/fornos.h:47)


def _size(solver):
    w, h = C.c_uint32(), C.c_uint32()
    check(lib().fornos_solver_map_size(solver.handle, C.byref(w), C.byref(h)))
    return w.value, h.value


def float_rgb8(solver, filter_range=(0.0, 0.0), normalize=False, dilate=0):
    """(rgb [h, w, 3] uint8, (min, max)) as exportFloatImage's 8-bit branch."""
    w, h = _size(solver)
    rgb = np.empty((h, w, 3), np.uint8)
    mm = np.empty(2, np.float32)
    check(lib().fornos_export_float_rgb8(solver.handle, float(filter_range[0]), float(filter_range[1]),
                                         int(bool(normalize)), int(dilate), rgb.ctypes.data_as(C.c_void_p),
                                         mm.ctypes.data_as(C.c_void_p)))
    return rgb, (float(mm[0]), float(mm[1]))


def float_f32(solver, filter_range=(0.0, 0.0), normalize=False):
    w, h = _size(solver)
    img = np.empty((h, w), np.float32)
    mm = np.empty(2, np.float32)
    check(lib().fornos_export_float_f32(solver.handle, float(filter_range[0]), float(filter_range[1]),
                                        int(bool(normalize)), img.ctypes.data_as(C.c_void_p),
                                        mm.ctypes.data_as(C.c_void_p)))
    return img, (float(mm[0]), float(mm[1]))


def normal_rgb8(solver, dilate=0):
    w, h = _size(solver)
    rgb = np.empty((h, w, 3), np.uint8)
    check(lib().fornos_export_normal_rgb8(solver.handle, int(dilate), rgb.ctypes.data_as(C.c_void_p)))
    return rgb


def vector_planes(solver):
    w, h = _size(solver)
    planes = np.empty((3, h, w), np.float32)
    check(lib().fornos_export_vector_f32(solver.handle, planes.ctypes.data_as(C.c_void_p)))
    return planes


def exportFloatImage(solver, path, filterRange=(0.0, 0.0), normalize=False, dilate=0):
    mm = np.empty(2, np.float32)
    check(lib().fornos_export_float_image(solver.handle, str(path).encode(), float(filterRange[0]),
                                          float(filterRange[1]), int(bool(normalize)), int(dilate),
                                          mm.ctypes.data_as(C.c_void_p)))
    return float(mm[0]), float(mm[1])


def exportNormalImage(solver, path, dilate=0):
    check(lib().fornos_export_normal_image(solver.handle, str(path).encode(), int(dilate)))


def exportVectorImage(solver, path):
    check(lib().fornos_export_vector_image(solver.handle, str(path).encode()))


def export_task(solver, path, dilation=DEFAULT_DILATION):
    """What each solver's Task::finish exports: AO and thickness normalised
    with no filter (Src/solver_ao.cpp:161, solver_thickness.cpp:162), height
    filtered to [0, maxDistance] and normalised per its parameter
    (solver_height.cpp:109-116), normals and bent normals as normal images
    (solver_normals.cpp:121, solver_bentnormals.cpp:172), positions as a
    vector EXR (solver_position.cpp:109).  Returns the value range for the
    float maps."""
    k = solver.kind
    if k in (_abi.SOLVER_AO, _abi.SOLVER_THICKNESS):
        return exportFloatImage(solver, path, (0.0, 0.0), True, dilation)
    if k == _abi.SOLVER_HEIGHT:
        p = solver.params
        return exportFloatImage(solver, path, (0.0, p.max_distance), bool(p.normalize_output), dilation)
    if k in (_abi.SOLVER_NORMALS, _abi.SOLVER_BENT_NORMALS):
        return exportNormalImage(solver, path, dilation)
    return exportVectorImage(solver, path)
