"""ctypes binding of libfornos_hip.so (include/fornos_*.h).

The library is built in-tree by fornos_amd.build; loading fails loudly when it
is missing, so no code path silently runs without the native library.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfornos_hip.so")

OK, INVALID_ARG, OOM, HIP_ERROR, UNSUPPORTED = 0, 1, 2, 3, 4

GEN_UV_SPHERE, GEN_ICOSPHERE_WAVES, GEN_HEIGHTFIELD, GEN_HEIGHTFIELD_LOW, GEN_TORUS_KNOT = range(5)
SOLVER_HEIGHT, SOLVER_POSITION, SOLVER_NORMALS, SOLVER_AO, SOLVER_BENT_NORMALS, SOLVER_THICKNESS = range(6)
BVH_REFERENCE = 0


class MeshView(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("n_positions", C.c_uint32),
                ("normals", C.c_void_p), ("n_normals", C.c_uint32),
                ("texcoords", C.c_void_p), ("n_texcoords", C.c_uint32),
                ("tangents", C.c_void_p), ("bitangents", C.c_void_p),
                ("vertices", C.c_void_p), ("n_vertices", C.c_uint32),
                ("triangles", C.c_void_p), ("n_triangles", C.c_uint32)]


class BVHParams(C.Structure):
    _fields_ = [("layout", C.c_int), ("max_tris_per_leaf", C.c_uint32), ("max_depth", C.c_uint32)]


class SceneInfo(C.Structure):
    _fields_ = [("n_nodes", C.c_uint32), ("n_triangles", C.c_uint32), ("n_leaves", C.c_uint32),
                ("max_leaf_tris", C.c_uint32), ("depth", C.c_uint32),
                ("build_seconds", C.c_double), ("device_bytes", C.c_size_t), ("ao_groups", C.c_uint32),
                ("sah", C.c_double), ("sah_ao", C.c_double), ("leaf_group_nodes", C.c_uint32)]


class TexelView(C.Structure):
    _fields_ = [("count", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32),
                ("positions", C.c_void_p), ("directions", C.c_void_p), ("normals", C.c_void_p),
                ("tangents", C.c_void_p), ("bitangents", C.c_void_p), ("indices", C.c_void_p)]


class SolverParams(C.Structure):
    _fields_ = [("sample_count", C.c_uint32), ("min_distance", C.c_float),
                ("max_distance", C.c_float), ("tangent_space", C.c_int),
                ("normalize_output", C.c_int)]


class GenParams(C.Structure):
    _fields_ = [("a", C.c_float), ("b", C.c_uint32), ("c", C.c_uint32)]


class ParamsShared(C.Structure):
    """fornos_params_shared (include/fornos_runner.h)."""
    _fields_ = [("lo_poly_mesh_path", C.c_char_p), ("hi_poly_mesh_path", C.c_char_p),
                ("lo_poly_mesh_normal", C.c_int), ("hi_poly_mesh_normal", C.c_int),
                ("bvh_tris_per_node", C.c_int), ("tex_width", C.c_int), ("tex_height", C.c_int),
                ("tex_dilation", C.c_int), ("ignore_backfaces", C.c_int), ("mapping", C.c_int),
                ("mapping_edge", C.c_float), ("sample_seed", C.c_uint32)]


class ParamsMap(C.Structure):
    """fornos_params_map (include/fornos_runner.h)."""
    _fields_ = [("enabled", C.c_int), ("output_path", C.c_char_p), ("sample_count", C.c_int),
                ("min_distance", C.c_float), ("max_distance", C.c_float), ("tangent_space", C.c_int),
                ("normalize_output", C.c_int)]


class Parameters(C.Structure):
    """fornos_parameters (include/fornos_runner.h)."""
    _fields_ = [("shared", ParamsShared), ("height", ParamsMap), ("positions", ParamsMap), ("normals", ParamsMap),
                ("ao", ParamsMap), ("bent_normals", ParamsMap), ("thickness", ParamsMap)]


NORMALS_IMPORT, NORMALS_PER_FACE, NORMALS_PER_VERTEX = range(3)
MAPPING_SMOOTH, MAPPING_LOWPOLY_NORMALS, MAPPING_HYBRID = range(3)

H = C.c_void_p
P = C.c_void_p
u32p = C.POINTER(C.c_uint32)

# (name, restype, argtypes); restype None means an int status.
_SIGS = [
    ("fornos_last_error", C.c_char_p, []),
    ("fornos_version", C.c_char_p, []),
    ("fornos_device_count", C.c_int, []),
    # host
    ("fornos_mesh_generate", None, [C.c_int, C.c_uint64, C.POINTER(GenParams), C.POINTER(H)]),
    ("fornos_mesh_from_view", None, [C.POINTER(MeshView), C.POINTER(H)]),
    ("fornos_mesh_copy", None, [H, C.POINTER(H)]),
    ("fornos_mesh_view_get", None, [H, C.POINTER(MeshView)]),
    ("fornos_mesh_compute_vertex_normals", None, [H]),
    ("fornos_mesh_compute_vertex_normals_aggressive", None, [H]),
    ("fornos_mesh_compute_tangent_space", None, [H]),
    ("fornos_mesh_compute_face_normals", None, [H]),
    ("fornos_mesh_load", None, [C.c_char_p, C.POINTER(H)]),
    ("fornos_mesh_destroy", "void", [H]),
    ("fornos_texels_create", None, [H, H, C.c_uint32, C.c_uint32, C.POINTER(H)]),
    ("fornos_texels_create_hybrid", None, [H, H, C.c_uint32, C.c_uint32, C.c_float, C.POINTER(H)]),
    ("fornos_texels_view_get", None, [H, C.POINTER(TexelView)]),
    ("fornos_texels_export_mapuv", None, [H, P, P, P, P, P]),
    ("fornos_texels_destroy", "void", [H]),
    ("fornos_bvh_reference_export", None, [H, C.c_uint32, C.c_uint32, u32p, P, P]),
    ("fornos_bvh_reference_links", None, [H, C.c_uint32, C.c_uint32, C.c_uint32, P, P]),
    ("fornos_sample_table", None, [C.c_uint32, C.c_uint32, C.c_uint32, P, P]),
    # device
    ("fornos_scene_create", None, [C.POINTER(MeshView), C.POINTER(BVHParams), C.POINTER(H)]),
    ("fornos_scene_info_get", None, [H, C.POINTER(SceneInfo)]),
    ("fornos_scene_export_reference_nodes", None, [H, P, P]),
    ("fornos_scene_export_soup", None, [H, P, P]),
    ("fornos_scene_destroy", "void", [H]),
    ("fornos_meshmap_create", None, [H, C.POINTER(TexelView), C.c_int, C.POINTER(H)]),
    ("fornos_meshmap_create_from_mesh", None, [H, C.POINTER(MeshView), C.POINTER(MeshView), C.c_uint32,
                                                C.c_uint32, C.c_int, C.POINTER(H)]),
    ("fornos_meshmap_create_from_mesh_hybrid", None, [H, C.POINTER(MeshView), C.POINTER(MeshView), C.c_uint32,
                                                       C.c_uint32, C.c_float, C.c_int, C.POINTER(H)]),
    ("fornos_meshmap_read_texels", None, [H, P, P, P, P, P, P]),
    ("fornos_meshmap_run_step", None, [H, C.POINTER(C.c_int)]),
    ("fornos_meshmap_run_all", None, [H, C.POINTER(C.c_float)]),
    ("fornos_meshmap_progress", None, [H, C.POINTER(C.c_float)]),
    ("fornos_meshmap_reset", None, [H]),
    ("fornos_meshmap_set_predicate", None, [H, C.c_int]),
    ("fornos_meshmap_set_traversal", None, [H, C.c_int, C.c_uint32]),
    ("fornos_meshmap_redo_counts", None, [H, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    ("fornos_meshmap_submit", None, [H, C.c_uint32]),
    ("fornos_meshmap_poll", None, [H, C.POINTER(C.c_float), C.POINTER(C.c_int)]),
    ("fornos_meshmap_read", None, [H, P, P]),
    ("fornos_meshmap_texel_count", None, [H, u32p]),
    ("fornos_meshmap_destroy", "void", [H]),
    ("fornos_solver_create", None, [C.c_int, C.POINTER(SolverParams), H, P, C.c_uint32, C.POINTER(H)]),
    ("fornos_solver_run_step", None, [H, C.POINTER(C.c_int)]),
    ("fornos_solver_run_all", None, [H, C.POINTER(C.c_float)]),
    ("fornos_solver_progress", None, [H, C.POINTER(C.c_float)]),
    ("fornos_solver_reset", None, [H]),
    ("fornos_solver_pair", None, [H, H]),
    ("fornos_solver_submit", None, [H, C.c_uint32]),
    ("fornos_solver_poll", None, [H, C.POINTER(C.c_float), C.POINTER(C.c_int)]),
    ("fornos_solver_read", None, [H, P, C.c_size_t]),
    ("fornos_solver_read_counts", None, [H, P]),
    ("fornos_solver_last_timing", None, [H, C.POINTER(C.c_float), C.POINTER(C.c_float), u32p]),
    ("fornos_solver_count_work", None, [H, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                        C.POINTER(C.c_uint64)]),
    ("fornos_solver_work_counters", None, [H, P, C.c_uint32]),
    ("fornos_solver_run_all_noalloc", None, [H, P, P]),
    ("fornos_solver_run_texels", None, [H, P, C.c_uint32, C.POINTER(C.c_float)]),
    ("fornos_solver_fallback_flags", None, [H, P]),
    ("fornos_solver_set_launch", None, [H, C.c_uint32, C.c_uint32, C.c_uint32]),
    ("fornos_solver_get_launch", None, [H, u32p, u32p, u32p]),
    ("fornos_solver_set_ray_order", None, [H, C.c_int]),
    ("fornos_solver_set_tile", None, [H, C.c_uint32]),
    ("fornos_solver_get_tile", None, [H, u32p]),
    ("fornos_solver_autotune", None, [H, C.c_float, u32p, u32p, u32p, C.POINTER(C.c_float)]),
    ("fornos_solver_get_ray_order", None, [H, C.POINTER(C.c_int)]),
    ("fornos_solver_effective_ray_order", None, [H, C.POINTER(C.c_int)]),
    ("fornos_solver_destroy", "void", [H]),
    ("fornos_batch_create", None, [C.POINTER(H), C.c_uint32, C.POINTER(H)]),
    ("fornos_batch_set_launch", None, [H, C.c_uint32, C.c_uint32, C.c_uint32]),
    ("fornos_batch_get_launch", None, [H, u32p, u32p, u32p]),
    ("fornos_batch_autotune", None, [H, C.c_float, u32p, u32p, u32p, C.POINTER(C.c_float)]),
    ("fornos_batch_get_tile", None, [H, H, u32p]),
    ("fornos_batch_set_tile_shift", None, [H, C.c_int]),
    ("fornos_batch_get_tile_shift", None, [H, C.POINTER(C.c_int)]),
    ("fornos_batch_run", None, [H, C.POINTER(C.c_float)]),
    ("fornos_batch_destroy", "void", [H]),
    # denoiser (include/fornos_denoise.h)
    ("fornos_denoiser_create", None, [C.c_uint32, C.c_uint32, C.c_uint64, C.POINTER(H)]),
    ("fornos_denoiser_param_count", None, [H, C.POINTER(C.c_uint64)]),
    ("fornos_denoiser_layer_count", None, [H, u32p]),
    ("fornos_denoiser_layer", None, [H, C.c_uint32, u32p, u32p, u32p, u32p, C.POINTER(C.c_uint64)]),
    ("fornos_denoiser_get_params", None, [H, P, C.c_uint64]),
    ("fornos_denoiser_set_params", None, [H, P, C.c_uint64]),
    ("fornos_denoiser_forward", None, [H, P, C.c_uint32, C.c_uint32, C.c_uint32, P, C.POINTER(C.c_float)]),
    ("fornos_denoiser_backward", None, [H, P, P, C.c_uint32, C.c_uint32, C.c_uint32, P, C.POINTER(C.c_float)]),
    ("fornos_denoiser_get_grads", None, [H, P, C.c_uint64]),
    ("fornos_denoiser_adam", None, [H, C.c_float, C.c_float, C.c_float, C.c_float]),
    ("fornos_denoiser_destroy", "void", [H]),
    ("fornos_denoise_gather", None, [H, H, H, H, H, H, P, C.c_float, C.c_float, P]),
    ("fornos_denoise_scatter", None, [H, P, P, P]),
    ("fornos_denoise_check_tiles", None, [P, P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float,
                                          C.c_float, P]),
    # export (include/fornos_export.h)
    ("fornos_export_float_rgb8", None, [H, C.c_float, C.c_float, C.c_int, C.c_int, P, P]),
    ("fornos_export_float_f32", None, [H, C.c_float, C.c_float, C.c_int, P, P]),
    ("fornos_export_normal_rgb8", None, [H, C.c_int, P]),
    ("fornos_export_vector_f32", None, [H, P]),
    ("fornos_solver_map_size", None, [H, u32p, u32p]),
    ("fornos_export_float_image", None, [H, C.c_char_p, C.c_float, C.c_float, C.c_int, C.c_int, P]),
    ("fornos_export_normal_image", None, [H, C.c_char_p, C.c_int]),
    ("fornos_export_vector_image", None, [H, C.c_char_p]),
    ("fornos_write_png", None, [C.c_char_p, C.c_uint32, C.c_uint32, P]),
    ("fornos_write_tga", None, [C.c_char_p, C.c_uint32, C.c_uint32, P]),
    ("fornos_write_exr", None, [C.c_char_p, C.c_uint32, C.c_uint32, C.c_uint32, P, P]),
    ("fornos_float_to_half", "void", [P, P, C.c_size_t]),
    ("fornos_parameters_defaults", "void", [C.POINTER(Parameters)]),
    ("fornos_runner_create", None, [C.POINTER(H)]),
    ("fornos_runner_start", None, [H, C.POINTER(Parameters)]),
    ("fornos_runner_pending", C.c_int, [H]),
    ("fornos_runner_run", None, [H]),
    ("fornos_runner_run_all", None, [H]),
    ("fornos_runner_current_task", None, [H, C.POINTER(C.c_char_p), C.POINTER(C.c_float)]),
    ("fornos_runner_exported_range", C.c_int, [H, C.c_int, P]),
    ("fornos_runner_destroy", "void", [H]),
    ("fornos_dn_conv3x3", None, [P, C.c_uint32, C.c_int, P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, P, P,
                                 C.c_uint32, C.c_int, P]),
    ("fornos_dn_conv3x3_wgrad", None, [P, C.c_uint32, C.c_int, P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                       P, C.c_uint32, P, P]),
]

SYMBOLS = [s[0] for s in _SIGS]


class FornosError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("fornos status %d: %s" % (status, msg))
        self.status = status


_lib = None


def lib():
    """Loads the native library (once); raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FornosError(UNSUPPORTED, "%s is missing: run `python -m fornos_amd.build` first" % LIB_PATH)
    # torch's ROCm wheel carries its own libamdhip64 with the same soname as
    # ROCm's.  Whichever is loaded first serves the whole process, and torch
    # cannot initialise its device on the other one.  Load torch's first when
    # torch is installed, so device tensors (used by the denoiser for memory)
    # and this library share one HIP runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    for name, res, args in _SIGS:
        f = getattr(L, name)
        f.argtypes = args
        if res is None:
            f.restype = C.c_int
        elif res == "void":
            f.restype = None
        else:
            f.restype = res
    _lib = L
    return L


def check(status):
    if status != OK:
        raise FornosError(status, lib().fornos_last_error().decode(errors="replace"))


def ptr(a):
    """Data pointer of a C-contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(C.c_void_p)


def view_array(addr, count, dtype, shape):
    """Copies `count` elements at a native address into a new numpy array."""
    if not addr or count == 0:
        return np.zeros(shape, dtype=dtype)
    n = int(np.prod(shape))
    buf = (C.c_byte * (n * np.dtype(dtype).itemsize)).from_address(addr)
    return np.frombuffer(buf, dtype=dtype, count=n).reshape(shape).copy()
