"""Python mirror of the reference's baking interface, over the C ABI.

Names follow the reference classes: MeshMapping.init/runStep/progress
(Src/meshmapping.h:42-57) and the solvers' init/runStep/getResults/progress
(Src/solver_*.h).  Arrays cross as numpy; device memory is owned by the
native handles.
"""
import ctypes as C
import functools

import numpy as np

from . import _abi
from ._abi import check, lib, ptr, view_array

SAMPLE_PERM_COUNT = 64 * 64  # k_samplePermCount, Src/solver_ao.cpp:34


class _Handle:
    _destroy = None

    def __init__(self):
        self._h = C.c_void_p()

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h and self._h.value:
            getattr(lib(), self._destroy)(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Mesh(_Handle):
    """The reference Mesh (Src/mesh.h:36-77), held by the native library."""
    _destroy = "fornos_mesh_destroy"

    @classmethod
    def generate(cls, generator, seed=0, a=0.0, b=0, c=0):
        m = cls()
        gp = _abi.GenParams(float(a), int(b), int(c))
        check(lib().fornos_mesh_generate(int(generator), int(seed), C.byref(gp), C.byref(m._h)))
        return m

    @classmethod
    def load(cls, path):
        """Mesh::loadFile (Src/mesh.cpp:612-617): .obj or .ply."""
        m = cls()
        check(lib().fornos_mesh_load(str(path).encode(), C.byref(m._h)))
        return m

    @classmethod
    def from_arrays(cls, positions, vertices, triangles, normals=None, texcoords=None):
        m = cls()
        keep = [np.ascontiguousarray(x, dtype=t) if x is not None else None for x, t in
                ((positions, np.float32), (vertices, np.uint32), (triangles, np.uint32),
                 (normals, np.float32), (texcoords, np.float32))]
        p, v, t, n, uv = keep
        mv = _abi.MeshView(ptr(p), len(p), ptr(n), 0 if n is None else len(n), ptr(uv),
                           0 if uv is None else len(uv), None, None, ptr(v), len(v), ptr(t), len(t))
        check(lib().fornos_mesh_from_view(C.byref(mv), C.byref(m._h)))
        return m

    def copy(self):
        m = Mesh()
        check(lib().fornos_mesh_copy(self._h, C.byref(m._h)))
        return m

    def mesh_view(self):
        mv = _abi.MeshView()
        check(lib().fornos_mesh_view_get(self._h, C.byref(mv)))
        return mv

    def arrays(self):
        """Copies of the mesh arrays as a dict of numpy arrays."""
        v = self.mesh_view()
        out = {
            "positions": view_array(v.positions, v.n_positions * 3, np.float32, (v.n_positions, 3)),
            "normals": view_array(v.normals, v.n_normals * 3, np.float32, (v.n_normals, 3)),
            "texcoords": view_array(v.texcoords, v.n_texcoords * 2, np.float32, (v.n_texcoords, 2)),
            "vertices": view_array(v.vertices, v.n_vertices * 3, np.uint32, (v.n_vertices, 3)),
            "triangles": view_array(v.triangles, v.n_triangles * 3, np.uint32, (v.n_triangles, 3)),
        }
        if v.tangents:
            out["tangents"] = view_array(v.tangents, v.n_vertices * 3, np.float32, (v.n_vertices, 3))
            out["bitangents"] = view_array(v.bitangents, v.n_vertices * 3, np.float32, (v.n_vertices, 3))
        return out

    @property
    def n_triangles(self):
        return self.mesh_view().n_triangles

    def computeFaceNormals(self):
        check(lib().fornos_mesh_compute_face_normals(self._h))

    def computeVertexNormals(self):
        check(lib().fornos_mesh_compute_vertex_normals(self._h))

    def computeVertexNormalsAggressive(self):
        check(lib().fornos_mesh_compute_vertex_normals_aggressive(self._h))

    def computeTangentSpace(self):
        check(lib().fornos_mesh_compute_tangent_space(self._h))


class CompressedMapUV(_Handle):
    """MapUV + CompressedMapUV (Src/compute.h:214-259) of a low-poly mesh."""
    _destroy = "fornos_texels_destroy"

    def __init__(self, lo, dirs, width, height, edge=0.0):
        """edge > 0: the Hybrid mapping method (MapUV::fromMeshes_Hybrid)."""
        super().__init__()
        check(lib().fornos_texels_create_hybrid(lo.handle, dirs.handle if dirs is not None else None,
                                                int(width), int(height), float(edge), C.byref(self._h)))
        self.width, self.height = int(width), int(height)

    def texel_view(self):
        tv = _abi.TexelView()
        check(lib().fornos_texels_view_get(self._h, C.byref(tv)))
        return tv

    @property
    def count(self):
        return self.texel_view().count

    def arrays(self):
        tv = self.texel_view()
        n = tv.count
        f = lambda a: view_array(a, n * 3, np.float32, (n, 3))
        return {"positions": f(tv.positions), "directions": f(tv.directions), "normals": f(tv.normals),
                "tangents": f(tv.tangents), "bitangents": f(tv.bitangents),
                "indices": view_array(tv.indices, n, np.uint32, (n,))}

    def mapuv(self):
        npx = self.width * self.height
        outs = [np.empty((npx, 3), np.float32) for _ in range(5)]
        check(lib().fornos_texels_export_mapuv(self._h, *[ptr(o) for o in outs]))
        return dict(zip(["positions", "directions", "normals", "tangents", "bitangents"], outs))


class TexelArrays:
    """A texel set held in numpy arrays (the layout of CompressedMapUV,
    Src/compute.h:245-259): positions, directions, normals, tangents,
    bitangents [n, 3] float32 and indices [n] uint32 into a width x height
    map.  Used for sub-sets of a map (windows, tiles)."""

    def __init__(self, arrays, width, height):
        self.a = {k: np.ascontiguousarray(arrays[k], np.float32)
                  for k in ("positions", "directions", "normals", "tangents", "bitangents")}
        self.a["indices"] = np.ascontiguousarray(arrays["indices"], np.uint32)
        self.width, self.height = int(width), int(height)

    @property
    def count(self):
        return len(self.a["indices"])

    def texel_view(self):
        a = self.a
        return _abi.TexelView(self.count, self.width, self.height, ptr(a["positions"]), ptr(a["directions"]),
                              ptr(a["normals"]), ptr(a["tangents"]), ptr(a["bitangents"]), ptr(a["indices"]))

    def arrays(self):
        return dict(self.a)

    def subset(self, select, indices=None, width=None, height=None):
        """The texels where `select` holds, optionally re-indexed into another map."""
        d = {k: v[select] for k, v in self.a.items()}
        if indices is not None:
            d["indices"] = indices
        return TexelArrays(d, width or self.width, height or self.height)


def sample_table(sample_count, n_perm=SAMPLE_PERM_COUNT, seed=1234):
    """(table[n_perm*S, 3] float32, rand[n_perm, 2] uint32), Src/math.h:363-382."""
    t = np.empty((n_perm * sample_count, 3), np.float32)
    r = np.empty((n_perm, 2), np.uint32)
    check(lib().fornos_sample_table(int(sample_count), int(n_perm), int(seed), ptr(t), ptr(r)))
    return t, r


@functools.lru_cache(maxsize=4)
def _shared_table(sample_count, seed):
    """The seeded table of a sample count, computed once for the solvers that
    use it (read-only)."""
    t, _ = sample_table(sample_count, SAMPLE_PERM_COUNT, seed)
    t.setflags(write=False)
    return t


class Scene(_Handle):
    """Hi-poly mesh + BVH resident in HBM (BVH::createBinary + fillMeshData)."""
    _destroy = "fornos_scene_destroy"

    def __init__(self, mesh, bvh_tris_per_node=8, max_depth=8192):
        super().__init__()
        self.mesh = mesh
        bp = _abi.BVHParams(_abi.BVH_REFERENCE, int(bvh_tris_per_node), int(max_depth))
        mv = mesh.mesh_view()
        check(lib().fornos_scene_create(C.byref(mv), C.byref(bp), C.byref(self._h)))

    def info(self):
        i = _abi.SceneInfo()
        check(lib().fornos_scene_info_get(self._h, C.byref(i)))
        return {k: getattr(i, k) for k, _ in i._fields_}

    def reference_nodes(self):
        """(nodes[n,9] uint32 in BVHGPUData layout, mesh triangle id per soup triangle)."""
        i = self.info()
        nodes = np.empty((i["n_nodes"], 9), np.uint32)
        ids = np.empty(i["n_triangles"], np.uint32)
        check(lib().fornos_scene_export_reference_nodes(self._h, ptr(nodes), ptr(ids)))
        return nodes, ids

    def soup(self):
        i = self.info()
        p = np.empty((3 * i["n_triangles"], 4), np.float32)
        n = np.empty((3 * i["n_triangles"], 4), np.float32)
        check(lib().fornos_scene_export_soup(self._h, ptr(p), ptr(n)))
        return p, n


class MeshMapping(_Handle):
    """MeshMapping (Src/meshmapping.h:42-75)."""
    _destroy = "fornos_meshmap_destroy"

    def init(self, texels, scene, cullBackfaces=False):
        self.close()
        self.texels, self.scene = texels, scene
        tv = texels.texel_view()
        check(lib().fornos_meshmap_create(scene.handle, C.byref(tv), int(bool(cullBackfaces)), C.byref(self._h)))
        return self

    def init_from_mesh(self, lo, dirs, width, height, scene, cullBackfaces=False, edge=0.0):
        """MeshMapping::init with the texel set rasterised on the device;
        edge > 0: the Hybrid mapping method (MapUV::fromMeshes_Hybrid)."""
        self.close()
        self.texels, self.scene = None, scene
        lv = lo.mesh_view()
        dv = dirs.mesh_view() if dirs is not None else None
        check(lib().fornos_meshmap_create_from_mesh_hybrid(scene.handle, C.byref(lv), C.byref(dv) if dv else None,
                                                           int(width), int(height), float(edge),
                                                           int(bool(cullBackfaces)), C.byref(self._h)))
        return self

    def read_texels(self):
        n = self.count
        outs = [np.empty((n, 3), np.float32) for _ in range(5)]
        idx = np.empty(n, np.uint32)
        check(lib().fornos_meshmap_read_texels(self._h, *[ptr(o) for o in outs], ptr(idx)))
        d = dict(zip(["positions", "directions", "normals", "tangents", "bitangents"], outs))
        d["indices"] = idx
        return d

    def runStep(self):
        done = C.c_int()
        check(lib().fornos_meshmap_run_step(self._h, C.byref(done)))
        return bool(done.value)

    def run_all(self):
        ms = C.c_float()
        check(lib().fornos_meshmap_run_all(self._h, C.byref(ms)))
        return ms.value

    def progress(self):
        p = C.c_float()
        check(lib().fornos_meshmap_progress(self._h, C.byref(p)))
        return p.value

    def reset(self):
        check(lib().fornos_meshmap_reset(self._h))

    PREDICATE_REFERENCE, PREDICATE_WATERTIGHT = 0, 1

    def set_predicate(self, predicate):
        """PREDICATE_REFERENCE (default, the reference's triangle test) or
        PREDICATE_WATERTIGHT (fornos_meshmap_set_predicate)."""
        check(lib().fornos_meshmap_set_predicate(self._h, int(predicate)))

    TRAVERSAL_WIDE, TRAVERSAL_REFERENCE = 0, 1

    def set_traversal(self, traversal, stack=0):
        """TRAVERSAL_WIDE (default) or TRAVERSAL_REFERENCE: the tree the
        reference predicate's passes walk; `stack`: entries of the wide
        traversal's stack (3..24, 0 = 24).  Results are identical
        (fornos_meshmap_set_traversal)."""
        check(lib().fornos_meshmap_set_traversal(self._h, int(traversal), int(stack)))

    def redo_counts(self):
        """{"undecided": passes redone on the reference tree for a hit whose
        reachability the wide traversal could not decide, "overflow": for a
        full stack}, since creation or the last reset."""
        u, o = C.c_uint32(), C.c_uint32()
        check(lib().fornos_meshmap_redo_counts(self._h, C.byref(u), C.byref(o)))
        return {"undecided": u.value, "overflow": o.value}

    @property
    def count(self):
        n = C.c_uint32()
        check(lib().fornos_meshmap_texel_count(self._h, C.byref(n)))
        return n.value

    def read(self):
        """(coords[n,4] = t, b0, b1, b2; tidx[n] soup vertex index or 0xFFFFFFFF)."""
        n = self.count
        coords = np.empty((n, 4), np.float32)
        tidx = np.empty(n, np.uint32)
        check(lib().fornos_meshmap_read(self._h, ptr(coords), ptr(tidx)))
        return coords, tidx


class Solver(_Handle):
    """One of the reference solvers (Src/solver_*.h), selected by kind."""
    _destroy = "fornos_solver_destroy"
    kind = None
    comps = 1

    def __init__(self, sample_count=0, min_distance=0.0, max_distance=0.0, tangent_space=False,
                 normalize_output=False, seed=1234, table=None):
        super().__init__()
        self.params = _abi.SolverParams(int(sample_count), float(min_distance), float(max_distance),
                                        int(bool(tangent_space)), int(bool(normalize_output)))
        self.seed = seed
        self.table = table

    def init(self, texels, meshMapping):
        self.close()
        self.meshMapping = meshMapping
        tab, n_perm = None, 0
        if self.kind >= _abi.SOLVER_AO:
            if self.table is None:
                self.table = _shared_table(self.params.sample_count, self.seed)
            tab = np.ascontiguousarray(self.table, np.float32)
            n_perm = len(tab) // self.params.sample_count
        check(lib().fornos_solver_create(self.kind, C.byref(self.params), meshMapping.handle,
                                         ptr(tab), n_perm, C.byref(self._h)))
        return self

    def runStep(self):
        done = C.c_int()
        check(lib().fornos_solver_run_step(self._h, C.byref(done)))
        return bool(done.value)

    def run_all(self):
        ms = C.c_float()
        check(lib().fornos_solver_run_all(self._h, C.byref(ms)))
        return ms.value

    def progress(self):
        p = C.c_float()
        check(lib().fornos_solver_progress(self._h, C.byref(p)))
        return p.value

    def run_texels(self, texels):
        """The pass for the listed texels only (distinct global indices, in
        the order their rays are handed out); their results equal run_all's
        (fornos_solver_run_texels)."""
        t = np.ascontiguousarray(texels, np.uint32)
        ms = C.c_float()
        check(lib().fornos_solver_run_texels(self._h, ptr(t), len(t), C.byref(ms)))
        return ms.value

    def reset(self):
        check(lib().fornos_solver_reset(self._h))

    def getResults(self, out=None):
        """The solver's results (one float or three per texel); `out`, when
        given, is the float32 host array to fill (page-locked memory makes
        the read a plain DMA)."""
        n = self.meshMapping.count
        shape = (n,) if self.comps == 1 else (n, self.comps)
        if out is None:
            out = np.empty(shape, np.float32)
        elif out.shape != shape or out.dtype != np.float32 or not out.flags["C_CONTIGUOUS"]:
            raise ValueError("getResults: out must be a C-contiguous float32 array of shape %s" % (shape,))
        check(lib().fornos_solver_read(self._h, ptr(out), out.nbytes))
        return out

    def counts(self):
        out = np.empty(self.meshMapping.count, np.uint32)
        check(lib().fornos_solver_read_counts(self._h, ptr(out)))
        return out

    def last_timing(self):
        """(k_rays ms, k_reduce ms, k_rays launches) of the last run_all."""
        a, b, n = C.c_float(), C.c_float(), C.c_uint32()
        check(lib().fornos_solver_last_timing(self._h, C.byref(a), C.byref(b), C.byref(n)))
        return a.value, b.value, n.value

    def count_work(self):
        a, b, c, d = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_uint64()
        check(lib().fornos_solver_count_work(self._h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        out = np.zeros(12, np.uint64)
        check(lib().fornos_solver_work_counters(self._h, ptr(out), 12))
        w = {"node_tests": a.value, "tri_tests": b.value, "rays": c.value, "stack_overflows": d.value}
        for k, v in zip(["walk_steps", "walk_lanes", "leaf_steps", "leaf_lanes", "refills", "refill_lanes",
                         "redo_rays", "slow_rays"], out[4:]):
            w[k] = int(v)
        w["reference_tree_rays"] = w["redo_rays"] + w["slow_rays"]
        return w

    FB_OVERFLOW, FB_REDO, FB_SLOW = 1, 2, 4

    def fallback_flags(self):
        """Per-texel flags of the last count_work run: FB_OVERFLOW (a ray was
        redone on the stackless tree), FB_REDO (a thickness ray was redone on
        the reference tree), FB_SLOW (a ray was traced on the reference tree
        for a direction component below 2^-40)."""
        out = np.empty(self.meshMapping.count, np.uint32)
        check(lib().fornos_solver_fallback_flags(self._h, ptr(out)))
        return out

    def set_launch(self, blocks_per_cu=8, refill=32, stack=16):
        check(lib().fornos_solver_set_launch(self._h, int(blocks_per_cu), int(refill), int(stack)))

    def get_launch(self):
        a, b, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(lib().fornos_solver_get_launch(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def set_tile(self, texels=0):
        """Texels a wave takes from the pass's counter at once (1..32; 0:
        256 / sampleCount, the default) (fornos_solver_set_tile).  Results do
        not depend on it."""
        check(lib().fornos_solver_set_tile(self._h, int(texels)))

    def get_tile(self):
        """The hand-out tile in effect (the default resolved)."""
        t = C.c_uint32()
        check(lib().fornos_solver_get_tile(self._h, C.byref(t)))
        return t.value

    def autotune(self, fraction=0.05):
        """Times launch shapes and hand-out tiles on the first `fraction` of
        the pass's texels, then the texel and Morton ray orders on the same
        texels, and keeps the fastest (a current setting gives way only to
        one at least 1 % faster); returns (blocks_per_cu, refill, stack,
        ms) (fornos_solver_autotune; get_tile and get_ray_order give the
        rest).  Results never depend on any of these."""
        a, b, c, ms = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_float()
        check(lib().fornos_solver_autotune(self._h, float(fraction), C.byref(a), C.byref(b), C.byref(c), C.byref(ms)))
        return a.value, b.value, c.value, ms.value

    RAY_ORDER_TEXELS, RAY_ORDER_MORTON, RAY_ORDER_AUTO = 0, 1, 2

    def set_ray_order(self, order):
        """RAY_ORDER_MORTON: texels' rays handed out by the Morton code of
        their hit positions; RAY_ORDER_TEXELS: the texel set's own order;
        RAY_ORDER_AUTO (default): Morton for rays that cross the scene
        (fornos_solver_set_ray_order).  Results do not depend on it."""
        check(lib().fornos_solver_set_ray_order(self._h, int(order)))

    def get_ray_order(self):
        t = C.c_int()
        check(lib().fornos_solver_get_ray_order(self._h, C.byref(t)))
        return t.value

    def effective_ray_order(self):
        """RAY_ORDER_TEXELS or RAY_ORDER_MORTON: what the next pass uses."""
        t = C.c_int()
        check(lib().fornos_solver_effective_ray_order(self._h, C.byref(t)))
        return t.value


class HeightSolver(Solver):
    kind = _abi.SOLVER_HEIGHT

    def __init__(self, normalizeOutput=True, maxDistance=0.0):
        super().__init__(max_distance=maxDistance, normalize_output=normalizeOutput)


class PositionSolver(Solver):
    kind, comps = _abi.SOLVER_POSITION, 3


class NormalsSolver(Solver):
    kind, comps = _abi.SOLVER_NORMALS, 3

    def __init__(self, tangentSpace=True):
        super().__init__(tangent_space=tangentSpace)


class AmbientOcclusionSolver(Solver):
    kind = _abi.SOLVER_AO

    def __init__(self, sampleCount=256, minDistance=0.01, maxDistance=10.0, seed=1234, table=None):
        super().__init__(sampleCount, minDistance, maxDistance, seed=seed, table=table)

    def pair(self, bentNormals):
        """Traces each ray once for this solver and a BentNormalsSolver with the
        same rays (fornos_solver_pair); None unpairs.  Results are unchanged."""
        check(lib().fornos_solver_pair(self._h, bentNormals.handle if bentNormals is not None else None))
        self._partner = bentNormals  # the native pair holds no reference
        return self


def shares_rays(ao_params, bn_params):
    """Whether AO and bent-normal solver keyword sets cast the same rays
    (sample count, distances, sample seed / table), so they can be paired."""
    f32 = lambda p, k, d: np.float32(p.get(k, d)).view(np.uint32)  # as the native side compares them
    same = (int(ao_params.get("sampleCount", 256)) == int(bn_params.get("sampleCount", 256)) and
            int(ao_params.get("seed", 1234)) == int(bn_params.get("seed", 1234)) and
            f32(ao_params, "minDistance", 0.01) == f32(bn_params, "minDistance", 0.01) and
            f32(ao_params, "maxDistance", 10.0) == f32(bn_params, "maxDistance", 10.0))
    return same and ao_params.get("table") is None and bn_params.get("table") is None


class BentNormalsSolver(Solver):
    kind, comps = _abi.SOLVER_BENT_NORMALS, 3

    def __init__(self, sampleCount=256, minDistance=0.01, maxDistance=10.0, tangentSpace=True,
                 seed=1234, table=None):
        super().__init__(sampleCount, minDistance, maxDistance, tangent_space=tangentSpace,
                         seed=seed, table=table)


class ThicknessSolver(Solver):
    kind = _abi.SOLVER_THICKNESS

    def __init__(self, sampleCount=256, minDistance=0.01, maxDistance=10.0, seed=1234, table=None):
        super().__init__(sampleCount, minDistance, maxDistance, seed=seed, table=table)


class Batch(_Handle):
    """Several AO solvers run as one persistent launch over a device job table,
    captured in a HIP graph (include/fornos_hip.h, fornos_batch_*).  No
    reference counterpart: Src/fornos.cpp:239-251 runs one solver step per
    frame.  Each solver's results equal its own run_all bit for bit and are
    read with its getResults()."""
    _destroy = "fornos_batch_destroy"

    def __init__(self, solvers):
        super().__init__()
        self.solvers = list(solvers)  # for reading results; the native batch holds its own references
        arr = (C.c_void_p * len(self.solvers))(*[s.handle.value for s in self.solvers])
        check(lib().fornos_batch_create(arr, len(self.solvers), C.byref(self._h)))

    def set_launch(self, blocks_per_cu, refill, stack=16):
        check(lib().fornos_batch_set_launch(self._h, int(blocks_per_cu), int(refill), int(stack)))

    def get_launch(self):
        a, b, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(lib().fornos_batch_get_launch(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def autotune(self, fraction=0.05):
        """Times launch shapes and the jobs' hand-out tiles (halved, kept,
        doubled) on a fraction of each job's texels, keeps the fastest (a
        current setting gives way only to one at least 1 % faster); returns
        (blocks_per_cu, refill, stack, ms)."""
        a, b, c, ms = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_float()
        check(lib().fornos_batch_autotune(self._h, float(fraction), C.byref(a), C.byref(b), C.byref(c),
                                          C.byref(ms)))
        return a.value, b.value, c.value, ms.value

    def set_tile_shift(self, shift):
        """Jobs' hand-out tiles: their solvers' halved (-1), kept (0) or
        doubled (1).  Results do not depend on it."""
        check(lib().fornos_batch_set_tile_shift(self._h, int(shift)))

    def get_tile_shift(self):
        t = C.c_int()
        check(lib().fornos_batch_get_tile_shift(self._h, C.byref(t)))
        return t.value

    def get_tile(self, solver):
        """The hand-out tile a job (one of this batch's solvers) runs with."""
        t = C.c_uint32()
        check(lib().fornos_batch_get_tile(self._h, solver.handle, C.byref(t)))
        return t.value

    def run(self):
        ms = C.c_float()
        check(lib().fornos_batch_run(self._h, C.byref(ms)))
        return ms.value


def device_count():
    return lib().fornos_device_count()
