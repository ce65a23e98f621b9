"""FornosParameters and FornosRunner (Src/fornos.h:35-149, Src/fornos.cpp:53-251)
over the native runner of include/fornos_runner.h.

    p = FornosParameters()
    p.shared.loPolyMeshPath, p.shared.hiPolyMeshPath = "low.obj", "high.ply"
    p.ao.enabled, p.ao.outputPath = True, "ao.png"
    r = FornosRunner()
    ok, errors = r.start(p)
    while r.pending():
        r.run()                      # never waits on the device
        t = r.currentTask()          # name() / progress() of the running task

Field names and defaults are the reference's.  run() submits a task's whole
work on its first call and polls it afterwards, instead of one runStep per
call; run_all() drives every task to the end from a headless caller.
"""
import ctypes as C
from dataclasses import dataclass, field

from . import _abi
from ._abi import check, lib


class NormalImport:
    """Src/fornos.h:35"""
    Import, ComputePerFace, ComputePerVertex = _abi.NORMALS_IMPORT, _abi.NORMALS_PER_FACE, _abi.NORMALS_PER_VERTEX


class MeshMappingMethod:
    """Src/fornos.h:36"""
    Smooth, LowPolyNormals, Hybrid = _abi.MAPPING_SMOOTH, _abi.MAPPING_LOWPOLY_NORMALS, _abi.MAPPING_HYBRID


@dataclass
class FornosParameters_Shared:
    """Src/fornos.h:38-51; sampleSeed seeds the hemisphere sample tables (the
    reference draws them from std::random_device)."""
    loPolyMeshPath: str = ""
    hiPolyMeshPath: str = ""
    loPolyMeshNormal: int = NormalImport.Import
    hiPolyMeshNormal: int = NormalImport.Import
    bvhTrisPerNode: int = 8
    texWidth: int = 2048
    texHeight: int = 2048
    texDilation: int = 16
    ignoreBackfaces: bool = True
    mapping: int = MeshMappingMethod.Smooth
    mappingEdge: float = 0.05
    sampleSeed: int = 1234


@dataclass
class FornosParameters_SolverHeight:
    """Src/fornos.h:53-61"""
    enabled: bool = False
    outputPath: str = ""
    normalizeOutput: bool = True
    maxDistance: float = 0.0


@dataclass
class FornosParameters_SolverPositions:
    """Src/fornos.h:63-69"""
    enabled: bool = False
    outputPath: str = ""


@dataclass
class FornosParameters_SolverNormals:
    """Src/fornos.h:71-78"""
    enabled: bool = False
    tangentSpace: bool = True
    outputPath: str = ""


@dataclass
class FornosParameters_SolverAO:
    """Src/fornos.h:80-89"""
    enabled: bool = False
    sampleCount: int = 256
    minDistance: float = 0.01
    maxDistance: float = 10.0
    outputPath: str = ""


@dataclass
class FornosParameters_SolverBentNormals:
    """Src/fornos.h:91-101"""
    enabled: bool = False
    sampleCount: int = 256
    minDistance: float = 0.01
    maxDistance: float = 10.0
    tangentSpace: bool = True
    outputPath: str = ""


@dataclass
class FornosParameters_SolverThickness:
    """Src/fornos.h:103-112"""
    enabled: bool = False
    sampleCount: int = 256
    minDistance: float = 0.01
    maxDistance: float = 10.0
    outputPath: str = ""


@dataclass
class FornosParameters:
    """Src/fornos.h:114-123"""
    shared: FornosParameters_Shared = field(default_factory=FornosParameters_Shared)
    height: FornosParameters_SolverHeight = field(default_factory=FornosParameters_SolverHeight)
    positions: FornosParameters_SolverPositions = field(default_factory=FornosParameters_SolverPositions)
    normals: FornosParameters_SolverNormals = field(default_factory=FornosParameters_SolverNormals)
    ao: FornosParameters_SolverAO = field(default_factory=FornosParameters_SolverAO)
    bentNormals: FornosParameters_SolverBentNormals = field(default_factory=FornosParameters_SolverBentNormals)
    thickness: FornosParameters_SolverThickness = field(default_factory=FornosParameters_SolverThickness)

    def to_c(self):
        """The fornos_parameters struct (its strings are kept alive on it)."""
        c = _abi.Parameters()
        lib().fornos_parameters_defaults(C.byref(c))
        keep = []

        def s(v):
            b = str(v or "").encode()
            keep.append(b)
            return b

        sh = self.shared
        c.shared.lo_poly_mesh_path = s(sh.loPolyMeshPath)
        c.shared.hi_poly_mesh_path = s(sh.hiPolyMeshPath)
        c.shared.lo_poly_mesh_normal = int(sh.loPolyMeshNormal)
        c.shared.hi_poly_mesh_normal = int(sh.hiPolyMeshNormal)
        c.shared.bvh_tris_per_node = int(sh.bvhTrisPerNode)
        c.shared.tex_width = int(sh.texWidth)
        c.shared.tex_height = int(sh.texHeight)
        c.shared.tex_dilation = int(sh.texDilation)
        c.shared.ignore_backfaces = int(bool(sh.ignoreBackfaces))
        c.shared.mapping = int(sh.mapping)
        c.shared.mapping_edge = float(sh.mappingEdge)
        c.shared.sample_seed = int(sh.sampleSeed)
        for name, cname in (("height", "height"), ("positions", "positions"), ("normals", "normals"),
                            ("ao", "ao"), ("bentNormals", "bent_normals"), ("thickness", "thickness")):
            src, dst = getattr(self, name), getattr(c, cname)
            dst.enabled = int(bool(src.enabled))
            dst.output_path = s(src.outputPath)
            for a, b in (("sampleCount", "sample_count"), ("minDistance", "min_distance"),
                         ("maxDistance", "max_distance")):
                if hasattr(src, a):
                    setattr(dst, b, type(getattr(dst, b))(getattr(src, a)))
            if hasattr(src, "tangentSpace"):
                dst.tangent_space = int(bool(src.tangentSpace))
            if hasattr(src, "normalizeOutput"):
                dst.normalize_output = int(bool(src.normalizeOutput))
        c._keep = keep
        return c


class FornosTask:
    """currentTask() of the runner (FornosTask, Src/fornos.h:129-137)."""

    def __init__(self, name, progress):
        self._name, self._progress = name, progress

    def name(self):
        return self._name

    def progress(self):
        return self._progress


class FornosRunner:
    """Src/fornos.h:139-149 over fornos_runner_*."""

    def __init__(self):
        self._h = C.c_void_p()
        check(lib().fornos_runner_create(C.byref(self._h)))

    def start(self, params):
        """(ok, errors) as bool start(const FornosParameters&, std::string& errors)."""
        self._c = params.to_c()
        st = lib().fornos_runner_start(self._h, C.byref(self._c))
        if st != _abi.OK:
            return False, lib().fornos_last_error().decode()
        return True, ""

    def pending(self):
        return bool(lib().fornos_runner_pending(self._h))

    def run(self):
        check(lib().fornos_runner_run(self._h))

    def run_all(self):
        check(lib().fornos_runner_run_all(self._h))

    def currentTask(self):
        name, prog = C.c_char_p(), C.c_float()
        check(lib().fornos_runner_current_task(self._h, C.byref(name), C.byref(prog)))
        if not name.value:
            return None
        return FornosTask(name.value.decode(), prog.value)

    def exported_range(self, kind):
        """(min, max) a finished float-map task (AO, thickness, height) exported, or None."""
        import numpy as np
        mm = np.zeros(2, np.float32)
        if lib().fornos_runner_exported_range(self._h, int(kind), mm.ctypes.data_as(C.c_void_p)):
            return float(mm[0]), float(mm[1])
        return None

    def close(self):
        if self._h:
            lib().fornos_runner_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
