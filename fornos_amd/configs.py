"""Bake workloads and the host preparation a bake runs before the GPU.

prepare() restates FornosRunner::start's assembly order (Src/fornos.cpp:53-236):
the low-poly mesh (a seeded generator or an .obj / .ply file) and its normal
import mode, the hi-poly mesh (or the low-poly one) and its mode, the
tangent frame when a tangent-space map is baked, the mapping mesh of the
method (Smooth and Hybrid: a copy of the low-poly mesh with aggressive
vertex normals unless the low-poly normals are computed per vertex;
LowPolyNormals: none), MapUV -> CompressedMapUV, then the hi-poly BVH.

Workload keys beyond the meshes, the map size and the maps, with the
reference defaults (Src/fornos.h:38-123): "lowNormals" / "highNormals"
("import", "face", "vertex"; import), "mapping" ("smooth", "lowpoly",
"hybrid"; smooth), "mappingEdge" (0.05), "width" / "height" (else "size"),
"bvhTrisPerNode" (8), "cull" (ignoreBackfaces, true).
"""
import os

from . import _abi
from .api import CompressedMapUV, Mesh, Scene

# name -> workload description
CONFIGS = {
    # Plain method, small (reference-runnable on CPU).
    "sphere512": {
        "low": (_abi.GEN_UV_SPHERE, 0, 1.0, 32, 16),
        "high": (_abi.GEN_ICOSPHERE_WAVES, 7, 0.02, 6, 0),
        "size": 512,
        "cull": True,
        "maps": {"height": {}, "normals": {"tangentSpace": False},
                 "ao": {"sampleCount": 32, "minDistance": 0.0, "maxDistance": 0.5}},
    },
    # Plain method, production size: the AO-256 rays/s workload.
    "heightfield2048": {
        "low": (_abi.GEN_HEIGHTFIELD_LOW, 11, 0.05, 64, 0),
        "high": (_abi.GEN_HEIGHTFIELD, 11, 0.05, 1024, 200),
        "size": 2048,
        "cull": True,
        "maps": {"normals": {"tangentSpace": True}, "position": {},
                 "ao": {"sampleCount": 256, "minDistance": 0.01, "maxDistance": 0.2},
                 "bentnormals": {"sampleCount": 256, "minDistance": 0.01, "maxDistance": 0.2,
                                 "tangentSpace": True},
                 "thickness": {"sampleCount": 128, "minDistance": 0.0, "maxDistance": 0.1}},
    },
    # Large, with overhangs: 8.4M-triangle torus-knot tube.
    "knot4096": {
        "low": (_abi.GEN_TORUS_KNOT, 0, 0.0, 128, 16),
        "high": (_abi.GEN_TORUS_KNOT, 23, 0.03, 2048, 2048),
        "size": 4096,
        "cull": True,
        "maps": {"ao": {"sampleCount": 512, "minDistance": 0.01, "maxDistance": 10.0},
                 "thickness": {"sampleCount": 256, "minDistance": 0.01, "maxDistance": 10.0}},
    },
}


def scaled(name, size=None, high_quads=None):
    """A copy of a workload at a smaller map size / hi-poly resolution (tests)."""
    import copy
    c = copy.deepcopy(CONFIGS[name])
    if size is not None:
        c["size"] = size
    if high_quads is not None:
        g = list(c["high"])
        g[3] = high_quads
        if g[0] == _abi.GEN_TORUS_KNOT:
            g[4] = high_quads
        c["high"] = tuple(g)
    return c


def make_mesh(spec):
    """A generator tuple (generator, seed, a, b, c) or a mesh file path."""
    if isinstance(spec, (str, os.PathLike)):
        return Mesh.load(spec)
    gen, seed, a, b, c = spec
    return Mesh.generate(gen, seed, a, b, c)


NORMAL_MODES = ("import", "face", "vertex")
MAPPINGS = ("smooth", "lowpoly", "hybrid")


def apply_normals(mesh, mode):
    """NormalImport (Src/fornos.cpp:59-64)."""
    if mode not in NORMAL_MODES:
        raise ValueError("normal import mode must be one of %s, not %r" % (NORMAL_MODES, mode))
    if mode == "face":
        mesh.computeFaceNormals()
    elif mode == "vertex":
        mesh.computeVertexNormals()


def tex_size(cfg):
    return int(cfg.get("width", cfg["size"])), int(cfg.get("height", cfg["size"]))


def needs_tangent_space(cfg):
    m = cfg["maps"]
    return (m.get("normals", {}).get("tangentSpace", False) or
            m.get("bentnormals", {}).get("tangentSpace", False))


def prepare(cfg, with_scene=True):
    """Meshes, texel set and (optionally) the device scene of a workload."""
    if isinstance(cfg, str):
        cfg = CONFIGS[cfg]
    lo_mode = cfg.get("lowNormals", "import")
    lo = make_mesh(cfg["low"])
    apply_normals(lo, lo_mode)
    if cfg.get("high") is None:
        hi = lo
    else:
        hi = make_mesh(cfg["high"])
        apply_normals(hi, cfg.get("highNormals", "import"))
    if needs_tangent_space(cfg):
        lo.computeTangentSpace()
    mapping = cfg.get("mapping", "smooth")
    if mapping not in MAPPINGS:
        raise ValueError("mapping must be one of %s, not %r" % (MAPPINGS, mapping))
    edge = 0.0
    if mapping == "hybrid":
        edge = float(cfg.get("mappingEdge", 0.05))
        if not edge > 0:
            raise ValueError("the Hybrid mapping needs mappingEdge > 0")
    if mapping == "lowpoly":
        dirs = None
    elif lo_mode == "vertex":
        dirs = lo
    else:
        dirs = lo.copy()
        dirs.computeVertexNormalsAggressive()
    W, H = tex_size(cfg)
    texels = CompressedMapUV(lo, dirs, W, H, edge=edge)
    out = {"cfg": cfg, "low": lo, "dirs": dirs, "high": hi, "texels": texels, "width": W, "height": H,
           "edge": edge}
    if with_scene:
        out["scene"] = Scene(hi, bvh_tris_per_node=int(cfg.get("bvhTrisPerNode", 8)))
    return out
