"""Seeded bake workloads and the host preparation a bake runs before the GPU.

prepare() restates FornosRunner::start's assembly order (Src/fornos.cpp:53-236)
for the procedural meshes: low-poly normals as generated ("Import"), the
tangent frame when a tangent-space map is baked, the Smooth mapping method
(a copy of the low-poly mesh with aggressive vertex normals gives the ray
directions), MapUV -> CompressedMapUV, then the hi-poly BVH.  Parameters
that a workload does not set take the reference defaults (Src/fornos.h:38-123):
minDistance 0.01, maxDistance 10, bvhTrisPerNode 8, ignoreBackfaces true.
"""
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
    gen, seed, a, b, c = spec
    return Mesh.generate(gen, seed, a, b, c)


def needs_tangent_space(cfg):
    m = cfg["maps"]
    return (m.get("normals", {}).get("tangentSpace", False) or
            m.get("bentnormals", {}).get("tangentSpace", False))


def prepare(cfg, with_scene=True):
    """Meshes, texel set and (optionally) the device scene of a workload."""
    if isinstance(cfg, str):
        cfg = CONFIGS[cfg]
    lo = make_mesh(cfg["low"])
    hi = make_mesh(cfg["high"])
    if needs_tangent_space(cfg):
        lo.computeTangentSpace()
    dirs = lo.copy()
    dirs.computeVertexNormalsAggressive()
    texels = CompressedMapUV(lo, dirs, cfg["size"], cfg["size"])
    out = {"cfg": cfg, "low": lo, "dirs": dirs, "high": hi, "texels": texels}
    if with_scene:
        out["scene"] = Scene(hi)
    return out
