"""Shared pieces of the GPU parity tests: a prepared workload with the
oracle's own BVH beside it, bitwise float comparison, and the oracle check of
hemisphere results over chosen texels."""
import numpy as np

import fornos_amd as fa
import oracle as orc

MISS = 0xFFFFFFFF

SOLVERS = {"ao": (fa.AmbientOcclusionSolver, orc.KIND_AO), "bentnormals": (fa.BentNormalsSolver, orc.KIND_BN),
           "thickness": (fa.ThicknessSolver, orc.KIND_THICK)}


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def same_bits(a, b):
    """Bitwise float equality, with every NaN equal to every NaN."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    nan = np.isnan(a) & np.isnan(b)
    return np.array_equal(bits(a)[~nan], bits(b)[~nan]) and np.array_equal(np.isnan(a), np.isnan(b))


class Bake:
    """A prepared workload (device scene, texel set, mapping) with the
    oracle's BVH and flatten of the same hi-poly mesh."""

    def __init__(self, cfg, cull=None):
        self._setup(fa.prepare(cfg), cull)

    @classmethod
    def from_meshes(cls, lo, hi, size, cull=True, maps=None):
        """A workload from two meshes: texel set of `lo` (its own normals as
        ray directions) on a size x size map, `hi` as the scene."""
        b = cls.__new__(cls)
        texels = fa.CompressedMapUV(lo, None, size, size)
        prep = {"cfg": {"maps": maps or {}, "cull": cull, "size": size}, "low": lo, "high": hi,
                "texels": texels, "scene": fa.Scene(hi)}
        b._setup(prep, cull)
        return b

    def _setup(self, prep, cull):
        self.prep = prep
        self.cfg = prep["cfg"]
        self.cull = self.cfg["cull"] if cull is None else cull
        self.tex = prep["texels"].arrays()
        self.scene = prep["scene"]
        hi = prep["high"].arrays()
        self.nodes, self.spos, self.snrm, self.stri = orc.bvh_build(hi)
        self.mm = fa.MeshMapping().init(prep["texels"], self.scene, self.cull)
        self.ms = self.mm.run_all()
        self.coords, self.tidx = self.mm.read()

    def oracle_map(self, begin=0, end=None, cull=None):
        cull = self.cull if cull is None else cull
        return orc.meshmapping(self.tex["positions"], self.tex["directions"], self.spos, self.nodes, cull,
                               begin=begin, end=end)

    def solver(self, kind, **over):
        """The workload's solver of `kind` ("ao", "bentnormals", "thickness"),
        parameters overridden by `over`."""
        p = dict(self.cfg["maps"].get(kind, {}))
        p.update(over)
        cls = SOLVERS[kind][0]
        kw = {"tangentSpace": p.get("tangentSpace", False)} if kind == "bentnormals" else {}
        if "table" in p:
            kw["table"] = p["table"]
        return cls(p["sampleCount"], p["minDistance"], p["maxDistance"], **kw).init(self.prep["texels"], self.mm)

    def oracle_hemi(self, kind, solver, lo, hi):
        """Oracle (values, per-texel counts) of texels [lo, hi) for a solver's
        parameters and table, tangent space applied as the solver does."""
        p = solver.params
        k = SOLVERS[kind][1]
        out, oc = orc.hemisphere(k, p.sample_count, p.min_distance, p.max_distance, solver.table, self.coords,
                                 self.tidx, self.spos, self.snrm, self.nodes, lo, hi)
        if k == orc.KIND_BN and p.tangent_space:
            hit = self.tidx[lo:hi] != MISS
            out[hit] = orc.tangentspace(self.tex["normals"], self.tex["tangents"], self.tex["bitangents"], out,
                                        begin=lo)[hit]
        return out, oc

    def check_texels(self, kind, solver, texels, res=None, cnt=None):
        """Compares the solver's results with the oracle on the given texel
        indices (grouped into contiguous runs); returns the texels checked."""
        res = solver.getResults() if res is None else res
        cnt = solver.counts() if cnt is None else cnt
        texels = np.unique(np.asarray(texels, np.int64))
        if len(texels) == 0:
            return 0
        breaks = np.flatnonzero(np.diff(texels) != 1) + 1
        for run in np.split(texels, breaks):
            lo, hi = int(run[0]), int(run[-1]) + 1
            out, oc = self.oracle_hemi(kind, solver, lo, hi)
            assert np.array_equal(cnt[lo:hi], oc), "%s: per-texel ray counts differ in [%d,%d)" % (kind, lo, hi)
            assert same_bits(res[lo:hi], out), "%s: values differ in [%d,%d)" % (kind, lo, hi)
        return len(texels)


def windows(n, k, width):
    """k windows of `width` texels spread evenly over [0, n)."""
    starts = np.linspace(0, max(0, n - width), k).astype(np.int64)
    return np.unique(np.concatenate([np.arange(s, min(n, s + width)) for s in starts]))
