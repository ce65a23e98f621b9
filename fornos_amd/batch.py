"""Batched AO bake jobs: 64 seeded jobs of uneven size run as one persistent
launch (fornos_batch_* in include/fornos_hip.h), against the same jobs run one
after another through each solver's own run_all.

A job draws, from its seed: a generator family (sphere, heightfield or knot,
the families of the plain workloads), a low-poly triangle count in [200, 20k]
and a high-poly count in [50k, 4M] (log-uniform, rounded to what the family's
grid can make), a map size from {256, 512, 1024, 2048} and an AO sample count
from {32, 64, 128, 256}.  `scale` shrinks map sizes and high-poly counts for
quick runs; scale 1 is the full workload.

The launch shape picked by the autotuner is cached in a JSON file keyed by the
device and the job table, so a repeated bake skips the search.
"""
import json
import math
import os

import numpy as np

from . import _abi
from .api import AmbientOcclusionSolver, Batch, Mesh, MeshMapping, Scene

SIZES = (256, 512, 1024, 2048)
SAMPLES = (32, 64, 128, 256)
FAMILIES = ("sphere", "heightfield", "knot")
LOW_TRIS = (200, 20_000)
HIGH_TRIS = (50_000, 4_000_000)
# AO distances per family: those of the plain workloads of the same family
MAX_DISTANCE = {"sphere": 0.5, "heightfield": 0.2, "knot": 10.0}


def _log_uniform(rng, lo, hi):
    return math.exp(rng.uniform(math.log(lo), math.log(hi)))


def job_spec(seed, scale=1.0):
    """The seeded description of one job (generator arguments, size, S)."""
    rng = np.random.default_rng(seed)
    fam = FAMILIES[int(rng.integers(len(FAMILIES)))]
    lo_t = _log_uniform(rng, *LOW_TRIS)
    hi_t = _log_uniform(rng, *HIGH_TRIS) * scale
    size = int(SIZES[int(rng.integers(len(SIZES)))] * math.sqrt(scale))
    size = max(64, size - size % 8)
    S = SAMPLES[int(rng.integers(len(SAMPLES)))]
    if fam == "sphere":
        # UV sphere: 2 * seg * (rings - 1) triangles, seg = 2 * rings
        rings = max(3, int(round(math.sqrt(lo_t / 4.0))))
        low = (_abi.GEN_UV_SPHERE, 0, 1.0, 2 * rings, rings)
        # icosphere: 20 * 4^k triangles
        k = max(1, min(8, int(round(math.log(max(hi_t, 20) / 20.0, 4)))))
        high = (_abi.GEN_ICOSPHERE_WAVES, 1000 + seed, 0.02, k, 0)
        cull = True
    elif fam == "heightfield":
        b = max(4, int(round(math.sqrt(lo_t / 2.0))))
        low = (_abi.GEN_HEIGHTFIELD_LOW, 1000 + seed, 0.05, b, 0)
        q = max(8, int(round(math.sqrt(hi_t / 2.0))))
        high = (_abi.GEN_HEIGHTFIELD, 1000 + seed, 0.05, q, int(200 * min(1.0, q / 1024.0)) + 8)
        cull = True
    else:
        # tube grid b x c with b = 8 c: 2 * b * c = 16 c^2 triangles
        c = max(3, int(round(math.sqrt(lo_t / 16.0))))
        low = (_abi.GEN_TORUS_KNOT, 0, 0.0, 8 * c, c)
        c = max(8, int(round(math.sqrt(hi_t / 2.0))))
        high = (_abi.GEN_TORUS_KNOT, 1000 + seed, 0.03, c, c)
        cull = True
    return {"seed": seed, "family": fam, "low": low, "high": high, "size": size, "cull": cull,
            "ao": {"sampleCount": S, "minDistance": 0.01, "maxDistance": MAX_DISTANCE[fam]}}


def job_specs(n_jobs=64, scale=1.0, first_seed=0):
    return [job_spec(first_seed + i, scale) for i in range(n_jobs)]


class Job:
    """A prepared job: device scene, device texel set (GPU raster), mapping
    result and an AO solver whose buffers the batch writes."""

    def __init__(self, spec):
        self.spec = spec
        lo = Mesh.generate(*spec["low"])
        hi = Mesh.generate(*spec["high"])
        dirs = lo.copy()
        dirs.computeVertexNormalsAggressive()
        self.scene = Scene(hi)
        self.mapping = MeshMapping().init_from_mesh(lo, dirs, spec["size"], spec["size"], self.scene,
                                                    spec["cull"])
        self.map_ms = self.mapping.run_all()
        ao = spec["ao"]
        self.solver = AmbientOcclusionSolver(ao["sampleCount"], ao["minDistance"], ao["maxDistance"])
        self.solver.init(None, self.mapping)
        self.high_tris = hi.n_triangles
        self.low_tris = lo.n_triangles
        _, tidx = self.mapping.read()
        self.live = int(np.count_nonzero(tidx != 0xFFFFFFFF))  # texels whose mapping ray hit

    @property
    def rays(self):
        return self.live * self.spec["ao"]["sampleCount"]


def _cache_key(jobs):
    import hashlib
    h = hashlib.sha1(json.dumps([j.spec for j in jobs], sort_keys=True).encode()).hexdigest()[:16]
    return f"{_abi.lib().fornos_device_count()}:{len(jobs)}:{h}"


def tune(batch, jobs, cache_path=None, fraction=0.05):
    """Launch shape and tile shift for `batch`: from the cache when present,
    else autotuned (and written back)."""
    key = _cache_key(jobs)
    cache = {}
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            cache = json.load(f)
    if key in cache:
        c = cache[key]
        batch.set_launch(c["blocks_per_cu"], c["refill"], c["stack"])
        batch.set_tile_shift(c.get("tile_shift", 0))
        return {"blocks_per_cu": c["blocks_per_cu"], "refill": c["refill"], "stack": c["stack"],
                "tile_shift": c.get("tile_shift", 0), "cached": True}
    bpc, refill, stack, ms = batch.autotune(fraction)
    shift = batch.get_tile_shift()
    if cache_path:
        cache[key] = {"blocks_per_cu": bpc, "refill": refill, "stack": stack, "tile_shift": shift, "sample_ms": ms}
        os.makedirs(os.path.dirname(os.path.abspath(cache_path)), exist_ok=True)
        with open(cache_path, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    return {"blocks_per_cu": bpc, "refill": refill, "stack": stack, "tile_shift": shift, "sample_ms": ms,
            "cached": False}


def main(argv=None):
    """Bakes the batched AO workload twice, as one graph launch and as one
    run_all per job, checks that every job's results are identical and prints
    one JSON line with both timings."""
    import argparse
    import sys
    import time

    ap = argparse.ArgumentParser(prog="python -m fornos_amd.batch")
    ap.add_argument("--jobs", type=int, default=64)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--first-seed", type=int, default=0)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--tune-fraction", type=float, default=0.05)
    ap.add_argument("--tune-cache", default=None)
    args = ap.parse_args(argv)

    t0 = time.perf_counter()
    jobs = []
    for i, spec in enumerate(job_specs(args.jobs, args.scale, args.first_seed)):
        jobs.append(Job(spec))
        j = jobs[-1]
        print("# job %d: %s low %d high %d tris, %d^2, S=%d, %d texels" % (
            i, spec["family"], j.low_tris, j.high_tris, spec["size"], spec["ao"]["sampleCount"],
            j.mapping.count), file=sys.stderr, flush=True)
    prep_s = time.perf_counter() - t0
    rays = sum(j.rays for j in jobs)

    # sequential: one run_all per job (each ends with a device synchronisation)
    seq_ms, ref = [], []
    for rep in range(args.steps + 1):
        t = time.perf_counter()
        for j in jobs:
            j.solver.reset()
            j.solver.run_all()
        if rep:
            seq_ms.append(1e3 * (time.perf_counter() - t))
    for j in jobs:
        ref.append(j.solver.getResults().view(np.uint32).copy())

    b = Batch([j.solver for j in jobs])
    tuned = tune(b, jobs, args.tune_cache, args.tune_fraction)
    b.run()
    bat_ms, graph_ms = [], []
    for _ in range(args.steps):
        t = time.perf_counter()
        graph_ms.append(b.run())
        bat_ms.append(1e3 * (time.perf_counter() - t))
    identical = all(np.array_equal(j.solver.getResults().view(np.uint32), r) for j, r in zip(jobs, ref))
    seq, bat = float(np.median(seq_ms)), float(np.median(bat_ms))
    print(json.dumps({
        "metric": "AO rays/s (batched jobs)", "value": rays / (bat * 1e-3), "unit": "rays/s",
        "jobs": len(jobs), "scale": args.scale, "rays": rays, "steps": args.steps,
        "batched_ms": bat, "graph_ms": float(np.median(graph_ms)), "sequential_ms": seq,
        "speedup": seq / bat, "identical": bool(identical), "launch": tuned, "prepare_s": prep_s,
        "families": {f: sum(j.spec["family"] == f for j in jobs) for f in FAMILIES},
    }), flush=True)
    return 0 if identical else 1


if __name__ == "__main__":
    raise SystemExit(main())
