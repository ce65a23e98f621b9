"""Headless bake.

Seeded workloads run FornosRunner::start's assembly in Python
(fornos_amd.configs.prepare) and bake every map they name on the GPU,
exported the way each Task::finish does (Src/solver_*.cpp) with the shared
dilation (Src/fornos.h:47):

    python -m fornos_amd.bake --workload heightfield2048 --out bake_dir [--dilation 16] [--ext png]

Mesh files go through the native runner (include/fornos_runner.h), the
reference's FornosRunner with its parameters:

    python -m fornos_amd.bake --low low.obj [--high high.ply] --out bake_dir \\
        --maps ao,normals,height [--size 2048] [--mapping smooth|lowpoly|hybrid] [--edge 0.05] \\
        [--low-normals import|face|vertex] [--high-normals ...] [--samples 256] \\
        [--min-distance 0.01] [--max-distance 10] [--object-space]
"""
import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

from . import _abi, configs, export
from .api import (AmbientOcclusionSolver, BentNormalsSolver, HeightSolver, MeshMapping, NormalsSolver,
                  PositionSolver, ThicknessSolver, shares_rays)
from .runner import FornosParameters, FornosRunner, MeshMappingMethod, NormalImport

SOLVERS = {"height": HeightSolver, "position": PositionSolver, "normals": NormalsSolver,
           "ao": AmbientOcclusionSolver, "bentnormals": BentNormalsSolver, "thickness": ThicknessSolver}
EXT = {"position": "exr"}


def bake(cfg, out_dir=None, ext="png", dilation=export.DEFAULT_DILATION, prep=None, pair=True):
    """Runs the mapping and every map of `cfg`; exports when out_dir is set.
    AO and bent normals casting the same rays share one traversal (pair).
    Returns per-map kernel ms and export seconds."""
    prep = prep or configs.prepare(cfg)
    cfg = prep["cfg"]
    report = {}
    t0 = time.perf_counter()
    mm = MeshMapping().init_from_mesh(prep["low"], prep["dirs"], prep["width"], prep["height"], prep["scene"],
                                      cfg.get("cull", True), edge=prep["edge"])
    report["mapping_ms"] = mm.run_all()
    solvers = {name: SOLVERS[name](**kw).init(None, mm) for name, kw in cfg["maps"].items()}
    maps = cfg["maps"]
    if pair and "ao" in maps and "bentnormals" in maps and shares_rays(maps["ao"], maps["bentnormals"]):
        solvers["ao"].pair(solvers["bentnormals"])  # one traversal for both maps
        report["paired"] = ["ao", "bentnormals"]
    # each map is exported on a host thread (its device part on the scene's
    # export stream, one export at a time; the encodes overlap) while the
    # next maps are baked: heightfield2048 with exports 0.44-0.49 s with one
    # thread, 0.38-0.42 s with three
    pool = ThreadPoolExecutor(max_workers=3) if out_dir else None
    if out_dir:
        os.makedirs(out_dir, exist_ok=True)
    exports = []
    for name, s in solvers.items():
        report[name + "_ms"] = s.run_all()
        if pool:
            path = os.path.join(out_dir, "%s.%s" % (name, EXT.get(name, ext)))
            exports.append((name, pool.submit(export.export_task, s, path, dilation)))
    report["bake_s"] = time.perf_counter() - t0
    if pool:
        t1 = time.perf_counter()
        for name, fut in exports:
            r = fut.result()
            if r is not None:
                report[name + "_range"] = r
        pool.shutdown()
        report["export_wait_s"] = time.perf_counter() - t1  # export time not hidden behind the bake
        report["total_s"] = time.perf_counter() - t0
    return report


def runner_parameters(args):
    """FornosParameters from the command line (file inputs)."""
    p = FornosParameters()
    sh = p.shared
    sh.loPolyMeshPath, sh.hiPolyMeshPath = args.low, args.high or ""
    modes = {"import": NormalImport.Import, "face": NormalImport.ComputePerFace,
             "vertex": NormalImport.ComputePerVertex}
    sh.loPolyMeshNormal, sh.hiPolyMeshNormal = modes[args.low_normals], modes[args.high_normals]
    sh.texWidth = args.width or args.size
    sh.texHeight = args.height or args.size
    sh.texDilation = args.dilation
    sh.ignoreBackfaces = not args.keep_backfaces
    sh.mapping = {"smooth": MeshMappingMethod.Smooth, "lowpoly": MeshMappingMethod.LowPolyNormals,
                  "hybrid": MeshMappingMethod.Hybrid}[args.mapping]
    sh.mappingEdge = args.edge
    targets = {"height": p.height, "position": p.positions, "normals": p.normals, "ao": p.ao,
               "bentnormals": p.bentNormals, "thickness": p.thickness}
    for name in [m for m in args.maps.split(",") if m]:
        if name not in targets:
            raise SystemExit("unknown map %r (one of %s)" % (name, ",".join(sorted(targets))))
        t = targets[name]
        t.enabled = True
        t.outputPath = os.path.join(args.out, "%s.%s" % (name, EXT.get(name, args.ext))) if args.out else ""
        if hasattr(t, "sampleCount"):
            t.sampleCount, t.minDistance, t.maxDistance = args.samples, args.min_distance, args.max_distance
        if hasattr(t, "tangentSpace"):
            t.tangentSpace = not args.object_space
        if name == "height" and args.height_max_distance is not None:
            t.maxDistance = args.height_max_distance
    return p


def run_files(args):
    p = runner_parameters(args)
    if args.out:
        os.makedirs(args.out, exist_ok=True)
    r = FornosRunner()
    t0 = time.perf_counter()
    ok, errors = r.start(p)
    if not ok:
        print(json.dumps({"error": errors}), flush=True)
        return 1
    report = {"start_s": time.perf_counter() - t0, "tasks": []}
    while r.pending():
        t = r.currentTask()
        name, t1 = t.name(), time.perf_counter()
        while r.pending() and r.currentTask() is not None and r.currentTask().name() == name:
            r.run()
            time.sleep(0.0005)
        report["tasks"].append({"name": name, "seconds": time.perf_counter() - t1})
    for kind, name in ((_abi.SOLVER_AO, "ao"), (_abi.SOLVER_THICKNESS, "thickness"), (_abi.SOLVER_HEIGHT, "height")):
        rng = r.exported_range(kind)
        if rng is not None:
            report[name + "_range"] = rng
    report["total_s"] = time.perf_counter() - t0
    print(json.dumps(report), flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m fornos_amd.bake")
    ap.add_argument("--workload", default="heightfield2048", choices=sorted(configs.CONFIGS))
    ap.add_argument("--low", default=None, help="low-poly mesh (.obj / .ply): bake through the runner")
    ap.add_argument("--high", default=None, help="hi-poly mesh (default: the low-poly mesh)")
    ap.add_argument("--low-normals", default="import", choices=configs.NORMAL_MODES)
    ap.add_argument("--high-normals", default="import", choices=configs.NORMAL_MODES)
    ap.add_argument("--mapping", default="smooth", choices=configs.MAPPINGS)
    ap.add_argument("--edge", type=float, default=0.05, help="Hybrid mapping edge width")
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--width", type=int, default=None)
    ap.add_argument("--height", type=int, default=None)
    ap.add_argument("--maps", default="normals,ao", help="comma list of height,position,normals,ao,bentnormals,thickness")
    ap.add_argument("--samples", type=int, default=256)
    ap.add_argument("--min-distance", type=float, default=0.01)
    ap.add_argument("--max-distance", type=float, default=10.0)
    ap.add_argument("--height-max-distance", type=float, default=None)
    ap.add_argument("--object-space", action="store_true", help="normals / bent normals in object space")
    ap.add_argument("--keep-backfaces", action="store_true", help="ignoreBackfaces = false")
    ap.add_argument("--out", default=None)
    ap.add_argument("--ext", default="png", choices=["png", "tga", "exr"])
    ap.add_argument("--dilation", type=int, default=export.DEFAULT_DILATION)
    args = ap.parse_args(argv)
    if args.low:
        return run_files(args)
    rep = bake(args.workload, args.out, args.ext, args.dilation)
    print(json.dumps(rep), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
