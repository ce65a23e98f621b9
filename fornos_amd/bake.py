"""Headless bake: FornosRunner::start's assembly (Src/fornos.cpp:53-236) for
a seeded workload, every map it names baked on the GPU and exported the way
each Task::finish does (Src/solver_*.cpp), with the shared dilation
(Src/fornos.h:47).

    python -m fornos_amd.bake --workload heightfield2048 --out bake_dir [--dilation 16] [--ext png]
"""
import argparse
import json
import os
import sys
import time

from . import configs, export
from .api import (AmbientOcclusionSolver, BentNormalsSolver, HeightSolver, MeshMapping, NormalsSolver,
                  PositionSolver, ThicknessSolver)

SOLVERS = {"height": HeightSolver, "position": PositionSolver, "normals": NormalsSolver,
           "ao": AmbientOcclusionSolver, "bentnormals": BentNormalsSolver, "thickness": ThicknessSolver}


def bake(cfg, out_dir=None, ext="png", dilation=export.DEFAULT_DILATION, prep=None):
    """Runs the mapping and every map of `cfg`; exports when out_dir is set.
    Returns per-map kernel ms and export seconds."""
    prep = prep or configs.prepare(cfg)
    cfg = prep["cfg"]
    report = {}
    t0 = time.perf_counter()
    mm = MeshMapping().init_from_mesh(prep["low"], prep["dirs"], cfg["size"], cfg["size"], prep["scene"],
                                      cfg["cull"])
    report["mapping_ms"] = mm.run_all()
    solvers = {}
    for name, kw in cfg["maps"].items():
        s = SOLVERS[name](**kw).init(None, mm)
        report[name + "_ms"] = s.run_all()
        solvers[name] = s
    report["bake_s"] = time.perf_counter() - t0
    if out_dir:
        os.makedirs(out_dir, exist_ok=True)
        t1 = time.perf_counter()
        for name, s in solvers.items():
            e = "exr" if name == "position" else ext
            r = export.export_task(s, os.path.join(out_dir, "%s.%s" % (name, e)), dilation)
            if r is not None:
                report[name + "_range"] = r
        report["export_s"] = time.perf_counter() - t1
    return report


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m fornos_amd.bake")
    ap.add_argument("--workload", default="heightfield2048", choices=sorted(configs.CONFIGS))
    ap.add_argument("--out", default=None)
    ap.add_argument("--ext", default="png", choices=["png", "tga", "exr"])
    ap.add_argument("--dilation", type=int, default=export.DEFAULT_DILATION)
    args = ap.parse_args(argv)
    rep = bake(args.workload, args.out, args.ext, args.dilation)
    print(json.dumps(rep), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
