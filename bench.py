"""AO rays/s on the production workload (2048^2 map over a 2.1M-triangle
heightfield, AO at 256 samples), one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-bake]

A step is one AO-256 pass over every texel of the map (one launch of the
hemisphere kernel, ~1.07e9 rays), with the scene, texel set, mapping result
and sample table already resident in HBM.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOAD = "heightfield2048"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes of an AO pass: per texel, per ray, per box test and per
# triangle test (the test counts come from a counting run of the same kernel)
TEXEL_BYTES = 16 + 4 + 96 + 4  # coords + tidx in, hit-triangle positions+normals, AO out
RAY_BYTES = 16                 # one sample direction
NODE_BYTES = 32                # one node test
TRI_BYTES = 64                 # one triangle test (predicate record)


def load_traffic(workload):
    """HBM bytes per AO launch of this workload from the committed rocprofv3
    --pmc summary, if there is one (profiles/ao_<workload>_pmc.json)."""
    p = os.path.join(ROOT, "profiles", "ao_%s_pmc.json" % workload)
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    except Exception:
        return None


def cpu_baseline(prep, mm, solver, budget_s=15.0):
    """The oracle (C restatement of the shaders, OpenMP over texels) on a
    bounded sample of the same AO-256 workload: contiguous 256-texel chunks
    spread over the map until ~budget_s of wall time.  Its per-texel occluder
    counts and AO values are compared with the GPU pass's on the same texels."""
    import numpy as np
    import oracle as orc
    nodes, _ = prep["scene"].reference_nodes()
    spos, snrm = prep["scene"].soup()
    coords, tidx = mm.read()
    gpu_counts = solver.counts()
    gpu_ao = solver.getResults().view(np.uint32)
    p = prep["cfg"]["maps"]["ao"]
    S = p["sampleCount"]
    n = len(tidx)
    threads = os.cpu_count() or 1
    starts = np.linspace(0, n - 256, 1024).astype(int)
    order = np.concatenate([starts[k::32] for k in range(32)])
    rays, texels, t = 0, 0, 0.0
    count_mismatch, value_mismatch = 0, 0
    for s in order:
        t0 = time.perf_counter()
        out, cnt = orc.hemisphere(orc.KIND_AO, S, p["minDistance"], p["maxDistance"], solver.table, coords, tidx,
                                  spos, snrm, nodes, int(s), int(s) + 256, nthreads=threads)
        t += time.perf_counter() - t0
        count_mismatch += int(np.sum(cnt != gpu_counts[s:s + 256]))
        value_mismatch += int(np.sum(out.view(np.uint32) != gpu_ao[s:s + 256]))
        live = int(np.sum(tidx[s:s + 256] != 0xFFFFFFFF))
        rays += live * S
        texels += 256
        if t > budget_s:
            break
    base = {"value": rays / t, "unit": "rays/s", "cores": threads, "kind": "port",
            "sample": "%d texels (%d-texel chunks spread over the map) x %d AO samples = %d rays, %.1f s"
                      % (texels, 256, S, rays, t)}
    parity = {"texels": texels, "rays": rays, "count_mismatches": count_mismatch,
              "value_mismatches": value_mismatch, "against": "oracle, same texels and sample table"}
    return base, parity


def denoise_leg(prep):
    """Denoised 16-sample bake (AO + bent normals through the U-Net, with the
    per-tile fallback) against the plain 1024-sample bake of the same map:
    wall times and the errors of the denoised maps against the plain ones."""
    import numpy as np
    from fornos_amd import denoise as dn
    if not os.path.exists(dn.WEIGHTS):
        return None
    import torch
    cfg = prep["cfg"]
    spec = {"seed": 0, "family": "heightfield", "low": cfg["low"], "high": cfg["high"], "cull": cfg["cull"],
            "ao": cfg["maps"]["ao"]}
    job = dn.MeshJob(spec, cfg["size"], (prep["low"], prep["dirs"], prep["high"], prep["scene"]))
    net = dn.Denoiser.load(dn.WEIGHTS)
    net.forward(torch.zeros((1, cfg["size"], cfg["size"], dn.IN), dtype=torch.float32, device="cuda"))
    # both bakes read their maps back into page-locked buffers allocated here
    plain_host = dn.readback_buffers(cfg["size"] * cfg["size"])
    dn_host = dn.readback_buffers(cfg["size"] * cfg["size"])
    ra, rb, tp, mm = dn.plain_bake(job, host=plain_host)
    _, tidx = mm.read()
    hit = tidx != 0xFFFFFFFF
    da, db, td, info = dn.denoised_bake(job, net, host=dn_host)
    ae, be = dn.errors(da, db, ra, rb, hit)
    print("# denoise: plain %.2f s, denoised %.3f s, AO MAE %.4f, BN %.2f deg" % (tp, td, ae, be),
          file=sys.stderr, flush=True)
    return {"plain_1024_s": tp, "denoised_16_s": td, "speedup": tp / td, "ao_mae": ae, "bn_deg": be,
            "network_ms": info["network_ms"], "fallback_tiles": info["fallback_tiles"], "tiles": info["tiles"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-bake", action="store_true")
    ap.add_argument("--no-denoise", action="store_true")
    ap.add_argument("--workload", default=WORKLOAD)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run one 1 GiB device-to-device copy (torch), the known-size kernel "
                         "that scales the rocprofv3 FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: the path runs on one GPU only; --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import fornos_amd as fa

    if fa.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)
    if args.calibrate_copy:
        # one 1 GiB device-to-device hipMemcpy (the runtime's copy kernel)
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        n = ctypes.c_size_t(1 << 30)
        assert hip.hipMalloc(ctypes.byref(a), n) == 0 and hip.hipMalloc(ctypes.byref(b), n) == 0
        assert hip.hipMemset(a, 0, n) == 0
        assert hip.hipDeviceSynchronize() == 0
        assert hip.hipMemcpy(b, a, n, 3) == 0       # hipMemcpyDeviceToDevice
        assert hip.hipDeviceSynchronize() == 0
        hip.hipFree(a)
        hip.hipFree(b)
    t0 = time.perf_counter()
    prep = fa.prepare(args.workload)
    prep_s = time.perf_counter() - t0
    cfg = prep["cfg"]
    scene_info = prep["scene"].info()
    print("# prepared %s in %.1f s: %d hi tris, %d nodes, %d texels"
          % (args.workload, prep_s, scene_info["n_triangles"], scene_info["n_nodes"], prep["texels"].count),
          file=sys.stderr, flush=True)

    mm = fa.MeshMapping().init(prep["texels"], prep["scene"], cfg["cull"])
    map_ms = mm.run_all()
    p = cfg["maps"]["ao"]
    ao = fa.AmbientOcclusionSolver(p["sampleCount"], p["minDistance"], p["maxDistance"]).init(prep["texels"], mm)
    n = mm.count
    _, tidx = mm.read()
    live = int(np.sum(tidx != 0xFFFFFFFF))
    S = p["sampleCount"]
    rays_per_step = live * S

    for _ in range(args.warmup):
        ao.reset()
        ao.run_all()
    pass_ms, rays_ms, reduce_ms = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ao.reset()
        pass_ms.append(ao.run_all())  # ends with hipEventSynchronize on the launch stream
        r, d, launches = ao.last_timing()
        rays_ms.append(r)
        reduce_ms.append(d)
    wall = time.perf_counter() - t0
    ms_per_step = 1000.0 * wall / args.steps
    value = rays_per_step * args.steps / wall
    print("# AO-%d: %.2f ms/step, %.3e rays/s" % (S, ms_per_step, value), file=sys.stderr, flush=True)

    # roofline of the dominant kernel, k_rays<AO>: algorithmic bytes of one
    # launch over that launch's average duration (HIP events on its stream)
    work = ao.count_work()
    alg_pass = (n * TEXEL_BYTES + work["rays"] * RAY_BYTES + work["node_tests"] * NODE_BYTES
                + work["tri_tests"] * TRI_BYTES)
    launch_ms = float(np.mean(rays_ms)) / launches
    achieved = alg_pass / launches / (launch_ms * 1e-3) / 1e9
    traffic = load_traffic(args.workload)
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                "kernel": "k_rays<AO>", "launches_per_step": launches, "launch_ms": launch_ms,
                "algorithmic_bytes_per_launch": alg_pass / launches,
                "reduce_ms_per_step": float(np.mean(reduce_ms)), "pass_kernel_ms": float(np.mean(pass_ms)),
                "node_tests_per_ray": work["node_tests"] / max(1, work["rays"]),
                "tri_tests_per_ray": work["tri_tests"] / max(1, work["rays"])}
    if traffic:
        # the algorithmic bytes count every node and triangle fetch, most of
        # which L2 and the MALL serve, so `frac` can pass 1; DRAM bytes from
        # the rocprofv3 counters over the same launch time cannot
        roofline["dram_gbs"] = traffic / (launch_ms * 1e-3) / 1e9
        roofline["dram_frac"] = roofline["dram_gbs"] / HBM_PEAK_GBS

    bake = None
    if not args.no_bake:
        # seconds per full bake of the workload: mapping + every map it names,
        # and with the exports each reference task writes (8-bit PNG, dilation
        # 16), which overlap the later maps' bakes; AO and bent normals with
        # the same rays share one traversal
        import tempfile
        from fornos_amd import bake as fb
        with tempfile.TemporaryDirectory() as td:
            rep = fb.bake(cfg, td, "png", prep=prep)
        bake = {"seconds": rep["bake_s"], "with_export_seconds": rep["total_s"],
                "export_wait_seconds": rep["export_wait_s"],
                "kernel_ms": {k[:-3]: v for k, v in rep.items() if k.endswith("_ms")},
                "paired": rep.get("paired")}

    denoise = None if args.no_denoise else denoise_leg(prep)
    cpu, parity = (None, None) if args.no_cpu else cpu_baseline(prep, mm, ao)
    line = {
        "metric": "AO rays/s", "value": value, "unit": "rays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded procedural meshes)",
        "config": {"workload": "%s: %d^2 map, %d-triangle high-poly mesh, AO %d samples min %.2f max %.2f"
                               % (args.workload, cfg["size"], scene_info["n_triangles"], S, p["minDistance"],
                                  p["maxDistance"]),
                   "texels": n, "rays_per_step": rays_per_step,
                   "bvh": "reference binary tree (exact reachability) + 4-wide SAH culling tree"},
        "roofline": roofline, "cpu_baseline": cpu, "parity": parity, "mapping_ms": map_ms, "bake": bake, "denoise": denoise,
        "texels_per_s": live / (ms_per_step * 1e-3),
    }
    print(json.dumps(line), flush=True)
    if parity and (parity["count_mismatches"] or parity["value_mismatches"]):
        print("bench.py: the AO pass differs from the oracle on %d texels" % parity["count_mismatches"],
              file=sys.stderr)
        sys.exit(1)


if __name__ == "__main__":
    main()
