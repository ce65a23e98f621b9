"""AO rays/s on the production workload (2048^2 map over a 2.1M-triangle
heightfield, AO at 256 samples), one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-bake]
                    [--no-denoise] [--no-knot] [--no-batch] [--workload NAME]

A step is one AO-256 pass over every texel of the map (~1.07e9 rays, four
launches of the hemisphere kernel), with the scene, texel set, mapping result
and sample table already resident in HBM.  Before the warm-up, the solver's
launch shape, hand-out tile and ray order are autotuned on 5 % of the pass
(fornos_solver_autotune, reported with its time under config.launch_shape and
config.ray_order; --launch B,R,S fixes the shape with the default tile and
order instead, as the profiles do; --no-tune keeps the defaults).  Prints one JSON line, with beside
the headline: the 8.4M-triangle knot workload's AO-512 pass (`workloads`),
the 64 batched jobs against 64 sequential passes (`batch`), the full bake, the
denoised bake, and the oracle on a bounded sample of the headline pass (CPU
baseline and parity).

Roofline of the dominant kernel, k_rays<AO>, per launch (launch time from HIP
events on its stream, live): an issue roofline (wave64 VALU instructions per
launch, from the rocprofv3 SQ_INSTS_VALU counter, each counted as a full-width
fp32 FMA, against the published 157.3 TFLOP/s fp32 vector peak) and an HBM
roofline (DRAM bytes per launch, from the calibrated FETCH_SIZE and WRITE_SIZE
counters, against 8 TB/s); `bound` names the nearer ceiling.  The counters per launch come from profiles/ao_<workload>_pmc.json.
The traversal's node and triangle fetches, most of which L2 and the MALL
serve, are reported separately as an L2-level byte model.
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
# published fp32 vector peak: 256 CUs x 2.4 GHz x 128 FMA lanes x 2 flops, a
# figure that counts packed fp32 (two FMAs per lane per instruction), which this
# build does not use; the issue roofline counts every wave64 VALU instruction as
# a full-width FMA (128 flops), so its fraction is the VALU issue rate over the
# published one, and non-packed code tops out near half of it
VALU_PEAK_TFLOPS = 157.3
FLOPS_PER_VALU_INST = 128
# L2-level model of a pass's fetches: per texel, per ray, per box test and per
# triangle test (the test counts come from a counting run of the same kernel)
TEXEL_BYTES = 16 + 4 + 96 + 4  # coords + tidx in, hit-triangle positions+normals, AO out
RAY_BYTES = 16                 # one sample direction
NODE_BYTES = 32                # one child box of a 4-wide node
TRI_BYTES = 64                 # one triangle test (predicate record)


def load_profile(workload):
    """rocprofv3 counters per k_rays<AO> launch of this workload, from the
    committed summary profiles/ao_<workload>_pmc.json (None if absent)."""
    p = os.path.join(ROOT, "profiles", "ao_%s_pmc.json" % workload)
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f)


def ao_pass(fa, workload, steps, warmup, log, tune=True, launch=None):
    """Prepares a workload and times its AO pass: returns the figures and the
    live objects (prep, mapping, solver)."""
    import numpy as np
    t0 = time.perf_counter()
    prep = fa.prepare(workload)
    prep_s = time.perf_counter() - t0
    cfg = prep["cfg"]
    info = prep["scene"].info()
    log("# prepared %s in %.1f s: %d hi tris, %d nodes, %d texels"
        % (workload, prep_s, info["n_triangles"], info["n_nodes"], prep["texels"].count))
    mm = fa.MeshMapping().init(prep["texels"], prep["scene"], cfg["cull"])
    mm.run_all()
    mm.reset()
    map_ms = mm.run_all()  # the second run: the first one also loads the kernel
    p = cfg["maps"]["ao"]
    ao = fa.AmbientOcclusionSolver(p["sampleCount"], p["minDistance"], p["maxDistance"]).init(prep["texels"], mm)
    _, tidx = mm.read()
    live = int(np.sum(tidx != 0xFFFFFFFF))
    S = p["sampleCount"]
    tuning = None
    auto_order = ao.effective_ray_order()  # the order the profiles were taken in (the automatic one)
    if launch:
        ao.set_launch(*launch)
        tuning = {"blocks_per_cu": launch[0], "refill": launch[1], "stack": launch[2], "tile": ao.get_tile(),
                  "tuned": False}
    elif tune:
        # launch shape, hand-out tile and ray order timed on 5 % of the pass
        # (results never depend on them); the cost is set-up, outside the
        # timed steps, and reported
        t0 = time.perf_counter()
        bpc, refill, stack, tms = ao.autotune(0.05)
        tuning = {"blocks_per_cu": bpc, "refill": refill, "stack": stack, "tile": ao.get_tile(), "tuned": True,
                  "seconds": time.perf_counter() - t0, "fraction": 0.05}
        log("# %s autotuned launch: %d blocks/CU, refill at %d idle lanes, stack %d, %d texel(s) per hand-out, "
            "%s order (%.2f s)" % (workload, bpc, refill, stack, tuning["tile"],
                                   "Morton" if ao.effective_ray_order() == ao.RAY_ORDER_MORTON else "texel",
                                   tuning["seconds"]))
    for _ in range(warmup):
        ao.reset()
        ao.run_all()
    pass_ms, rays_ms, reduce_ms = [], [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        ao.reset()
        pass_ms.append(ao.run_all())  # ends with hipEventSynchronize on the launch stream
        r, d, launches = ao.last_timing()
        rays_ms.append(r)
        reduce_ms.append(d)
    wall = time.perf_counter() - t0
    res = {"workload": workload, "prep": prep, "mm": mm, "ao": ao, "cfg": cfg, "params": p, "S": S,
           "texels": mm.count, "live": live, "rays_per_step": live * S, "info": info, "map_ms": map_ms,
           "wall": wall, "steps": steps, "ms_per_step": 1000.0 * wall / steps,
           "value": live * S * steps / wall, "launches": launches,
           "launch_ms": float(np.mean(rays_ms)) / launches, "reduce_ms": float(np.mean(reduce_ms)),
           "pass_kernel_ms": float(np.mean(pass_ms)), "launch_shape": tuning or {"blocks_per_cu": 8, "refill": 32,
                                                                                 "stack": 16, "tile": ao.get_tile(),
                                                                                 "tuned": False},
           "ray_order": ("Morton code of the texels' hit positions"
                         if ao.effective_ray_order() == ao.RAY_ORDER_MORTON else "texel set order"),
           "default_tile_and_order": ao.get_tile() == (1 if S >= 256 else min(32, 256 // S)) and
                                     ao.effective_ray_order() == auto_order}
    log("# %s AO-%d: %.2f ms/step, %.3e rays/s" % (workload, S, res["ms_per_step"], res["value"]))
    return res


def roofline(res):
    """Issue and HBM rooflines of one k_rays<AO> launch: counters per launch
    from the committed rocprofv3 summary over the live launch time."""
    work = res["ao"].count_work()
    launches, t = res["launches"], res["launch_ms"] * 1e-3
    l2_bytes = (res["texels"] * TEXEL_BYTES + work["rays"] * RAY_BYTES + work["node_tests"] * NODE_BYTES
                + work["tri_tests"] * TRI_BYTES) / launches
    out = {"kernel": "k_rays<AO>", "launches_per_step": launches, "launch_ms": res["launch_ms"],
           "reduce_ms_per_step": res["reduce_ms"], "pass_kernel_ms": res["pass_kernel_ms"],
           "node_tests_per_ray": work["node_tests"] / max(1, work["rays"]),
           "tri_tests_per_ray": work["tri_tests"] / max(1, work["rays"]),
           "l2_model_bytes_per_launch": l2_bytes, "l2_model_gbs": l2_bytes / t / 1e9}
    prof = load_profile(res["workload"]) or {}
    hbm, valu = prof.get("hbm_bytes_per_launch"), prof.get("valu_insts_per_launch")
    shape = res["launch_shape"]
    pshape = prof.get("launch_shape")
    out["profile_launch_shape"] = pshape
    if pshape and ([shape["blocks_per_cu"], shape["refill"], shape["stack"]] != list(pshape) or
                   not res["default_tile_and_order"]):
        out["profile_note"] = ("counters taken at launch shape %s with the default hand-out tile and the automatic "
                               "ray order, this run used %s, tile %d, %s"
                               % (pshape, [shape["blocks_per_cu"], shape["refill"], shape["stack"]], shape["tile"],
                                  res["ray_order"]))
    cands = []
    if valu:
        tf = valu * FLOPS_PER_VALU_INST / t / 1e12
        issue = {"achieved": tf, "peak": VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                 "frac": tf / VALU_PEAK_TFLOPS, "insts_per_launch": valu, "insts_per_s": valu / t,
                 "lanes_active": prof.get("valu_lane_fraction"),
                 "note": "VALU issue: each wave64 VALU instruction counted as a 128-flop FMA"}
        out["issue"] = issue
        cands.append(("valu", issue))
    if hbm:
        mem = {"achieved": hbm / t / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "frac": hbm / t / 1e9 / HBM_PEAK_GBS, "bytes_per_launch": hbm}
        out["hbm"] = mem
        cands.append(("hbm", mem))
    if not cands:
        out.update({"bound": None, "achieved": None, "peak": None, "unit": None, "frac": None, "traffic": None,
                    "note": "no rocprofv3 counter summary for this workload in profiles/"})
        return out
    bound, top = max(cands, key=lambda c: c[1]["frac"])
    out.update({"bound": bound, "achieved": top["achieved"], "peak": top["peak"], "unit": top["unit"],
                "frac": top["frac"], "traffic": hbm, "profile": "profiles/ao_%s_pmc.json" % res["workload"]})
    return out


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
            "label": "CPU restatement, not llvmpipe (no headless GL context on the GPU box)",
            "sample": "%d texels (%d-texel chunks spread over the map) x %d AO samples = %d rays, %.1f s"
                      % (texels, 256, S, rays, t)}
    parity = {"texels": texels, "rays": rays, "count_mismatches": count_mismatch,
              "value_mismatches": value_mismatch, "against": "oracle, same texels and sample table"}
    return base, parity


def denoise_leg(prep, log):
    """Denoised 16-sample bake (AO + bent normals through the U-Net, with the
    per-tile fallback) against the plain 1024-sample bake of the same map:
    wall times and the errors of the denoised maps against the plain ones."""
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
    # the denoised bake is short (tens of ms): the median of three runs
    runs = [dn.denoised_bake(job, net, host=dn_host) for _ in range(3)]
    order = sorted(range(3), key=lambda i: runs[i][2])
    da, db, td, info = runs[order[1]]
    ae, be = dn.errors(da, db, ra, rb, hit)
    log("# denoise: plain %.2f s, denoised %.3f s (median of %s), AO MAE %.4f, BN %.2f deg"
        % (tp, td, ", ".join("%.3f" % r[2] for r in runs), ae, be))
    return {"plain_1024_s": tp, "denoised_16_s": td, "denoised_runs_s": [r[2] for r in runs], "speedup": tp / td,
            "ao_mae": ae, "bn_deg": be, "network_ms": info["network_ms"], "fallback_tiles": info["fallback_tiles"],
            "tiles": info["tiles"],
            "breakdown_s": {k: info[k] for k in ("bake16_s", "network_s", "check_s", "fallback_s")}}


def batch_leg(log, n_jobs=64):
    """The 64 seeded AO jobs of uneven size (sphere / heightfield / knot,
    256^2-2048^2 maps, 32-256 samples): one graph launch of the batch against
    one run_all per job, each after a warm-up; every job's results compared."""
    import numpy as np
    from fornos_amd import batch
    t0 = time.perf_counter()
    jobs = [batch.Job(s) for s in batch.job_specs(n_jobs)]
    prep_s = time.perf_counter() - t0
    rays = sum(j.rays for j in jobs)

    def sequential():
        t = time.perf_counter()
        for j in jobs:
            j.solver.reset()
            j.solver.run_all()
        return time.perf_counter() - t

    sequential()
    seq_s = sequential()
    ref = [j.solver.getResults().view(np.uint32).copy() for j in jobs]
    b = batch.Batch([j.solver for j in jobs])
    b.run()
    t = time.perf_counter()
    graph_ms = b.run()
    bat_s = time.perf_counter() - t
    identical = all(np.array_equal(j.solver.getResults().view(np.uint32), r) for j, r in zip(jobs, ref))
    log("# batch: %d jobs, %.3e rays: sequential %.3f s, batched %.3f s" % (len(jobs), rays, seq_s, bat_s))
    out = {"jobs": len(jobs), "rays": rays, "high_tris": sum(j.high_tris for j in jobs), "prepare_s": prep_s,
           "sequential_s": seq_s, "batched_s": bat_s, "graph_ms": graph_ms, "speedup": seq_s / bat_s,
           "rays_per_s": rays / bat_s, "identical": bool(identical), "launch": list(b.get_launch())}
    b.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-bake", action="store_true")
    ap.add_argument("--no-denoise", action="store_true")
    ap.add_argument("--no-knot", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="default launch shape instead of the autotuned one")
    ap.add_argument("--launch", default=None,
                    help="BLOCKS_PER_CU,REFILL,STACK: this launch shape for the headline workload instead of tuning "
                         "(the profiles are taken at the tuned shape this way)")
    ap.add_argument("--workload", default=WORKLOAD)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run one 1 GiB device-to-device copy, the known-size kernel that scales the "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: the path runs on one GPU only; --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    log = lambda s: print(s, file=sys.stderr, flush=True)

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

    launch = tuple(int(v) for v in args.launch.split(",")) if args.launch else None
    head = ao_pass(fa, args.workload, args.steps, args.warmup, log, tune=not args.no_tune, launch=launch)
    roof = roofline(head)

    workloads = {}
    if not args.no_knot and args.workload != "knot4096":
        k = ao_pass(fa, "knot4096", 1, 1, log, tune=not args.no_tune)
        workloads["knot4096"] = {
            "metric": "AO rays/s", "value": k["value"], "unit": "rays/s", "ms_per_step": k["ms_per_step"],
            "steps": 1, "warmup": 1,
            "config": "knot4096: %d^2 map, %d-triangle torus-knot tube, AO %d samples min %.2f max %.2f"
                      % (k["cfg"]["size"], k["info"]["n_triangles"], k["S"], k["params"]["minDistance"],
                         k["params"]["maxDistance"]),
            "texels": k["texels"], "rays_per_step": k["rays_per_step"], "mapping_ms": k["map_ms"],
            "ray_order": k["ray_order"], "launch_shape": k["launch_shape"],
            "roofline": roofline(k)}
        del k

    batch = None if args.no_batch else batch_leg(log)

    cfg = head["cfg"]
    bake = None
    if not args.no_bake:
        # seconds per full bake of the workload: mapping + every map it names,
        # and with the exports each reference task writes (8-bit PNG, dilation
        # 16), which overlap the later maps' bakes; AO and bent normals with
        # the same rays share one traversal
        import tempfile
        from fornos_amd import bake as fb
        with tempfile.TemporaryDirectory() as td:
            rep = fb.bake(cfg, td, "png", prep=head["prep"])
        bake = {"seconds": rep["bake_s"], "with_export_seconds": rep["total_s"],
                "export_wait_seconds": rep["export_wait_s"],
                "kernel_ms": {k[:-3]: v for k, v in rep.items() if k.endswith("_ms")},
                "paired": rep.get("paired")}

    denoise = None if args.no_denoise else denoise_leg(head["prep"], log)
    cpu, parity = (None, None) if args.no_cpu else cpu_baseline(head["prep"], head["mm"], head["ao"])
    p, S = head["params"], head["S"]
    line = {
        "metric": "AO rays/s", "value": head["value"], "unit": "rays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded procedural meshes)",
        "config": {"workload": "%s: %d^2 map, %d-triangle high-poly mesh, AO %d samples min %.2f max %.2f"
                               % (args.workload, cfg["size"], head["info"]["n_triangles"], S, p["minDistance"],
                                  p["maxDistance"]),
                   "texels": head["texels"], "rays_per_step": head["rays_per_step"],
                   "bvh": "reference binary tree (exact reachability) + 4-wide SAH culling tree",
                   "ray_order": head["ray_order"], "launch_shape": head["launch_shape"]},
        "roofline": roof, "cpu_baseline": cpu, "parity": parity, "mapping_ms": head["map_ms"],
        "texels_per_s": head["live"] / (head["ms_per_step"] * 1e-3),
        "workloads": workloads, "batch": batch, "bake": bake, "denoise": denoise,
    }
    print(json.dumps(line), flush=True)
    if parity and (parity["count_mismatches"] or parity["value_mismatches"]):
        print("bench.py: the AO pass differs from the oracle on %d texels" % parity["count_mismatches"],
              file=sys.stderr)
        sys.exit(1)
    if batch and not batch["identical"]:
        print("bench.py: a batched job differs from its sequential pass", file=sys.stderr)
        sys.exit(1)


if __name__ == "__main__":
    main()
