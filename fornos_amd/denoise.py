"""Learned AO / bent-normal denoiser: a 4-level, 32-channel U-Net run by
libfornos_hip.so (include/fornos_denoise.h) on the bf16 matrix cores, trained
on the GPU by the library's own forward / backward / Adam kernels.

No reference counterpart: the reference bakes AO and bent normals by brute
force only (Src/solver_ao.cpp, Src/solver_bentnormals.cpp).  The denoised
bake runs the plain GPU baker at 16 samples plus the position, normal and
mapping-distance passes, feeds them through the network, and falls back to
the plain 1024-sample bake on every 32 x 32 block whose output fails a
consistency check against its noisy input (those texels are then exactly the
plain bake's).

Training pairs are 16-sample inputs and 1024-sample targets baked by the plain
GPU method on windows of seeded procedural meshes (the sphere, heightfield and
knot generators of the batched workload).  torch is used for device memory
only (images and crops); every network computation is the library's.

    python -m fornos_amd.denoise train --seeds 1000:1200 --out weights.npz
    python -m fornos_amd.denoise eval --weights weights.npz --seeds 2000:2020
"""
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

from . import _abi
from ._abi import check, lib
from .api import (AmbientOcclusionSolver, BentNormalsSolver, HeightSolver, Mesh, MeshMapping, NormalsSolver,
                  PositionSolver, Scene, TexelArrays, _Handle)

IN, OUT, MASK = 12, 4, 11
LOW_S, HIGH_S = 16, 1024
PLAIN_SEED = 7  # sample-table seed of the plain 1024-sample bake (and so of the fallback)
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "denoiser_l4c32.npz")


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the denoiser needs a HIP device")
    return torch


def _p(t):
    return C.c_void_p(t.data_ptr())


class Denoiser(_Handle):
    """The network handle (fornos_denoiser_*)."""
    _destroy = "fornos_denoiser_destroy"

    def __init__(self, levels=4, base=32, seed=0):
        super().__init__()
        self.levels, self.base = levels, base
        check(lib().fornos_denoiser_create(levels, base, seed, C.byref(self._h)))
        n = C.c_uint64()
        check(lib().fornos_denoiser_param_count(self._h, C.byref(n)))
        self.n_params = n.value

    def layers(self):
        n = C.c_uint32()
        check(lib().fornos_denoiser_layer_count(self._h, C.byref(n)))
        out = []
        for i in range(n.value):
            v = [C.c_uint32() for _ in range(4)] + [C.c_uint64()]
            check(lib().fornos_denoiser_layer(self._h, i, *[C.byref(x) for x in v]))
            out.append(dict(zip(["c1", "c2", "cout", "up", "offset"], [x.value for x in v])))
        return out

    def get_params(self):
        p = np.empty(self.n_params, np.float32)
        check(lib().fornos_denoiser_get_params(self._h, p.ctypes.data_as(C.c_void_p), self.n_params))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        check(lib().fornos_denoiser_set_params(self._h, p.ctypes.data_as(C.c_void_p), len(p)))

    def grads(self):
        g = np.empty(self.n_params, np.float32)
        check(lib().fornos_denoiser_get_grads(self._h, g.ctypes.data_as(C.c_void_p), self.n_params))
        return g

    def forward(self, inp, out=None):
        """inp: float32 device tensor [n, h, w, 12] -> ([n, h, w, 4], device ms)."""
        torch = _torch()
        inp = inp.contiguous()
        n, h, w, c = inp.shape
        assert c == IN and inp.dtype == torch.float32 and inp.is_cuda
        if out is None:
            out = torch.empty((n, h, w, OUT), dtype=torch.float32, device=inp.device)
        ms = C.c_float()
        torch.cuda.synchronize()
        check(lib().fornos_denoiser_forward(self._h, _p(inp), n, h, w, _p(out), C.byref(ms)))
        return out, ms.value

    def backward(self, inp, target, want_output=False):
        """Loss of the batch; the gradient stays in the handle."""
        torch = _torch()
        inp, target = inp.contiguous(), target.contiguous()
        n, h, w, _ = inp.shape
        out = torch.empty((n, h, w, OUT), dtype=torch.float32, device=inp.device) if want_output else None
        loss = C.c_float()
        torch.cuda.synchronize()
        check(lib().fornos_denoiser_backward(self._h, _p(inp), _p(target), n, h, w,
                                             _p(out) if out is not None else None, C.byref(loss)))
        return (loss.value, out) if want_output else loss.value

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        check(lib().fornos_denoiser_adam(self._h, lr, beta1, beta2, eps))

    def save(self, path, meta=None):
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        np.savez_compressed(path, params=self.get_params(), levels=self.levels, base=self.base,
                            meta=json.dumps(meta or {}))

    @classmethod
    def load(cls, path=WEIGHTS):
        z = np.load(path, allow_pickle=False)
        d = cls(int(z["levels"]), int(z["base"]))
        d.set_params(z["params"])
        return d


# ---------------------------------------------------------------------------
# bakes on a texel set
# ---------------------------------------------------------------------------
class MeshJob:
    """A seeded mesh pair with its device scene and map-sized texel set."""

    def __init__(self, spec, size=2048, meshes=None):
        """meshes: an already prepared (low, dirs, high, scene), else built from spec."""
        self.spec = spec
        self.size = size
        if meshes is None:
            lo = Mesh.generate(*spec["low"])
            hi = Mesh.generate(*spec["high"])
            dirs = lo.copy()
            dirs.computeVertexNormalsAggressive()
            scene = Scene(hi)
        else:
            lo, dirs, hi, scene = meshes
        self.lo, self.dirs = lo, dirs
        self.scene = scene
        self.cull = spec["cull"]
        pos = hi.arrays()["positions"]
        a, b = pos.min(0), pos.max(0)
        self.center = ((a + b) * 0.5).astype(np.float32)
        self.diag = float(np.linalg.norm(b - a))
        self.high_tris = hi.n_triangles
        ao = spec["ao"]
        self.mind, self.maxd = ao["minDistance"], ao["maxDistance"]
        mm = MeshMapping().init_from_mesh(lo, dirs, size, size, self.scene, self.cull)
        self.texels = TexelArrays(mm.read_texels(), size, size)
        mm.close()

    def window(self, x0, y0, w):
        """Texels whose map pixel lies in [x0, x0+w) x [y0, y0+w), re-indexed into the window."""
        idx = self.texels.a["indices"].astype(np.int64)
        x, y = idx % self.size, idx // self.size
        sel = (x >= x0) & (x < x0 + w) & (y >= y0) & (y < y0 + w)
        local = ((y[sel] - y0) * w + (x[sel] - x0)).astype(np.uint32)
        return self.texels.subset(sel, local, w, w), sel

    def mapping(self, texels=None):
        """Mesh mapping of a texel sub-set, or of the whole map (texel set
        rasterised on the device, nothing uploaded)."""
        if texels is None:
            mm = MeshMapping().init_from_mesh(self.lo, self.dirs, self.size, self.size, self.scene, self.cull)
        else:
            mm = MeshMapping().init(texels, self.scene, self.cull)
        mm.run_all()
        return mm

    def solvers(self, mm, S, seed=1234):
        """AO and bent-normal solvers on the same rays, paired: one traversal
        computes both (results identical to two passes).  The seeded table
        is the process's shared copy (computed once per sample count and
        seed)."""
        ao = AmbientOcclusionSolver(S, self.mind, self.maxd, seed=seed).init(None, mm)
        bn = BentNormalsSolver(S, self.mind, self.maxd, tangentSpace=False, seed=seed).init(None, mm)
        ao.pair(bn)
        return ao, bn

    def gbuffers(self, mm):
        pos = PositionSolver().init(None, mm)
        nrm = NormalsSolver(tangentSpace=False).init(None, mm)
        hgt = HeightSolver(normalizeOutput=False).init(None, mm)
        for s in (pos, nrm, hgt):
            s.run_all()
        return pos, nrm, hgt

    def gather(self, mm, ao, bn, g, image):
        """Network input from the solvers; g: the (position, normals, height)
        solvers of gbuffers(), or None to compute their values from the
        mapping inside the gather (the same bits, no separate passes)."""
        hs = (None, None, None) if g is None else tuple(s.handle for s in g)
        c = np.ascontiguousarray(self.center, np.float32)
        check(lib().fornos_denoise_gather(mm.handle, ao.handle, bn.handle, *hs,
                                          c.ctypes.data_as(C.c_void_p), 2.0 / self.diag, 20.0 / self.diag,
                                          _p(image)))
        return image


def training_pair(job, x0, y0, w, seed):
    """(input [w, w, 12], target [w, w, 4]) device tensors for one window."""
    torch = _torch()
    sub, _ = job.window(x0, y0, w)
    mm = job.mapping(sub)
    g = job.gbuffers(mm)
    inp = torch.empty((w, w, IN), dtype=torch.float32, device="cuda")
    tgt = torch.empty((w, w, IN), dtype=torch.float32, device="cuda")
    lo_ao, lo_bn = job.solvers(mm, LOW_S, seed)
    lo_ao.run_all()
    lo_bn.run_all()
    job.gather(mm, lo_ao, lo_bn, g, inp)
    hi_ao, hi_bn = job.solvers(mm, HIGH_S, seed + 1)
    hi_ao.run_all()
    hi_bn.run_all()
    job.gather(mm, hi_ao, hi_bn, g, tgt)
    return inp, tgt[..., :OUT].contiguous()


def readback_buffers(n):
    """Page-locked host arrays (AO [n], bent normals [n, 3]) for the results of
    a timed bake of at most n texels, allocated before the timing starts, so
    that both bakes of a comparison read back with the same plain DMA."""
    torch = _torch()
    return (torch.empty(n, dtype=torch.float32, pin_memory=True),
            torch.empty((n, 3), dtype=torch.float32, pin_memory=True))


def plain_bake(job, S=HIGH_S, seed=PLAIN_SEED, host=None):
    """Mapping + AO + bent normals at S samples; host results and wall time.
    host: readback_buffers to read into (default: new arrays)."""
    t0 = time.perf_counter()
    mm = job.mapping()
    ao, bn = job.solvers(mm, S, seed)
    ao.run_all()
    bn.run_all()
    if host is None:
        a, b = ao.getResults(), bn.getResults()
    else:
        n = mm.count
        a, b = ao.getResults(out=host[0][:n].numpy()), bn.getResults(out=host[1][:n].numpy())
    t = time.perf_counter() - t0
    return a, b, t, mm


def denoised_bake(job, net, seed=11, tile=32, z=4.0, slack=0.004, fallback=True, host=None, masks=False):
    """Mapping + 16-sample AO / bent normals + G-buffers -> network -> per-tile
    check -> plain 1024-sample bake of the failing tiles.  Host results, wall
    time and a breakdown.  host: readback_buffers to read into.  masks: also
    return the failing blocks ("fail_blocks", [tiles, tiles] bool) and the
    texels that fell back ("fallback_mask", per texel) in the breakdown."""
    torch = _torch()
    W = job.size
    t0 = time.perf_counter()
    mm = job.mapping()
    ao, bn = job.solvers(mm, LOW_S, seed)
    ao.run_all()
    bn.run_all()
    inp = torch.empty((1, W, W, IN), dtype=torch.float32, device="cuda")
    job.gather(mm, ao, bn, None, inp)  # position, normal and distance from the mapping
    t1 = time.perf_counter()
    out, net_ms = net.forward(inp)
    t2 = time.perf_counter()
    n = mm.count
    d_ao = torch.empty(n, dtype=torch.float32, device="cuda")
    d_bn = torch.empty((n, 3), dtype=torch.float32, device="cuda")
    check(lib().fornos_denoise_scatter(mm.handle, _p(out), _p(d_ao), _p(d_bn)))
    nt = (W + tile - 1) // tile
    fail = torch.empty(nt * nt, dtype=torch.int32, device="cuda")
    check(lib().fornos_denoise_check_tiles(_p(inp), _p(out), W, W, tile, LOW_S, z, slack, _p(fail)))
    if host is None:
        a, b = d_ao.cpu().numpy(), d_bn.cpu().numpy()
    else:
        a, b = host[0][:n], host[1][:n]
        a.copy_(d_ao)
        b.copy_(d_bn)
        a, b = a.numpy(), b.numpy()
    bad = fail.cpu().numpy().reshape(nt, nt).astype(bool)
    t3 = time.perf_counter()
    n_fb = 0
    if fallback and bad.any():
        # the plain bake itself on the failing blocks' texels: same mapping,
        # same table, each texel its own sample permutation, so these texels
        # equal plain_bake's bit for bit
        idx = job.texels.a["indices"].astype(np.int64)
        sel = bad[(idx // W) // tile, (idx % W) // tile]
        n_fb = int(sel.sum())
        fao, fbn = job.solvers(mm, HIGH_S, PLAIN_SEED)
        fao.run_texels(np.flatnonzero(sel))
        a[sel] = fao.getResults()[sel]
        b[sel] = fbn.getResults()[sel]
    t4 = time.perf_counter()
    info = {"bake16_s": t1 - t0, "network_s": t2 - t1, "network_ms": net_ms, "check_s": t3 - t2,
            "fallback_s": t4 - t3, "fallback_tiles": int(bad.sum()), "tiles": int(nt * nt), "fallback_texels": n_fb}
    if masks:
        idx = job.texels.a["indices"].astype(np.int64)
        info["fail_blocks"] = bad
        info["fallback_mask"] = (bad[(idx // W) // tile, (idx % W) // tile] if fallback
                                 else np.zeros(len(idx), bool))
    return a, b, t4 - t0, info


def errors(ao, bn, ref_ao, ref_bn, hit):
    """AO mean absolute error and bent-normal mean angular error (degrees)
    over mapped texels; texels whose reference bent normal is undefined
    (every sample occluded) are left out of the angle."""
    ae = float(np.mean(np.abs(ao[hit] - ref_ao[hit])))
    ok = hit & np.all(np.isfinite(ref_bn), axis=1) & np.all(np.isfinite(bn), axis=1)
    rb = ref_bn[ok] / np.maximum(np.linalg.norm(ref_bn[ok], axis=1, keepdims=True), 1e-30)
    db = bn[ok] / np.maximum(np.linalg.norm(bn[ok], axis=1, keepdims=True), 1e-30)
    ang = np.degrees(np.arccos(np.clip(np.sum(rb * db, axis=1), -1.0, 1.0)))
    return ae, float(np.mean(ang)) if len(ang) else 0.0


# ---------------------------------------------------------------------------
# training
# ---------------------------------------------------------------------------
def _specs(seeds):
    from .batch import job_spec
    return [job_spec(s) for s in seeds]


def make_dataset(seeds, windows=2, wsize=512, size=2048, log=print):
    torch = _torch()
    data = []
    for k, spec in enumerate(_specs(seeds)):
        t = time.perf_counter()
        job = MeshJob(spec, size)
        rng = np.random.default_rng(spec["seed"])
        idx = job.texels.a["indices"].astype(np.int64)
        for wi in range(windows):
            # a window around a random texel, so it is not empty
            q = idx[int(rng.integers(len(idx)))] if len(idx) else 0
            x0 = int(np.clip(q % size - wsize // 2, 0, size - wsize)) // 8 * 8
            y0 = int(np.clip(q // size - wsize // 2, 0, size - wsize)) // 8 * 8
            inp, tgt = training_pair(job, x0, y0, wsize, 100 * spec["seed"] + wi)
            data.append((inp, tgt))
        log("# data %d/%d seed %d %s %d hi tris: %.1f s" % (k + 1, len(seeds), spec["seed"], spec["family"],
                                                            job.high_tris, time.perf_counter() - t))
    torch.cuda.synchronize()
    return data


def train(net, data, steps=3000, batch=16, crop=128, lr=1e-3, seed=0, log=print):
    torch = _torch()
    g = torch.Generator(device="cpu").manual_seed(seed)
    hist = []
    t0 = time.perf_counter()
    for it in range(steps):
        xs, ys = [], []
        for _ in range(batch):
            i = int(torch.randint(len(data), (1,), generator=g))
            inp, tgt = data[i]
            H = inp.shape[0]
            y0 = int(torch.randint(H - crop + 1, (1,), generator=g)) // 8 * 8
            x0 = int(torch.randint(H - crop + 1, (1,), generator=g)) // 8 * 8
            a, b = inp[y0:y0 + crop, x0:x0 + crop], tgt[y0:y0 + crop, x0:x0 + crop]
            f = int(torch.randint(8, (1,), generator=g))
            if f & 1:
                a, b = a.flip(0), b.flip(0)
            if f & 2:
                a, b = a.flip(1), b.flip(1)
            if f & 4:
                a, b = a.transpose(0, 1), b.transpose(0, 1)
            xs.append(a)
            ys.append(b)
        x = torch.stack(xs).contiguous()
        y = torch.stack(ys).contiguous()
        warm = min(1.0, (it + 1) / 100.0)
        rate = lr * warm * (0.5 * (1 + math.cos(math.pi * it / steps)) * 0.99 + 0.01)
        loss = net.backward(x, y)
        net.adam(rate)
        hist.append(loss)
        if (it + 1) % 100 == 0 or it == 0:
            log("# step %d loss %.5f (mean of last 100 %.5f) lr %.2e %.1f s" % (
                it + 1, loss, float(np.mean(hist[-100:])), rate, time.perf_counter() - t0))
    return hist


def conv_flops(net, n, h, w):
    """Algorithmic flops of one forward pass: 2 * 9 * cin * cout per output
    pixel of every 3x3 layer (cin of the first layer = the 12 real input
    channels) plus the 1x1 head."""
    total = 0
    for i, L in enumerate(net.layers()):
        level = 0
        # encoder layers 2l, 2l+1 sit at level l; decoder levels run L-2 .. 0
        if i < 2 * net.levels:
            level = i // 2
        else:
            level = net.levels - 2 - (i - 2 * net.levels) // 2
        cin = IN if i == 0 else L["c1"] + L["c2"]
        total += 2 * 9 * cin * L["cout"] * n * (h >> level) * (w >> level)
    return total + 2 * net.base * OUT * n * h * w


def time_network(net, n, h, w, reps=10):
    """Forward and training-step times on random inputs."""
    torch = _torch()
    x = torch.rand((n, h, w, IN), dtype=torch.float32, device="cuda")
    x[..., MASK] = 1.0
    y = torch.rand((n, h, w, OUT), dtype=torch.float32, device="cuda")
    for _ in range(3):  # warm-up (allocations, clocks)
        net.forward(x)
    fw = [net.forward(x)[1] for _ in range(reps)]
    net.backward(x, y)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        net.backward(x, y)
        net.adam(1e-5)
    step_ms = (time.perf_counter() - t) / reps * 1e3
    fl = conv_flops(net, n, h, w)
    fms = float(np.median(fw))
    return {"n": n, "h": h, "w": w, "forward_ms": fms, "forward_tflops": fl / (fms * 1e-3) / 1e12,
            "train_step_ms": step_ms, "train_tflops": 3 * fl / (step_ms * 1e-3) / 1e12, "forward_flops": fl}


def _seed_range(s):
    a, b = s.split(":")
    return list(range(int(a), int(b)))


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(prog="python -m fornos_amd.denoise")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--seeds", default="1000:1200")
    t.add_argument("--windows", type=int, default=2)
    t.add_argument("--steps", type=int, default=3000)
    t.add_argument("--batch", type=int, default=16)
    t.add_argument("--crop", type=int, default=128)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--init", default=None, help="start from these weights (fine-tuning)")
    t.add_argument("--out", default=WEIGHTS)
    tm = sub.add_parser("time")
    tm.add_argument("--weights", default=None)
    tm.add_argument("--shape", default="1x2048x2048", help="n x h x w")
    tm.add_argument("--reps", type=int, default=10)
    e = sub.add_parser("eval")
    e.add_argument("--weights", default=WEIGHTS)
    e.add_argument("--seeds", default="2000:2020")
    e.add_argument("--size", type=int, default=2048)
    e.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    log = lambda s: print(s, file=sys.stderr, flush=True)
    if args.cmd == "train":
        t0 = time.perf_counter()
        data = make_dataset(_seed_range(args.seeds), args.windows, log=log)
        t1 = time.perf_counter()
        net = Denoiser.load(args.init) if args.init else Denoiser(4, 32, seed=1)
        hist = train(net, data, args.steps, args.batch, args.crop, lr=args.lr, log=log)
        t2 = time.perf_counter()
        meta = {"seeds": args.seeds, "windows": args.windows, "init": args.init, "steps": args.steps, "batch": args.batch,
                "crop": args.crop,
                "lr": args.lr, "data_s": t1 - t0, "train_s": t2 - t1, "final_loss": float(np.mean(hist[-100:]))}
        net.save(args.out, meta)
        print(json.dumps(meta), flush=True)
        return 0
    if args.cmd == "time":
        net = Denoiser.load(args.weights) if args.weights else Denoiser(4, 32, seed=1)
        n, h, w = (int(v) for v in args.shape.split("x"))
        print(json.dumps(time_network(net, n, h, w, args.reps)), flush=True)
        return 0
    torch = _torch()
    net = Denoiser.load(args.weights)
    net.forward(torch.zeros((1, args.size, args.size, IN), dtype=torch.float32, device="cuda"))  # allocations
    rows = []
    plain_host, dn_host = readback_buffers(args.size * args.size), readback_buffers(args.size * args.size)
    for spec in _specs(_seed_range(args.seeds)):
        job = MeshJob(spec, args.size)
        ra, rb, tp, mm = plain_bake(job, host=plain_host)
        _, tidx = mm.read()
        hit = tidx != 0xFFFFFFFF
        a16, b16, _, _ = plain_bake(job, LOW_S, seed=11)
        da, db, td, info = denoised_bake(job, net, host=dn_host)
        e16 = errors(a16, b16, ra, rb, hit)
        edn = errors(da, db, ra, rb, hit)
        row = {"seed": spec["seed"], "family": spec["family"], "high_tris": job.high_tris,
               "texels": int(hit.sum()), "plain_s": tp, "denoised_s": td, "speedup": tp / td,
               "ao_mae_16": e16[0], "bn_deg_16": e16[1], "ao_mae": edn[0], "bn_deg": edn[1], **info}
        rows.append(row)
        log(json.dumps(row))
        del job
    summary = {"metric": "denoised vs plain bake time at AO MAE <= 0.01", "jobs": len(rows),
               "plain_s": sum(r["plain_s"] for r in rows), "denoised_s": sum(r["denoised_s"] for r in rows),
               "ao_mae": float(np.mean([r["ao_mae"] for r in rows])),
               "bn_deg": float(np.mean([r["bn_deg"] for r in rows])),
               "ao_mae_max": float(np.max([r["ao_mae"] for r in rows])),
               "bn_deg_max": float(np.max([r["bn_deg"] for r in rows])),
               "fallback_tiles": int(sum(r["fallback_tiles"] for r in rows)),
               "tiles": int(sum(r["tiles"] for r in rows))}
    summary["speedup"] = summary["plain_s"] / summary["denoised_s"]
    print(json.dumps(summary), flush=True)
    if args.out:
        with open(args.out, "w") as f:
            json.dump({"summary": summary, "rows": rows}, f, indent=1)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
