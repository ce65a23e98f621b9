"""Image export on the device against the oracle (oracle/export.py), on real
bakes: every quantised byte, every float and every dilated pixel identical;
and the files decode to those buffers."""
import numpy as np
import pytest

import fornos_amd as fa
from fornos_amd import export as fx
from oracle import export as ox

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bake(gpu):
    cfg = fa.configs.scaled("heightfield2048", size=192, high_quads=96)
    prep = fa.prepare(cfg)
    mm = fa.MeshMapping().init_from_mesh(prep["low"], prep["dirs"], cfg["size"], cfg["size"], prep["scene"], True)
    mm.run_all()
    idx = mm.read_texels()["indices"]
    out = {"mm": mm, "idx": idx, "w": cfg["size"]}
    m = cfg["maps"]
    ao = fa.AmbientOcclusionSolver(64, m["ao"]["minDistance"], m["ao"]["maxDistance"]).init(None, mm)
    bn = fa.BentNormalsSolver(64, 0.01, 0.2, tangentSpace=True).init(None, mm)
    th = fa.ThicknessSolver(32, 0.0, 0.1).init(None, mm)
    ht = fa.HeightSolver(True, 0.05).init(None, mm)
    nr = fa.NormalsSolver(True).init(None, mm)
    po = fa.PositionSolver().init(None, mm)
    for s in (ao, bn, th, ht, nr, po):
        s.run_all()
    out.update(ao=ao, bn=bn, th=th, ht=ht, nr=nr, po=po)
    return out


@pytest.mark.parametrize("name,rng,norm,dil", [("ao", (0, 0), True, 16), ("th", (0, 0), True, 16),
                                               ("ht", (0, 0.05), True, 16), ("ht", (0, 0.01), False, 3),
                                               ("ao", (0.5, 0.9), True, 0)])
def test_float_map_export_matches_oracle(bake, name, rng, norm, dil):
    s, w = bake[name], bake["w"]
    vals = s.getResults()
    rgb, mm = fx.float_rgb8(s, rng, norm, dil)
    ref, rmm = ox.float_rgb8(vals, bake["idx"], w, w, rng[0], rng[1], norm, dil)
    assert np.array_equal(rgb.reshape(-1), ref)
    assert mm == (float(rmm[0]), float(rmm[1]))
    f, _ = fx.float_f32(s, rng, norm)
    rf, _ = ox.float_f32(vals, bake["idx"], w, w, rng[0], rng[1], norm)
    assert np.array_equal(f.reshape(-1).view(np.uint32), rf.view(np.uint32))


@pytest.mark.parametrize("name,dil", [("bn", 16), ("nr", 16), ("nr", 0)])
def test_normal_export_matches_oracle(bake, name, dil):
    s, w = bake[name], bake["w"]
    rgb = fx.normal_rgb8(s, dil)
    assert np.array_equal(rgb.reshape(-1), ox.normal_rgb8(s.getResults(), bake["idx"], w, w, dil))


def test_vector_export_matches_oracle(bake):
    s, w = bake["po"], bake["w"]
    p = fx.vector_planes(s)
    assert np.array_equal(p.reshape(3, -1), ox.vector_planes(s.getResults(), bake["idx"], w, w))


def test_task_exports_write_decodable_files(bake, tmp_path):
    from PIL import Image
    w = bake["w"]
    for name, ext in (("ao", "png"), ("th", "tga"), ("bn", "png"), ("ht", "exr"), ("po", "exr")):
        s = bake[name]
        path = tmp_path / ("%s.%s" % (name, ext))
        fx.export_task(s, path)
        if ext == "exr":
            img = ox.read_exr(str(path))
            if name == "po":
                ref = ox.vector_planes(s.getResults(), bake["idx"], w, w)
                for k, c in enumerate("BGR"):
                    expect = ox.float_to_half(ref[k]).view(np.float16).astype(np.float32)
                    assert np.array_equal(img[c].reshape(-1), expect, equal_nan=True)
            else:
                p = s.params
                ref, _ = ox.float_f32(s.getResults(), bake["idx"], w, w, 0.0, p.max_distance, bool(p.normalize_output))
                expect = ox.float_to_half(ref).view(np.float16).astype(np.float32)
                assert np.array_equal(img["B"].reshape(-1), expect, equal_nan=True)
        else:
            back = np.asarray(Image.open(path).convert("RGB")).reshape(-1)
            if name == "bn":
                ref = ox.normal_rgb8(s.getResults(), bake["idx"], w, w, 16)
            else:
                ref, _ = ox.float_rgb8(s.getResults(), bake["idx"], w, w, 0, 0, True, 16)
            assert np.array_equal(back, ref), name


def test_export_overlapping_a_running_pass_equals_a_quiet_one(bake):
    # an export waits for its own solver's pass only: run a long AO pass on
    # the same scene while exporting the finished maps from another thread,
    # and the bytes equal the exports taken with nothing else running
    from concurrent.futures import ThreadPoolExecutor
    quiet = {k: fx.float_rgb8(bake[k], (0.0, 0.0), True, 16)[0] for k in ("ao", "th")}
    quiet["nr"] = fx.normal_rgb8(bake["nr"], 16)
    quiet["po"] = fx.vector_planes(bake["po"])
    busy = fa.AmbientOcclusionSolver(1024, 0.01, 0.2).init(None, bake["mm"])
    with ThreadPoolExecutor(max_workers=1) as pool:
        futs = {k: pool.submit(fx.float_rgb8, bake[k], (0.0, 0.0), True, 16) for k in ("ao", "th")}
        futs["nr"] = pool.submit(fx.normal_rgb8, bake["nr"], 16)
        futs["po"] = pool.submit(fx.vector_planes, bake["po"])
        busy.run_all()
        got = {k: f.result() for k, f in futs.items()}
    for k in ("ao", "th"):
        assert np.array_equal(got[k][0], quiet[k]), k
    assert np.array_equal(got["nr"], quiet["nr"])
    assert np.array_equal(got["po"].view(np.uint32), quiet["po"].view(np.uint32))
    # and a solver exported right after its own pass was queued (no wait in
    # between) gives its finished results
    s = fa.AmbientOcclusionSolver(64, 0.01, 0.2).init(None, bake["mm"])
    import ctypes as C
    from fornos_amd._abi import check, lib
    check(lib().fornos_solver_submit(s.handle, 4))
    first = fx.float_rgb8(s, (0.0, 0.0), True, 0)[0]
    assert np.array_equal(first, fx.float_rgb8(bake["ao"], (0.0, 0.0), True, 0)[0])
