"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Same seeded inputs on both sides; the oracle builds its own BVH and flatten
(restated from Src/bvh.cpp and Src/meshmapping.cpp:62-125) and runs the
restated shaders.  Integer and index results must be identical; float results
are compared bit for bit, because both sides follow the same operation order
with IEEE-correct division and square root and no FMA contraction.
"""
import numpy as np
import pytest

import fornos_amd as fa
import oracle as orc
from bakes import MISS, Bake, same_bits

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sphere(gpu):
    return Bake("sphere512")


@pytest.fixture(scope="module")
def field(gpu):
    return Bake(fa.configs.scaled("heightfield2048", size=256, high_quads=160))


def test_scene_matches_oracle_flatten(sphere):
    nodes, ids = sphere.scene.reference_nodes()
    assert np.array_equal(nodes, sphere.nodes)
    assert np.array_equal(ids, sphere.stri)
    sp, sn = sphere.scene.soup()
    assert same_bits(sp, sphere.spos) and same_bits(sn, sphere.snrm)


@pytest.mark.parametrize("which", ["sphere", "field"])
def test_meshmapping_culling_matches_oracle(which, request):
    b = request.getfixturevalue(which)
    coords, tidx = b.oracle_map(cull=True)
    n = len(tidx)
    miss = int(np.sum(tidx == MISS))
    print("%s: %d texels, %d misses, mapping %.3f ms" % (which, n, miss, b.ms))
    assert np.array_equal(b.tidx, tidx)
    assert same_bits(b.coords, coords)
    assert miss < n // 10


def test_meshmapping_nocull_matches_oracle(sphere):
    mm = fa.MeshMapping().init(sphere.prep["texels"], sphere.scene, False)
    mm.run_all()
    c, t = mm.read()
    oc, ot = sphere.oracle_map(cull=False)
    assert np.array_equal(t, ot)
    assert same_bits(c, oc)


def test_meshmapping_steps_equal_run_all(field):
    mm = fa.MeshMapping().init(field.prep["texels"], field.scene, True)
    steps = 0
    while not mm.runStep():
        steps += 1
    assert mm.progress() == 1.0
    c, t = mm.read()
    assert np.array_equal(t, field.tidx) and same_bits(c, field.coords)


def test_mapping_hits_agree_with_brute_force(sphere):
    # BVH-independent: the closest hit over all triangles, no tree.
    sel = np.arange(0, len(sphere.tidx), 997)
    bc, bt = orc.map_bruteforce(sphere.tex["positions"][sel], sphere.tex["directions"][sel], sphere.spos, True)
    gt = sphere.tidx[sel]
    ids_gpu = np.where(gt == MISS, -1, sphere.stri[np.minimum(gt, len(sphere.stri) * 3 - 1) // 3].astype(np.int64))
    ids_bf = np.where(bt == MISS, -1, sphere.stri[np.minimum(bt, len(sphere.stri) * 3 - 1) // 3].astype(np.int64))
    assert np.mean(ids_gpu == ids_bf) > 0.999
    ok = (gt != MISS) & (bt != MISS)
    diag = 2 * np.sqrt(3) * 1.2
    assert np.max(np.abs(sphere.coords[sel][ok, 0] - bc[ok, 0])) <= 1e-5 * diag


def test_height_position_normals_match_oracle(sphere, field):
    for b, tangent in ((sphere, False), (field, True)):
        h = fa.HeightSolver(True, 0.0).init(b.prep["texels"], b.mm)
        h.run_all()
        assert same_bits(h.getResults(), orc.height(b.coords))
        p = fa.PositionSolver().init(b.prep["texels"], b.mm)
        p.run_all()
        assert same_bits(p.getResults(), orc.texel_attr(0, b.coords, b.tidx, b.spos))
        nsv = fa.NormalsSolver(tangent).init(b.prep["texels"], b.mm)
        nsv.run_all()
        expect = orc.texel_attr(1, b.coords, b.tidx, b.snrm)
        if tangent:
            hit = b.tidx != MISS
            expect[hit] = orc.tangentspace(b.tex["normals"], b.tex["tangents"], b.tex["bitangents"], expect)[hit]
        assert same_bits(nsv.getResults(), expect)


def _ranges(n, k=3, width=1024):
    starts = np.linspace(0, max(0, n - width), k).astype(int)
    return [(int(s), int(min(n, s + width))) for s in starts]


def _hemi_check(b, solver, kind, S, mind, maxd, tangent=False, width=1024):
    res = solver.getResults()
    cnt = solver.counts()
    table = solver.table
    for lo, hi in _ranges(len(b.tidx), width=width):
        out, oc = orc.hemisphere(kind, S, mind, maxd, table, b.coords, b.tidx, b.spos, b.snrm, b.nodes, lo, hi)
        if kind == orc.KIND_BN and tangent:
            hit = b.tidx[lo:hi] != MISS
            out[hit] = orc.tangentspace(b.tex["normals"], b.tex["tangents"], b.tex["bitangents"], out, begin=lo)[hit]
        assert np.array_equal(cnt[lo:hi], oc), "per-texel ray counts differ in [%d,%d)" % (lo, hi)
        assert same_bits(res[lo:hi], out), "values differ in [%d,%d)" % (lo, hi)


def test_ao_matches_oracle_sphere(sphere):
    p = sphere.cfg["maps"]["ao"]
    s = fa.AmbientOcclusionSolver(p["sampleCount"], p["minDistance"], p["maxDistance"]).init(
        sphere.prep["texels"], sphere.mm)
    ms = s.run_all()
    print("sphere AO-32: %.3f ms" % ms)
    _hemi_check(sphere, s, orc.KIND_AO, p["sampleCount"], p["minDistance"], p["maxDistance"], width=2048)
    ao = s.getResults()
    assert np.all((ao >= 0) & (ao <= 1))


@pytest.mark.parametrize("kind", ["ao", "bentnormals", "thickness"])
def test_hemisphere_matches_oracle_field(field, kind):
    p = field.cfg["maps"][kind]
    cls = {"ao": fa.AmbientOcclusionSolver, "bentnormals": fa.BentNormalsSolver,
           "thickness": fa.ThicknessSolver}[kind]
    kw = {"tangentSpace": p["tangentSpace"]} if kind == "bentnormals" else {}
    s = cls(p["sampleCount"], p["minDistance"], p["maxDistance"], **kw).init(field.prep["texels"], field.mm)
    s.run_all()
    k = {"ao": orc.KIND_AO, "bentnormals": orc.KIND_BN, "thickness": orc.KIND_THICK}[kind]
    _hemi_check(field, s, k, p["sampleCount"], p["minDistance"], p["maxDistance"],
                tangent=p.get("tangentSpace", False), width=256)


def test_hemisphere_steps_equal_run_all(field):
    p = field.cfg["maps"]["thickness"]
    a = fa.ThicknessSolver(p["sampleCount"], p["minDistance"], p["maxDistance"]).init(field.prep["texels"], field.mm)
    a.run_all()
    b = fa.ThicknessSolver(p["sampleCount"], p["minDistance"], p["maxDistance"]).init(field.prep["texels"], field.mm)
    n = 0
    while not b.runStep():
        n += 1
    assert n >= 1
    assert same_bits(a.getResults(), b.getResults())


@pytest.mark.parametrize("S", [1, 7, 64, 300, 1024])
def test_ao_sample_counts(field, S):
    s = fa.AmbientOcclusionSolver(S, 0.01, 0.2).init(field.prep["texels"], field.mm)
    s.run_all()
    res, cnt = s.getResults(), s.counts()
    out, oc = orc.hemisphere(orc.KIND_AO, S, 0.01, 0.2, s.table, field.coords, field.tidx, field.spos,
                             field.snrm, field.nodes, 0, 128)
    assert np.array_equal(cnt[:128], oc) and same_bits(res[:128], out)


def test_solvers_on_tables_that_differ_in_one_sample_each(field):
    # the library keeps the ray-ordered copies of recent sample tables; two
    # tables that differ only in the last sample of every permutation (that
    # sample mirrored through the disk's centre) each get their own
    S = 32
    t1 = fa.sample_table(S, seed=5)[0]
    t2 = t1.copy()
    last = t2.reshape(-1, S, 3)[:, -1]
    last[:, :2] *= -1.0
    a = field.solver("ao", sampleCount=S, table=t1)
    b = field.solver("ao", sampleCount=S, table=t2)
    c = field.solver("ao", sampleCount=S, table=t1)
    for s in (a, b, c):
        s.run_all()
    assert np.array_equal(a.counts(), c.counts()) and same_bits(a.getResults(), c.getResults())
    assert not np.array_equal(a.counts(), b.counts())
    n = len(field.tidx)
    assert field.check_texels("ao", b, np.r_[0:1024, n // 2:n // 2 + 1024]) == 2048


@pytest.mark.parametrize("kind", ["ao", "bentnormals", "thickness"])
def test_launch_shape_and_stack_depth_do_not_change_results(field, kind):
    # the stack depth only decides which rays are redone on the stackless
    # tree; the launch shape only how rays are handed out
    p = field.cfg["maps"][kind]
    cls = {"ao": fa.AmbientOcclusionSolver, "bentnormals": fa.BentNormalsSolver,
           "thickness": fa.ThicknessSolver}[kind]
    s = cls(p["sampleCount"], p["minDistance"], p["maxDistance"]).init(field.prep["texels"], field.mm)
    assert s.get_launch() == (8, 32, 16) and s.get_ray_order() == s.RAY_ORDER_AUTO
    # AUTO: Morton order for rays at least a quarter of the scene box's diagonal long
    nodes, _ = field.scene.reference_nodes()
    box = nodes[0, :6].view(np.float32).astype(np.float64)
    long_rays = p["maxDistance"] >= 0.25 * np.float32(np.linalg.norm(box[3:] - box[:3]))
    assert s.effective_ray_order() == (s.RAY_ORDER_MORTON if long_rays else s.RAY_ORDER_TEXELS)
    default_tile = max(1 if p["sampleCount"] >= 256 else 256 // p["sampleCount"], 2 if kind == "thickness" else 1)
    assert s.get_tile() == default_tile
    s.run_all()
    ref = s.getResults().view(np.uint32).copy()
    live = int(np.sum(field.tidx != MISS))
    for k, shape in enumerate([(8, 16, 24), (8, 16, 32), (1, 1, 16), (3, 64, 24), (16, 7, 32)]):
        s.set_launch(*shape)
        s.set_tile((2, 1, 32, 0, 5)[k])  # texels per hand-out; 0: the default
        assert s.get_tile() == ((2, 1, 32, 0, 5)[k] or default_tile)
        s.set_ray_order((s.RAY_ORDER_TEXELS, s.RAY_ORDER_MORTON, s.RAY_ORDER_AUTO)[k % 3])
        assert s.effective_ray_order() in (s.RAY_ORDER_TEXELS, s.RAY_ORDER_MORTON)
        if k % 3 < 2:
            assert s.effective_ray_order() == k % 3
        assert s.get_launch() == shape
        s.reset()
        s.run_all()
        assert np.array_equal(s.getResults().view(np.uint32), ref), shape
        w = s.count_work()
        assert w["rays"] == live * p["sampleCount"]
        assert w["node_tests"] > w["rays"] and w["tri_tests"] > 0
        assert 0 <= w["stack_overflows"] <= w["rays"]
        assert np.array_equal(s.getResults().view(np.uint32), ref), shape
    with pytest.raises(fa.FornosError):
        s.set_launch(8, 16, 20)
    with pytest.raises(fa.FornosError):
        s.set_ray_order(3)
    with pytest.raises(fa.FornosError):
        s.set_tile(33)


@pytest.mark.parametrize("maxd", [0.05, 0.3])
def test_paired_ao_bn_equal_separate_passes(field, maxd):
    # one traversal for AO and bent normals (kAOBN): AO counts hits at
    # t < maxD in reachable leaves, BN any hit there; small maxD leaves many
    # BN-only hits beyond maxD
    S, tex = 64, field.prep["texels"]
    ao0 = fa.AmbientOcclusionSolver(S, 0.01, maxd).init(tex, field.mm)
    bn0 = fa.BentNormalsSolver(S, 0.01, maxd, True).init(tex, field.mm)
    ao0.run_all()
    bn0.run_all()
    ao1 = fa.AmbientOcclusionSolver(S, 0.01, maxd).init(tex, field.mm)
    bn1 = fa.BentNormalsSolver(S, 0.01, maxd, True).init(tex, field.mm)
    ao1.pair(bn1)
    ms = ao1.run_all()
    assert bn1.run_all() < ms  # nothing left to trace
    assert same_bits(ao0.getResults(), ao1.getResults()) and np.array_equal(ao0.counts(), ao1.counts())
    assert same_bits(bn0.getResults(), bn1.getResults()) and np.array_equal(bn0.counts(), bn1.counts())
    # the BN-only hits exist: BN occlusion exceeds AO occlusion somewhere
    assert np.any(bn1.counts() > ao1.counts())
    out, oc = orc.hemisphere(orc.KIND_BN, S, 0.01, maxd, bn1.table, field.coords, field.tidx, field.spos,
                             field.snrm, field.nodes, 0, 96)
    assert np.array_equal(bn1.counts()[:96], oc)


def test_paired_steps_orders_and_rejections(field):
    S, tex = 32, field.prep["texels"]
    ref_ao = fa.AmbientOcclusionSolver(S, 0.01, 0.2).init(tex, field.mm)
    ref_bn = fa.BentNormalsSolver(S, 0.01, 0.2, True).init(tex, field.mm)
    ref_ao.run_all()
    ref_bn.run_all()
    a, b = ref_ao.getResults().view(np.uint32).copy(), ref_bn.getResults().view(np.uint32).copy()
    ao = fa.AmbientOcclusionSolver(S, 0.01, 0.2).init(tex, field.mm)
    bn = fa.BentNormalsSolver(S, 0.01, 0.2, True).init(tex, field.mm)
    ao.pair(bn)
    # run_step pacing: each AO step advances the BN solver as well
    while not ao.runStep():
        pass
    assert bn.runStep()  # already done
    assert np.array_equal(ao.getResults().view(np.uint32), a) and np.array_equal(bn.getResults().view(np.uint32), b)
    # BN first: it runs alone, then AO runs alone (no stale pair state)
    ao.reset()
    bn.run_all()
    ao.run_all()
    assert np.array_equal(ao.getResults().view(np.uint32), a) and np.array_equal(bn.getResults().view(np.uint32), b)
    # submit + poll pacing of the pair
    import ctypes as C
    from fornos_amd._abi import check, lib
    ao.reset()
    check(lib().fornos_solver_submit(ao.handle, 3))
    prog, done = C.c_float(), C.c_int()
    while True:
        check(lib().fornos_solver_poll(bn.handle, C.byref(prog), C.byref(done)))
        if done.value:
            break
    assert np.array_equal(ao.getResults().view(np.uint32), a) and np.array_equal(bn.getResults().view(np.uint32), b)
    ao.pair(None)
    # rays that differ cannot be paired
    for other in (fa.BentNormalsSolver(S, 0.01, 0.3).init(tex, field.mm),
                  fa.BentNormalsSolver(S + 1, 0.01, 0.2).init(tex, field.mm),
                  fa.BentNormalsSolver(S, 0.01, 0.2, seed=99).init(tex, field.mm)):
        with pytest.raises(fa.FornosError):
            ao.pair(other)
    with pytest.raises(fa.FornosError):  # the AO solver comes first
        check(lib().fornos_solver_pair(bn.handle, ao.handle))


def test_ao_tree_of_split_leaves(field):
    # the heightfield's long reference leaves gain from AO's own tree of
    # triangle groups; AO results through it equal the oracle's (the AO tests
    # above run on this scene), while BN and thickness keep whole leaves
    info = field.scene.info()
    assert info["ao_groups"] > 0 and info["sah_ao"] < 0.9 * info["sah"]
    s = fa.AmbientOcclusionSolver(48, 0.01, 0.2).init(field.prep["texels"], field.mm)
    s.run_all()
    w = s.count_work()
    out, oc = orc.hemisphere(orc.KIND_AO, 48, 0.01, 0.2, s.table, field.coords, field.tidx, field.spos,
                             field.snrm, field.nodes, 0, 256)
    assert np.array_equal(s.counts()[:256], oc) and same_bits(s.getResults()[:256], out)
    b = fa.BentNormalsSolver(48, 0.01, 0.2).init(field.prep["texels"], field.mm)
    b.run_all()
    wb = b.count_work()
    assert w["tri_tests"] < wb["tri_tests"]  # the groups cut AO's triangle tests


@pytest.mark.parametrize("kind", ["ao", "thickness", "paired"])
def test_autotune_picks_a_shape_and_keeps_results(field, kind):
    # fornos_solver_autotune times launch shapes on a prefix of the pass;
    # the shape it keeps changes nothing in the results
    S, tex = 32, field.prep["texels"]
    if kind == "thickness":
        make = lambda: fa.ThicknessSolver(S, 0.01, 0.5).init(tex, field.mm)
    else:
        make = lambda: fa.AmbientOcclusionSolver(S, 0.01, 0.2).init(tex, field.mm)
    ref = make()
    ref_bn = fa.BentNormalsSolver(S, 0.01, 0.2, True).init(tex, field.mm) if kind == "paired" else None
    ref.run_all()
    if ref_bn:
        ref_bn.run_all()
    s = make()
    bn = None
    if kind == "paired":
        bn = fa.BentNormalsSolver(S, 0.01, 0.2, True).init(tex, field.mm)
        s.pair(bn)
        with pytest.raises(fa.FornosError):
            bn.autotune(0.1)  # tuned through its AO solver
    assert s.get_tile() == 8  # 256 / 32 samples
    bpc, refill, stack, ms = s.autotune(0.1)
    assert s.get_launch() == (bpc, refill, stack)
    assert bpc in (4, 6, 8, 12) and refill in (16, 24, 32, 40, 48) and stack in (16, 24, 32) and ms > 0
    # the hand-out tile around its default, and an explicit ray order
    assert s.get_tile() in (4, 8, 16)
    assert s.get_ray_order() in (s.RAY_ORDER_TEXELS, s.RAY_ORDER_MORTON)
    assert s.effective_ray_order() == s.get_ray_order()
    s.reset()
    s.run_all()
    assert same_bits(s.getResults(), ref.getResults())
    if bn:
        assert bn.run_all() >= 0 and same_bits(bn.getResults(), ref_bn.getResults())
    with pytest.raises(fa.FornosError):
        s.autotune(0.0)
    with pytest.raises(fa.FornosError):
        fa.HeightSolver(True, 0.0).init(tex, field.mm).autotune(0.1)
    with pytest.raises(fa.FornosError):
        fa.HeightSolver(True, 0.0).init(tex, field.mm).set_tile(2)
    with pytest.raises(fa.FornosError):
        fa.HeightSolver(True, 0.0).init(tex, field.mm).get_tile()
