"""Pins for the CPU oracle (CPU only).

The reference ships no tests, golden images or fixtures, its host C++ does
not build here without stand-in headers, and its GLSL cannot run here, so
each oracle function is pinned by what the mathematics fixes: closed forms,
hand-derived known answers, brute force over all triangles, and invariance
of BVH-independent results under different trees.  Each check is chosen so
that a plausible slip (a flipped sign or cross product, a dropped term, a
transposed frame, an off-by-one in the traversal links) fails it.

Function -> pins
  orc_vertex_normals            flat grid (exact +z), flipped winding (-z), octahedron
  orc_vertex_normals_aggressive octahedron with unshared corners (exact axes)
  orc_tangent_space             planar grid + planar UVs (exact +x, +y)
  orc_mapuv / orc_compress      inset quad (full coverage, analytic positions);
                                UV = 0 wrap of the reference's bound (empty set)
  orc_samples                   hand-derived S=4 table; unit length; E[z] = 2/3;
                                stratification of z^2
  orc_bvh_build                 leaf rules (max_tris, max_depth, unsplittable);
                                the cost's xy-twice / no-yz area term decides a
                                split that true SAH decides the other way;
                                structural invariants on a large mesh
  orc_meshmapping               plane offsets (t = delta, analytic barycentrics),
                                culling of flipped faces, tie rules, brute force
  orc_hemisphere (AO/BN/thick)  open plane (AO 1), enclosing sphere (AO 0),
                                wall (blocked fraction from the table), slab
                                thickness h/cos, BN of an open plane, NaN BN when
                                fully occluded, invariance under 4 different BVHs;
                                hits beyond maxD: counted by BN and returned by
                                thickness in a leaf the rays enter before maxD,
                                not in one entered beyond it (mutant oracles with
                                a t < maxD test in BN, or a clamp in thickness,
                                fail this pin)
  orc_tangentspace              identity and rotated orthonormal frames
  orc_height / orc_texel_attr   closed forms on a plane
"""
import math

import numpy as np
import pytest

import oracle as orc
from meshes import grid, icosphere, merge, octahedron, quad, soup_mesh, wall_x

MISS = orc.MISS
FMAX = float(orc.FLT_MAX)


def scene(mesh, max_tris=8, max_depth=8192):
    nodes, sp, sn, st = orc.bvh_build(mesh, max_tris, max_depth)
    return {"nodes": nodes, "spos": sp, "snrm": sn, "stri": st}


def plane_texels(n, z=0.0, d=(0, 0, 1), span=0.8):
    xs = np.linspace(-span / 2, span / 2, n, dtype=np.float32)
    X, Y = np.meshgrid(xs, xs)
    p = np.stack([X.ravel(), Y.ravel(), np.full(X.size, z, np.float32)], 1).astype(np.float32)
    return p, np.tile(np.float32(d), (len(p), 1))


# ---------------------------------------------------------------- normals --
def test_vertex_normals_flat_and_flipped():
    g = grid(6)
    n, vn = orc.vertex_normals(g["positions"], g["vertices"], g["triangles"])
    assert np.array_equal(n, np.tile(np.float32([0, 0, 1]), (len(n), 1)))
    assert np.array_equal(vn, g["vertices"][:, 0])
    flipped = g["triangles"][:, ::-1].copy()
    n2, _ = orc.vertex_normals(g["positions"], g["vertices"], flipped)
    assert np.array_equal(n2, -n)


def test_vertex_normals_octahedron():
    o = octahedron(shared=True)
    n, _ = orc.vertex_normals(o["positions"], o["vertices"], o["triangles"])
    assert np.allclose(n, o["positions"], atol=1e-7)


def test_aggressive_normals_merge_unshared_corners():
    o = octahedron(shared=False)
    # per position index: every corner keeps its own face normal
    n, _ = orc.vertex_normals(o["positions"], o["vertices"], o["triangles"])
    assert np.allclose(np.abs(n), 1 / math.sqrt(3), atol=1e-7)
    # by position value: the six apexes, exactly on the axes
    a, vn = orc.vertex_normals_aggressive(o["positions"], o["vertices"], o["triangles"])
    assert len(a) == 6
    assert np.array_equal(a[vn], o["positions"][o["vertices"][:, 0]])


def test_tangent_space_planar():
    g = grid(4)
    t, b = orc.tangent_space(g["positions"], g["normals"], g["texcoords"], g["vertices"], g["triangles"])
    assert np.allclose(t, [1, 0, 0], atol=1e-7) and np.allclose(b, [0, 1, 0], atol=1e-7)


# ------------------------------------------------------------- texel set --
def test_mapuv_inset_quad_covers_every_texel():
    W = H = 64
    e = 1.0 / 4096
    tris = [[(0, 0, 0), (1, 0, 0), (1, 1, 0)], [(0, 0, 0), (1, 1, 0), (0, 1, 0)]]
    uvs = [[(e, e), (1 - e, e), (1 - e, 1 - e)], [(e, e), (1 - e, 1 - e), (e, 1 - e)]]
    m = soup_mesh(tris, uvs=uvs)
    mp = orc.mapuv(m, W, H)
    idx = orc.compress(mp["directions"])
    assert len(idx) == W * H
    x = (np.arange(W) + 0.5) / W
    u = (x - e) / (1 - 2 * e)                 # texel centre back to the quad's x
    expect = np.stack(np.meshgrid(u, u), -1).reshape(-1, 2)
    assert np.allclose(mp["positions"][:, :2], expect, atol=2e-6)
    assert np.allclose(mp["directions"], [0, 0, 1], atol=1e-7)


def test_mapuv_uv_zero_skips_triangles():
    # uint32_t(roundf(-0.5)) wraps to 0xFFFFFFFF and is clamped to W, so a
    # triangle touching u = 0 or v = 0 rasterises nothing (Src/compute.cpp:136-139)
    tris = [[(0, 0, 0), (1, 0, 0), (1, 1, 0)], [(0, 0, 0), (1, 1, 0), (0, 1, 0)]]
    uvs = [[(0, 0), (1, 0), (1, 1)], [(0, 0), (1, 1), (0, 1)]]
    mp = orc.mapuv(soup_mesh(tris, uvs=uvs), 32, 32)
    assert len(orc.compress(mp["directions"])) == 0


def test_mapuv_missing_texcoords():
    assert orc.mapuv(soup_mesh([[(0, 0, 0), (1, 0, 0), (1, 1, 0)]]), 8, 8) is None


# ----------------------------------------------------------- sample table --
def test_samples_known_answer():
    d = orc.samples(4, np.zeros((1, 2), np.uint32))
    s = math.sqrt
    expect = [(0, 0, 1), (-0.5, 0, s(0.75)), (0, s(0.5), s(0.5)), (0, -s(0.75), 0.5)]
    assert np.allclose(d, expect, atol=2e-7)


def test_samples_cosine_distribution():
    rng = np.random.default_rng(3)
    S, P = 32, 4096
    d = orc.samples(S, rng.integers(0, 2 ** 32, (P, 2), dtype=np.uint64).astype(np.uint32))
    assert np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-6)
    assert np.all(d[:, 2] >= 0)
    assert abs(d[:, 2].mean() - 2 / 3) < 2e-3          # E[cos] under cosine weighting
    # x = frac(i/S + c) is stratified: sorted 1 - z^2 steps by exactly 1/S
    z2 = np.sort(1 - d[:S, 2].astype(np.float64) ** 2)
    assert np.allclose(np.diff(z2), 1 / S, atol=1e-5)


# ------------------------------------------------------------------- BVH --
def _walk(nodes):
    """(parent of each node, leaves) from the skip links."""
    n = len(nodes)
    parent = [-1] * n

    def rec(i, par):
        parent[i] = par
        j = i + 1
        if nodes[i, 6] != nodes[i, 7] or nodes[i, 8] == i + 1:
            assert nodes[i, 8] == i + 1
            return
        while j < nodes[i, 8]:
            rec(j, i)
            j = nodes[j, 8]
    rec(0, -1)
    return parent


def test_bvh_leaf_rules():
    g = grid(8)
    s = scene(g, max_tris=8, max_depth=0)
    assert len(s["nodes"]) == 1 and s["nodes"][0, 7] == 3 * len(g["triangles"])
    same = soup_mesh([[(0, 0, 0), (1, 0, 0), (0, 1, 0)]] * 5)
    s = scene(same, max_tris=1)
    assert len(s["nodes"]) == 1 and s["nodes"][0, 6] == 0 and s["nodes"][0, 7] == 15
    s = scene(g, max_tris=1)
    assert np.all(s["nodes"][:, 7] - s["nodes"][:, 6] <= 3)


def test_bvh_cost_area_term_decides_split():
    # Four clusters of four thin, z-tilted triangles at x in {0, 0.5},
    # y in {0, 1}.  True SAH splits along y (side areas ~2*0.5*h < 2*1*h);
    # the reference's 2(xy + xz + xy) has no yz term and splits along x.
    d, h = 0.001, 0.1
    tris = []
    for cx, cy in ((0, 0), (0.5, 0), (0, 1), (0.5, 1)):
        for k in range(4):
            z = 0.01 * k
            tris.append([(cx, cy, z), (cx + d, cy, z + h), (cx, cy + d, z + h / 2)])
    m = soup_mesh(tris)
    s = scene(m)
    nodes, stri = s["nodes"], s["stri"]
    assert len(nodes) == 3
    left = stri[nodes[1, 6] // 3: nodes[1, 7] // 3]
    assert sorted(left.tolist()) == [0, 1, 2, 3, 8, 9, 10, 11]   # the x = 0 clusters


def test_bvh_invariants_large():
    m = icosphere(1.0, 4)
    s = scene(m)
    nodes, stri = s["nodes"], s["stri"]
    assert sorted(stri.tolist()) == list(range(len(m["triangles"])))
    parent = _walk(nodes)
    lo = nodes[:, :3].view(np.float32)
    hi = nodes[:, 3:6].view(np.float32)
    sp = s["spos"][:, :3].reshape(-1, 3, 3)
    for i in range(len(nodes)):
        a, b = nodes[i, 6] // 3, nodes[i, 7] // 3
        if b > a:   # leaf: its triangles lie in its box up to the centre/size rounding
            tl, th = sp[a:b].min(axis=(0, 1)), sp[a:b].max(axis=(0, 1))
            assert np.all(lo[i] <= tl + 1e-6) and np.all(hi[i] >= th - 1e-6)
            assert b - a <= 8
        if parent[i] >= 0:
            p = parent[i]
            assert np.all(lo[p] <= lo[i] + 1e-6) and np.all(hi[p] >= hi[i] - 1e-6)
    # the soup holds the leaves' triangles in the order the stackless DFS
    # visits the leaves (node order): a smaller soup index is visited first,
    # which is how the wide mapping breaks ties in t
    leaves = [(nodes[i, 6], nodes[i, 7]) for i in range(len(nodes)) if nodes[i, 7] > nodes[i, 6]]
    assert leaves[0][0] == 0 and leaves[-1][1] == 3 * len(m["triangles"])
    assert all(a1 == b0 for (_, a1), (b0, _) in zip(leaves, leaves[1:]))


# ---------------------------------------------------------------- mapping --
@pytest.mark.parametrize("cull", [True, False])
def test_mapping_plane_offsets(cull):
    delta = np.float32(0.0625)
    p, d = plane_texels(9)
    for z, sign in ((delta, 1), (-delta, -1)):
        s = scene(soup_mesh(quad(float(z), 4.0)))
        c, t = orc.meshmapping(p, d, s["spos"], s["nodes"], cull)
        assert np.all(t != MISS)
        assert np.all(c[:, 0] == delta)
        # barycentrics reproduce the texel's xy on the hit triangle
        v = s["spos"][:, :3].reshape(-1, 3, 3)[t // 3]
        xy = (c[:, 1:, None] * v).sum(axis=1)
        assert np.allclose(xy[:, :2], p[:, :2], atol=1e-5)


def test_mapping_culls_back_faces():
    p, d = plane_texels(5)
    s = scene(soup_mesh(quad(0.1, 4.0, flip=True)))
    c, t = orc.meshmapping(p, d, s["spos"], s["nodes"], True)
    assert np.all(t == MISS) and np.all(c[:, 0] == FMAX)
    c, t = orc.meshmapping(p, d, s["spos"], s["nodes"], False)
    assert np.all(t != MISS) and np.allclose(c[:, 0], 0.1)


def test_mapping_tie_rules():
    p, d = plane_texels(5)
    # coincident copies: the earlier soup triangle wins
    s = scene(soup_mesh(quad(0.1, 4.0) + quad(0.1, 4.0)), max_tris=8)
    for cull in (True, False):
        _, t = orc.meshmapping(p, d, s["spos"], s["nodes"], cull)
        assert np.all(t < 6)
    # equal distances forward and backward: the forward hit wins
    s = scene(soup_mesh(quad(0.1, 4.0) + quad(-0.1, 4.0)))
    for cull in (True, False):
        c, t = orc.meshmapping(p, d, s["spos"], s["nodes"], cull)
        hitz = s["spos"][t, 2]
        assert np.all(hitz > 0) and np.allclose(c[:, 0], 0.1)


def test_axis_parallel_ray_through_zero_box_face_is_rejected():
    # RayAABB divides by the direction: a zero component with the origin on
    # a box face gives 0/0 = NaN, and GLSL min/max (y < x ? y : x) carry it
    # into both slab ends, so the box is rejected and the hit on the vertex
    # at (0,0,1) is lost, while a scan over all triangles finds it.  GLSL
    # leaves NaN in min/max undefined; this is the convention the oracle and
    # the kernels share, counted as a known class rather than pinned.
    s = scene(icosphere(1.0, 2))
    p, d = np.float32([[0, 0, 0.9]]), np.float32([[0, 0, 1]])
    c, t = orc.meshmapping(p, d, s["spos"], s["nodes"], True)
    bc, bt = orc.map_bruteforce(p, d, s["spos"], True)
    assert t[0] == MISS and bt[0] != MISS


def test_mapping_matches_brute_force():
    hi = icosphere(1.1, 3)
    rng = np.random.default_rng(1)
    dirs = rng.normal(size=(400, 3)).astype(np.float32)
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    p = (dirs * 0.97).astype(np.float32)
    for cull in (True, False):
        s = scene(hi, max_tris=4)
        c, t = orc.meshmapping(p, dirs, s["spos"], s["nodes"], cull)
        bc, bt = orc.map_bruteforce(p, dirs, s["spos"], cull)
        assert np.array_equal(t, bt) and np.array_equal(c, bc)
        assert np.all(t != MISS) and np.all((c[:, 0] > 0.05) & (c[:, 0] < 0.2))


# -------------------------------------------------------------- hemisphere --
def hemi(kind, s, coords, tidx, S=64, mind=1e-3, maxd=10.0, seed=5):
    rng = np.random.default_rng(seed)
    table = orc.samples(S, rng.integers(0, 2 ** 32, (16, 2), dtype=np.uint64).astype(np.uint32))
    out, cnt = orc.hemisphere(kind, S, mind, maxd, table, coords, tidx, s["spos"], s["snrm"], s["nodes"])
    return out, cnt, table


def mapped(hi, p, d, max_tris=8, cull=True):
    s = scene(hi, max_tris)
    c, t = orc.meshmapping(p, d, s["spos"], s["nodes"], cull)
    return s, c, t


def test_ao_open_plane_is_unoccluded():
    p, d = plane_texels(4, z=-0.05)
    s, c, t = mapped(soup_mesh(quad(0.0, 8.0)), p, d)
    ao, cnt, _ = hemi(orc.KIND_AO, s, c, t)
    assert np.all(ao == 1.0) and np.all(cnt == 0)


def test_ao_inside_enclosing_sphere_is_fully_occluded():
    inner, outer = icosphere(1.0, 3), icosphere(1.5, 3, inward=True)
    rng = np.random.default_rng(2)
    dirs = rng.normal(size=(50, 3)).astype(np.float32)
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    s, c, t = mapped(merge(inner, outer), (dirs * 0.9).astype(np.float32), dirs)
    assert np.allclose(c[:, 0], 0.1, atol=0.01)
    ao, cnt, _ = hemi(orc.KIND_AO, s, c, t, S=32)
    assert np.all(ao == 0.0) and np.all(cnt == 32)
    bn, bcnt, _ = hemi(orc.KIND_BN, s, c, t, S=32)
    assert np.all(np.isnan(bn)) and np.all(bcnt == 32)     # normalize(0), bentnormals_step2.comp:30


def test_ao_next_to_wall_blocks_the_samples_pointing_at_it():
    S = 256
    floor = quad(0.0, 200.0)
    m = soup_mesh(floor + wall_x(-0.01, 200.0))
    p = np.float32([[0.0, 0.0, -0.05]])
    d = np.float32([[0, 0, 1]])
    s, c, t = mapped(m, p, d)
    ao, cnt, table = hemi(orc.KIND_AO, s, c, t, S=S, maxd=1000.0)
    # frame of ao_step0.comp:49-50 for d = +z: ty = (0,1,0), tx = (-1,0,0);
    # a sample points at the wall (x < 0) iff its x component is > 0
    blocked = int(np.sum(table[:S, 0] > 0))
    assert abs(int(cnt[0]) - blocked) <= 2
    assert abs(ao[0] - (1 - cnt[0] / S)) == 0


def test_thickness_of_a_slab_is_h_over_cos():
    h = 0.05
    m = soup_mesh(quad(0.0, 200.0) + quad(-h, 200.0))
    p = np.float32([[0.0, 0.0, -0.01]])
    s, c, t = mapped(m, p, np.float32([[0, 0, 1]]))
    th, cnt, table = hemi(orc.KIND_THICK, s, c, t, S=128, mind=1e-4, maxd=50.0)
    assert cnt[0] == 128
    expect = np.mean(h / table[:128, 2].astype(np.float64))
    assert abs(th[0] - expect) < 1e-5 * expect


def test_bent_normal_of_open_plane():
    p, d = plane_texels(1, z=-0.05)
    s, c, t = mapped(soup_mesh(quad(0.0, 8.0)), p, d)
    bn, cnt, table = hemi(orc.KIND_BN, s, c, t, S=64)
    v = table[:64].astype(np.float64) * np.array([-1.0, 1.0, 1.0])   # tx = -x, ty = +y
    expect = v.sum(0) / np.linalg.norm(v.sum(0))
    assert cnt[0] == 0 and np.allclose(bn[0], expect, atol=1e-5)


def beyond_maxd(max_tris, S=64, maxd=1.0, library=None):
    """A floor at z = 0 (the texel maps onto it), a roof at z = 2 facing down
    and a basement at z = -2, both 200 wide, with maxD = 1: every roof or
    basement hit lies beyond maxD.  Returns (tree leaves' z sets, AO, BN,
    thickness results and counts, sample table)."""
    m = soup_mesh(quad(0.0, 8.0) + quad(2.0, 200.0, flip=True) + quad(-2.0, 200.0))
    s, c, t = mapped(m, np.float32([[0.0, 0.0, -0.05]]), np.float32([[0, 0, 1]]), max_tris=max_tris)
    assert t[0] != MISS and c[0, 0] == np.float32(0.05)
    leaves = [set(np.round(s["spos"][n[6]:n[7], 2], 3)) for n in s["nodes"] if n[7] > n[6]]
    out = {}
    for kind in (orc.KIND_AO, orc.KIND_BN, orc.KIND_THICK):
        rng = np.random.default_rng(9)
        table = orc.samples(S, rng.integers(0, 2 ** 32, (16, 2), dtype=np.uint64).astype(np.uint32))
        out[kind] = orc.hemisphere(kind, S, 1e-3, maxd, table, c, t, s["spos"], s["snrm"], s["nodes"],
                                   library=library)
    return leaves, out, table


def check_beyond_maxd(library=None):
    S, maxd = 64, 1.0
    # frame for the +z normal: tx = -x, ty = +y (ao_step0.comp:47-50); a
    # sample's ray reaches the 200-wide roof (up) or basement (down) at
    # height 2 when |2 x / z| and |2 y / z| stay below 100
    # (a) one leaf holds everything: its box contains the ray origins, so the
    # reference enters it at t <= 0 < maxD and tests every triangle in it
    leaves, out, table = beyond_maxd(8, S, maxd, library)
    assert len(leaves) == 1
    r = table[:S].astype(np.float64)
    reach = (np.abs(2 * r[:, 0] / r[:, 2]) < 100) & (np.abs(2 * r[:, 1] / r[:, 2]) < 100)
    assert np.all(np.abs(np.abs(2 * r[:, :2] / r[:, 2:]) - 100) > 1)  # no sample near the edge
    ao, ao_cnt = out[orc.KIND_AO]
    assert ao_cnt[0] == 0 and ao[0] == 1.0            # AO: hits at t >= maxD do not count
    bn, bn_cnt = out[orc.KIND_BN]
    assert bn_cnt[0] == int(reach.sum()) > S // 2     # BN: any hit in a reached leaf (bentnormals_step1.comp:170)
    th, th_cnt = out[orc.KIND_THICK]
    t = np.where(reach, 2.0 / (r[:, 2] / np.linalg.norm(r, axis=1)), maxd)
    assert th_cnt[0] == int(reach.sum())              # thickness: closest t, not clamped (thick_step1.comp:158)
    assert abs(th[0] - t.mean()) <= 1e-5 * t.mean() and th[0] > 1.5
    # (b) two triangles per leaf: roof and basement sit in leaves of their
    # own whose boxes the rays enter at t >= 2 > maxD, so the reference never
    # tests them: no occlusion, thickness maxD
    leaves, out, _ = beyond_maxd(2, S, maxd, library)
    assert all(len(z) == 1 for z in leaves) and len(leaves) == 3
    bn, bn_cnt = out[orc.KIND_BN]
    assert bn_cnt[0] == 0 and np.all(np.isfinite(bn[0]))
    th, th_cnt = out[orc.KIND_THICK]
    assert th_cnt[0] == 0 and th[0] == np.float32(maxd)


def test_bn_and_thickness_count_hits_beyond_maxd_only_in_reached_leaves():
    check_beyond_maxd()


@pytest.mark.parametrize("name", ["bn_maxd", "thick_clamp"])
def test_beyond_maxd_pins_reject_planted_mistakes(name):
    # the oracle built with a t < maxD test in bent normals, or with
    # thickness hits clamped to maxD, fails the pins above
    with pytest.raises(AssertionError):
        check_beyond_maxd(orc.mutant(name))


def test_missed_texels_get_defined_values():
    p, d = plane_texels(2)
    s, c, t = mapped(soup_mesh(quad(0.1, 4.0, flip=True)), p, d)
    assert np.all(t == MISS)
    ao, _, _ = hemi(orc.KIND_AO, s, c, t)
    bn, _, _ = hemi(orc.KIND_BN, s, c, t)
    th, _, _ = hemi(orc.KIND_THICK, s, c, t, maxd=3.0)
    assert np.all(ao == 1) and np.all(bn == 0) and np.all(th == 3.0)
    assert np.all(orc.height(c) == 0)


def test_hemisphere_results_do_not_depend_on_the_bvh():
    # With every hit closer than maxD, AO, bent normals and thickness are
    # functions of the geometry only; four different trees must agree.
    hi = merge(icosphere(1.0, 3), icosphere(1.4, 2, inward=True))
    rng = np.random.default_rng(4)
    dirs = rng.normal(size=(60, 3)).astype(np.float32)
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    p = (dirs * 0.95).astype(np.float32)
    res = []
    for mt in (1, 3, 8, 32):
        s, c, t = mapped(hi, p, dirs, max_tris=mt)
        out = [hemi(k, s, c, t, S=32, maxd=10.0)[:2] for k in (orc.KIND_AO, orc.KIND_BN, orc.KIND_THICK)]
        res.append((c, t, out))
    for c, t, out in res[1:]:
        assert np.array_equal(c, res[0][0])
        for (a, ca), (b, cb) in zip(out, res[0][2]):
            assert np.array_equal(ca, cb)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_ao_matches_brute_force_any_hit():
    # AO with minD < t < maxD decided by a scan over every triangle
    hi = merge(icosphere(1.0, 2), icosphere(1.3, 1, inward=True))
    s, c, t = mapped(hi, np.float32([[0.01, 0.02, 0.9]]), np.float32([[0, 0, 1]]))
    assert t[0] != MISS
    S, mind, maxd = 64, 1e-3, 0.35
    ao, cnt, table = hemi(orc.KIND_AO, s, c, t, S=S, mind=mind, maxd=maxd)
    o = (c[0, 1:, None] * s["spos"][t[0]:t[0] + 3, :3]).sum(0)
    nrm = c[0, 1:, None] * s["snrm"][t[0]:t[0] + 3, :3]
    n = nrm.sum(0) / np.linalg.norm(nrm.sum(0))
    ty = np.array([n[2], 0, -n[0]]) if abs(n[0]) > abs(n[1]) else np.array([0, n[2], -n[1]])
    ty /= np.linalg.norm(ty)
    tx = np.cross(n, ty)
    hits = 0
    for k in range(S):
        r = table[k]
        dr = tx * r[0] + ty * r[1] + n * r[2]
        dr = (dr / np.linalg.norm(dr)).astype(np.float32)
        tt = orc.hemi_bruteforce(o.astype(np.float32), dr, mind, s["spos"])
        hits += tt < maxd
    assert abs(int(cnt[0]) - hits) <= 1


# ------------------------------------------------------------ texel passes --
def test_tangentspace_frames():
    rng = np.random.default_rng(0)
    v = rng.normal(size=(20, 3)).astype(np.float32)
    n = np.tile(np.float32([0, 0, 1]), (20, 1))
    t = np.tile(np.float32([1, 0, 0]), (20, 1))
    b = np.tile(np.float32([0, 1, 0]), (20, 1))
    out = orc.tangentspace(n, t, b, v)
    assert np.allclose(out, v / np.linalg.norm(v, axis=1, keepdims=True), atol=1e-6)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] *= -1
    T, B, N = q[:, 0], q[:, 1], q[:, 2]
    out = orc.tangentspace(np.tile(N, (20, 1)), np.tile(T, (20, 1)), np.tile(B, (20, 1)), v)
    w = v @ q                                        # (v.T, v.B, v.N)
    assert np.allclose(out, w / np.linalg.norm(w, axis=1, keepdims=True), atol=1e-5)


def test_height_position_normal_closed_forms():
    p, d = plane_texels(5)
    s, c, t = mapped(soup_mesh(quad(0.25, 4.0)), p, d)
    assert np.all(orc.height(c) == np.float32(0.25))
    pos = orc.texel_attr(0, c, t, s["spos"])
    assert np.allclose(pos[:, :2], p[:, :2], atol=1e-6) and np.all(pos[:, 2] == np.float32(0.25))
    nrm = orc.texel_attr(1, c, t, s["snrm"])
    assert np.allclose(nrm, [0, 0, 1], atol=1e-7)
