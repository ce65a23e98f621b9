"""The reachability links of the reference tree (reference_chain_links in
fornos_amd/csrc/bvh.cpp) skip every ancestor whose box contains its child's:
the reference's slab test (RayAABB, Shaders/ao_step1.comp:43-52) then passes
the parent whenever it passes the child.  Checked here with float32
arithmetic, one IEEE rounding per operation as in the shader, on boxes nested
to the last unit in the last place and rays that graze them."""
import numpy as np
import pytest

FLT_MAX = np.float32(3.4028234663852886e38)


def ray_aabb_pass(o, d, lo, hi, T):
    """RayAABB(o, d, lo, hi) < T, elementwise over rows (float32)."""
    t1 = (lo - o) / d
    t2 = (hi - o) / d
    a = np.minimum(t1, t2).max(axis=1)
    b = np.maximum(t1, t2).min(axis=1)
    da = np.where((b >= 0) & (a <= b), a, FLT_MAX)
    return da < T


def _grow(x, ulps, toward):
    for _ in range(ulps):
        x = np.nextafter(x, toward).astype(np.float32)
    return x


def _cases(rng, n):
    lo = rng.uniform(-4, 4, (n, 3)).astype(np.float32)
    hi = (lo + rng.uniform(0, 2, (n, 3)).astype(np.float32) * rng.choice([0, 1e-6, 1], (n, 3))).astype(np.float32)
    # rays from outside towards a point on the child's surface, so that the
    # slab planes are met at or near their ends
    u = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    face = rng.integers(0, 3, n)
    side = rng.integers(0, 2, n)
    target = (lo + u * (hi - lo)).astype(np.float32)
    target[np.arange(n), face] = np.where(side, hi[np.arange(n), face], lo[np.arange(n), face])
    d = rng.normal(size=(n, 3)).astype(np.float32)
    tiny = rng.random((n, 3)) < 0.05
    d[tiny] = (np.sign(d[tiny]) * rng.choice([2.0 ** -40, 2.0 ** -30, 1e-7], tiny.sum())).astype(np.float32)
    d[d == 0] = np.float32(2.0 ** -40)
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    d[np.abs(d) < 2.0 ** -40] = np.float32(2.0 ** -40)
    dist = rng.uniform(0, 3, (n, 1)).astype(np.float32)
    o = (target - d * dist).astype(np.float32)
    T = (dist[:, 0] * rng.uniform(0.98, 1.02, n)).astype(np.float32)
    return o, d, lo, hi, T


def test_nested_parent_passes_with_child():
    rng = np.random.default_rng(7)
    for ulps in (0, 1, 2, 50):
        o, d, lo, hi, T = _cases(rng, 400000)
        plo = lo.copy()
        phi = hi.copy()
        grow = rng.random(lo.shape) < 0.5
        plo[grow] = _grow(lo[grow], ulps, -np.inf)
        phi[grow] = _grow(hi[grow], ulps, np.inf)
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            child = ray_aabb_pass(o, d, lo, hi, T)
            parent = ray_aabb_pass(o, d, plo, phi, T)
        assert child.sum() > 10000
        assert not np.any(child & ~parent), "a nested parent failed where its child passed"


def test_parent_one_ulp_inside_can_fail():
    """The containment condition is needed: a parent plane one unit inside
    the child's loses rays the child keeps."""
    rng = np.random.default_rng(11)
    o, d, lo, hi, T = _cases(rng, 400000)
    plo = lo.copy()
    phi = hi.copy()
    plo[:, 0] = _grow(lo[:, 0], 1, np.inf)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        child = ray_aabb_pass(o, d, lo, hi, T)
        parent = ray_aabb_pass(o, d, plo, phi, T)
    assert np.any(child & ~parent)


def _tree(name):
    import fornos_amd as fa
    if name == "sphere512":
        cfg = fa.CONFIGS["sphere512"]
    elif name == "heightfield_small":
        cfg = fa.configs.scaled("heightfield2048", size=384, high_quads=96)
    else:
        cfg = fa.configs.scaled("knot4096", size=256, high_quads=64)
    return fa.configs.make_mesh(cfg["high"])


@pytest.mark.parametrize("name", ["sphere512", "heightfield_small", "knot_small"])
def test_links_decide_every_leaf_as_its_whole_chain(name):
    """On real reference trees (fornos_bvh_reference_links), for rays leaving
    the surface: a leaf's link chain passes the reference test exactly when
    every node from the root to it does, at several bounds T."""
    import ctypes as C
    from fornos_amd._abi import check, lib, ptr
    mesh = _tree(name)
    n = C.c_uint32()
    check(lib().fornos_bvh_reference_export(mesh.handle, 8, 8192, C.byref(n), None, None))
    N = n.value
    nodes = np.empty((N, 9), np.uint32)
    ids = np.empty(mesh.n_triangles, np.uint32)
    check(lib().fornos_bvh_reference_export(mesh.handle, 8, 8192, C.byref(n), ptr(nodes), ptr(ids)))
    parent = np.empty(N, np.uint32)
    links = np.empty(N, np.uint32)
    check(lib().fornos_bvh_reference_links(mesh.handle, 8, 8192, N, ptr(parent), ptr(links)))
    lo = np.ascontiguousarray(nodes[:, 0:3]).view(np.float32)
    hi = np.ascontiguousarray(nodes[:, 3:6]).view(np.float32)
    leaf = nodes[:, 6] != nodes[:, 7]
    miss = np.uint32(0xFFFFFFFF)
    # parent links agree with the skip links; every link is an ancestor
    skip = nodes[:, 8]
    for i in np.nonzero(~leaf)[0]:
        c = i + 1
        while c < skip[i]:
            assert parent[c] == i
            c = c + 1 if leaf[c] else skip[c]
    # rays from points on random triangles, in random directions
    rng = np.random.default_rng(3)
    a = mesh.arrays()
    pos = np.asarray(a["positions"], np.float32).reshape(-1, 3)
    tri = np.asarray(a["triangles"], np.int64).reshape(-1, 3)
    verts = np.asarray(a["vertices"], np.int64).reshape(-1, 3)
    R = 256
    t = rng.integers(0, len(tri), R)
    w = rng.dirichlet([1, 1, 1], R).astype(np.float32)
    corners = pos[verts[tri[t], 0]]
    o = (corners * w[:, :, None]).sum(1).astype(np.float32)
    d = rng.normal(size=(R, 3)).astype(np.float32)
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    d[np.abs(d) < 2.0 ** -40] = np.float32(2.0 ** -40)
    diag = float(np.linalg.norm(pos.max(0) - pos.min(0)))
    for T in (np.float32(0.02 * diag), np.float32(0.3 * diag), np.float32(3e38)):
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            p = np.stack([ray_aabb_pass(o, d, np.broadcast_to(lo[x], o.shape), np.broadcast_to(hi[x], o.shape), T)
                          for x in range(N)], axis=1)
        full = np.zeros_like(p)
        linked = np.zeros_like(p)
        for x in range(N):  # DFS order: ancestors come first
            full[:, x] = p[:, x] & (full[:, parent[x]] if parent[x] != miss else True)
            linked[:, x] = p[:, x] & (linked[:, links[x]] if links[x] != miss else True)
        assert np.array_equal(full[:, leaf], linked[:, leaf])
        assert full[:, leaf].any()
    depth = np.zeros(N, np.int64)
    hops = np.zeros(N, np.int64)
    for x in range(N):
        depth[x] = 1 + (depth[parent[x]] if parent[x] != miss else 0)
        hops[x] = 1 + (hops[links[x]] if links[x] != miss else 0)
    assert hops[leaf].mean() < 0.5 * depth[leaf].mean()
