"""The reachability links of the reference tree (reference_chain_links in
fornos_amd/csrc/bvh.cpp) skip every ancestor whose box contains its child's:
the reference's slab test (RayAABB, Shaders/ao_step1.comp:43-52) then passes
the parent whenever it passes the child.  Checked here with float32
arithmetic, one IEEE rounding per operation as in the shader, on boxes nested
to the last unit in the last place and rays that graze them."""
import numpy as np

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
