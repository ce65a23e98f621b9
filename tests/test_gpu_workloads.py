"""The batched-jobs and denoiser workloads at full size, on the GPU.

- 64 seeded AO jobs at scale 1 (sphere / heightfield / knot, 256^2-2048^2,
  32-256 samples): the one-launch batch equals 64 sequential run_all calls
  on every job, and three jobs' mappings and AO equal the oracle on windows.
- the denoiser on held-out seeds at 2048^2 with the committed weights: AO
  mean error <= 0.01 and bent-normal angle <= 2 degrees against the plain
  1024-sample bake, and both several times below the plain 16-sample bake's
  (what the network gains); with the network damaged, every block that fails
  the check equals the plain 1024-sample bake bit for bit.
- fornos_solver_run_texels, which that fallback uses: listed texels equal
  run_all's results.
- the network input computed from the mapping inside the gather equals the
  one gathered from the position, normals and height passes, bit for bit.
"""
import numpy as np
import pytest

import fornos_amd as fa
import oracle as orc
from bakes import MISS, Bake, same_bits, windows
from fornos_amd import batch

pytestmark = pytest.mark.gpu


# ---- fornos_solver_run_texels ------------------------------------------------

@pytest.fixture(scope="module")
def field(gpu):
    return Bake(fa.configs.scaled("heightfield2048", size=256, high_quads=160))


@pytest.mark.parametrize("kind", ["ao", "bentnormals", "thickness", "paired"])
def test_run_texels_equals_run_all(field, kind):
    n = field.mm.count
    rng = np.random.default_rng(5)
    texels = np.unique(np.concatenate([rng.choice(n, 3000, replace=False), np.arange(1000, 1300)]))
    k = "ao" if kind == "paired" else kind
    full = field.solver(k)
    part = field.solver(k)
    if kind == "paired":
        bn_full, bn_part = field.solver("bentnormals"), field.solver("bentnormals")
        full.pair(bn_full)
        part.pair(bn_part)
    full.run_all()
    part.run_texels(texels)
    assert same_bits(part.getResults()[texels], full.getResults()[texels])
    assert np.array_equal(part.counts()[texels], full.counts()[texels])
    if kind == "paired":
        assert same_bits(bn_part.getResults()[texels], bn_full.getResults()[texels])
        assert np.array_equal(bn_part.counts()[texels], bn_full.counts()[texels])
    # any order: the same results
    again = field.solver(k)
    again.run_texels(np.random.default_rng(6).permutation(texels))
    assert same_bits(again.getResults()[texels], full.getResults()[texels])
    with pytest.raises(fa.FornosError):
        part.run_texels(np.concatenate([texels, texels[:1]]))  # listed twice
    with pytest.raises(fa.FornosError):
        part.run_texels([n])  # beyond the texel set


# ---- batched jobs at scale 1 -------------------------------------------------

@pytest.fixture(scope="module")
def jobs64(gpu):
    return [batch.Job(s) for s in batch.job_specs(64, scale=1.0)]


@pytest.mark.timeout(600)
def test_batch_64_jobs_equal_sequential(jobs64):
    assert len(jobs64) == 64
    assert sum(j.high_tris for j in jobs64) > 40_000_000
    assert {j.spec["family"] for j in jobs64} == set(batch.FAMILIES)
    seq = []
    for j in jobs64:
        j.solver.reset()
        j.solver.run_all()
        seq.append((j.solver.getResults().view(np.uint32).copy(), j.solver.counts().copy()))
    b = batch.Batch([j.solver for j in jobs64])
    for rep in range(2):
        ms = b.run()
        for i, (j, (r, c)) in enumerate(zip(jobs64, seq)):
            assert np.array_equal(j.solver.counts(), c), (rep, i, j.spec)
            assert np.array_equal(j.solver.getResults().view(np.uint32), r), (rep, i, j.spec)
    print("64 jobs, %d rays: batch %.0f ms" % (sum(j.rays for j in jobs64), ms))
    b.close()


def test_batch_jobs_match_oracle(jobs64):
    # the three jobs with the fewest hi-poly triangles, one per family where possible
    picked = {}
    for j in sorted(jobs64, key=lambda j: j.high_tris):
        picked.setdefault(j.spec["family"], j)
    assert len(picked) == 3
    for fam, j in picked.items():
        hi = j.scene.mesh.arrays()
        nodes, spos, snrm, _ = orc.bvh_build(hi)
        tex = j.mapping.read_texels()
        coords, tidx = j.mapping.read()
        n = len(tidx)
        sel = windows(n, 8, 128)
        for a, b in zip(sel[::128], sel[127::128]):
            oc, ot = orc.meshmapping(tex["positions"], tex["directions"], spos, nodes, j.spec["cull"],
                                     begin=int(a), end=int(b) + 1)
            assert np.array_equal(ot[a:b + 1], tidx[a:b + 1]), (fam, a)
            assert same_bits(oc[a:b + 1], coords[a:b + 1]), (fam, a)
        p = j.solver.params
        j.solver.reset()
        j.solver.run_all()
        res, cnt = j.solver.getResults(), j.solver.counts()
        for a, b in zip(sel[::128], sel[127::128]):
            out, ocnt = orc.hemisphere(orc.KIND_AO, p.sample_count, p.min_distance, p.max_distance, j.solver.table,
                                       coords, tidx, spos, snrm, nodes, int(a), int(b) + 1)
            assert np.array_equal(cnt[a:b + 1], ocnt), (fam, a)
            assert same_bits(res[a:b + 1], out), (fam, a)
        print("%s job: %d hi tris, %d^2, S=%d: %d texels checked" % (fam, j.high_tris, j.spec["size"],
                                                                      p.sample_count, len(sel)))


# ---- the denoiser at 2048^2 ----------------------------------------------------

@pytest.fixture(scope="module")
def net(gpu):
    from fornos_amd import denoise as dn
    import torch
    d = dn.Denoiser.load(dn.WEIGHTS)
    d.forward(torch.zeros((1, 2048, 2048, dn.IN), dtype=torch.float32, device="cuda"))
    return d


@pytest.mark.parametrize("seed", [2007, 2010])
def test_denoiser_gain_on_held_out_seed(net, seed):
    from fornos_amd import denoise as dn
    job = dn.MeshJob(batch.job_spec(seed), 2048)
    ra, rb, t_plain, mm = dn.plain_bake(job)
    _, tidx = mm.read()
    hit = tidx != MISS
    a16, b16, _, _ = dn.plain_bake(job, dn.LOW_S, seed=11)
    da, db, t_dn, info = dn.denoised_bake(job, net)
    ae16, be16 = dn.errors(a16, b16, ra, rb, hit)
    ae, be = dn.errors(da, db, ra, rb, hit)
    print("seed %d %s: 16 samples AO MAE %.4f BN %.2f deg; denoised AO MAE %.4f BN %.2f deg; %.2f s vs %.3f s "
          "(%.1fx), %d fallback blocks" % (seed, job.spec["family"], ae16, be16, ae, be, t_plain, t_dn,
                                           t_plain / t_dn, info["fallback_tiles"]))
    assert ae <= 0.01 and be <= 2.0
    assert ae * 3 < ae16 and be * 3 < be16  # what the network gains over its input
    assert t_dn * 5 < t_plain


def test_gather_from_mapping_equals_gather_from_passes(gpu):
    # the denoised bake gathers position, normal and distance from the
    # mapping itself; the three passes give the same network input
    import torch
    from fornos_amd import denoise as dn
    job = dn.MeshJob(batch.job_spec(2007), 512)
    mm = job.mapping()
    ao, bn = job.solvers(mm, dn.LOW_S, seed=11)
    ao.run_all()
    bn.run_all()
    W = job.size
    a = torch.full((1, W, W, dn.IN), 7.0, dtype=torch.float32, device="cuda")
    b = torch.full((1, W, W, dn.IN), -7.0, dtype=torch.float32, device="cuda")
    job.gather(mm, ao, bn, job.gbuffers(mm), a)
    job.gather(mm, ao, bn, None, b)
    ai, bi = a.cpu().numpy().view(np.uint32), b.cpu().numpy().view(np.uint32)
    assert np.array_equal(ai, bi)
    _, tidx = mm.read()
    assert 0 < int(np.sum(tidx != MISS)) < W * W  # hits, misses and pixels outside the texel set
    # all three passes or none; and the mapping must have been run
    pos = fa.PositionSolver().init(None, mm)
    c = np.zeros(3, np.float32)
    with pytest.raises(fa.FornosError, match="normals must be a solver"):
        fa._abi.check(fa.lib().fornos_denoise_gather(mm.handle, ao.handle, bn.handle, pos.handle, None, None,
                                                     c.ctypes.data, 1.0, 1.0, dn._p(a)))
    import ctypes as C
    with pytest.raises(fa.FornosError, match="16-byte aligned"):  # checked before any launch
        fa._abi.check(fa.lib().fornos_denoise_gather(mm.handle, ao.handle, bn.handle, None, None, None,
                                                     c.ctypes.data, 1.0, 1.0, C.c_void_p(b.data_ptr() + 4)))
    fresh = fa.MeshMapping().init_from_mesh(job.lo, job.dirs, W, W, job.scene, job.cull)
    ao2, bn2 = job.solvers(fresh, dn.LOW_S, seed=11)
    with pytest.raises(fa.FornosError, match="has not been run"):
        job.gather(fresh, ao2, bn2, None, b)


def test_denoiser_fallback_blocks_equal_plain_bake(net):
    from fornos_amd import denoise as dn
    job = dn.MeshJob(batch.job_spec(2010), 2048)
    ra, rb, _, _ = dn.plain_bake(job)
    good = net.get_params()
    damaged = good.copy()
    damaged[-4] += 0.03  # the head's AO bias: every denoised AO 0.03 too high
    try:
        net.set_params(damaged)
        da, db, _, info = dn.denoised_bake(job, net, masks=True)
        na, nb, _, _ = dn.denoised_bake(job, net, fallback=False)
    finally:
        net.set_params(good)
    sel, bad = info["fallback_mask"], info["fail_blocks"]
    print("damaged network: %d of %d blocks fell back, %d of %d texels"
          % (info["fallback_tiles"], info["tiles"], int(sel.sum()), len(sel)))
    assert info["fallback_tiles"] == int(bad.sum()) > 0
    assert int(sel.sum()) == info["fallback_texels"] > 0
    # the texels of failing blocks carry the plain 1024-sample bake's results
    assert same_bits(da[sel], ra[sel]) and same_bits(db[sel], rb[sel])
    # the rest carry the network's
    assert same_bits(da[~sel], na[~sel]) and same_bits(db[~sel], nb[~sel])
