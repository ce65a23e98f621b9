"""Handles may be destroyed in any order: a mapping holds a reference to its
scene, a solver to its mapping and a batch to its solvers, as the reference's
shared_ptr members do (Src/solver_ao.h:93).  A garbage collector finalizing
objects of a reference cycle calls their finalizers in no fixed order, so the
scene's handle may go before the mapping's; the objects it still serves must
keep working and be freed last."""
import numpy as np
import pytest

import fornos_amd as fa
from bakes import same_bits
from meshes import icosphere

pytestmark = pytest.mark.gpu


def _rig(n=4096, seed=5):
    m = icosphere(1.0, 2)
    hi = fa.Mesh.from_arrays(m["positions"], m["vertices"], m["triangles"], m["normals"], m["texcoords"])
    scene = fa.Scene(hi)
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    p = (0.5 * d).astype(np.float32)  # inside the sphere, looking out
    up = np.where(np.abs(d[:, :1]) < 0.9, [[1, 0, 0]], [[0, 1, 0]])
    t = np.cross(d, up)
    t = (t / np.linalg.norm(t, axis=1, keepdims=True)).astype(np.float32)
    b = np.cross(d, t).astype(np.float32)
    tex = fa.TexelArrays({"positions": p, "directions": d, "normals": d, "tangents": t, "bitangents": b,
                          "indices": np.arange(n, dtype=np.uint32)}, n, 1)
    mm = fa.MeshMapping().init(tex, scene, False)
    ao = fa.AmbientOcclusionSolver(32, 0.0, 0.5).init(tex, mm)
    return hi, scene, tex, mm, ao


def _bake(mm, ao):
    mm.reset()
    mm.run_all()
    ao.reset()
    ao.run_all()
    return mm.read(), ao.getResults()


def test_scene_and_mesh_closed_first(gpu):
    hi, scene, tex, mm, ao = _rig()
    (c0, t0), r0 = _bake(mm, ao)
    assert np.all(t0 != 0xFFFFFFFF)
    hi.close()
    scene.close()  # the mapping keeps the scene alive
    (c1, t1), r1 = _bake(mm, ao)
    assert np.array_equal(t0, t1) and same_bits(c0, c1) and same_bits(r0, r1)
    mm.close()  # the solver keeps the mapping alive
    ao.reset()
    ao.run_all()
    ao.close()  # frees solver, then mapping, then scene


def test_batch_keeps_its_solvers(gpu):
    hi, scene, tex, mm, ao = _rig(seed=6)
    (_, _), r0 = _bake(mm, ao)
    ao2 = fa.AmbientOcclusionSolver(32, 0.0, 0.5).init(tex, mm)
    b = fa.Batch([ao, ao2])
    ao2.close()  # the batch still holds it
    scene.close()
    ao.reset()
    b.run()
    assert same_bits(ao.getResults(), r0)
    b.close()
    ao.close()
    mm.close()
