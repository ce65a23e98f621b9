"""The procedural meshes face outward: vertex normals of the closed meshes
(UV sphere, displaced icosphere, torus-knot tube) point away from the centre
(the tube's centre line for the knot), heightfield normals point up.  An
inward-facing closed mesh occludes every hemisphere ray (AO = 0)."""
import numpy as np
import pytest

import fornos_amd as fa
from fornos_amd import _abi


def vertex_normals(m):
    a = m.arrays()
    n = np.zeros_like(a["positions"])
    n[a["vertices"][:, 0]] = a["normals"][a["vertices"][:, 2]]
    return a["positions"], n


@pytest.mark.parametrize("spec", [(_abi.GEN_UV_SPHERE, 0, 1.0, 32, 16), (_abi.GEN_ICOSPHERE_WAVES, 7, 0.02, 4, 0)])
def test_spheres_face_outward(spec):
    p, n = vertex_normals(fa.Mesh.generate(*spec))
    assert np.all(np.sum(p * n, axis=1) > 0)


@pytest.mark.parametrize("spec", [(_abi.GEN_TORUS_KNOT, 3, 0.03, 256, 32), (_abi.GEN_TORUS_KNOT, 0, 0.0, 128, 16)])
def test_knot_faces_outward(spec):
    p, n = vertex_normals(fa.Mesh.generate(*spec))
    ring = spec[4]
    p, n = p.reshape(-1, ring, 3), n.reshape(-1, ring, 3)
    assert np.all(np.sum((p - p.mean(axis=1, keepdims=True)) * n, axis=2) > 0)


@pytest.mark.parametrize("spec", [(_abi.GEN_HEIGHTFIELD, 1, 0.05, 64, 20), (_abi.GEN_HEIGHTFIELD_LOW, 1, 0.05, 16, 0)])
def test_heightfields_face_up(spec):
    _, n = vertex_normals(fa.Mesh.generate(*spec))
    assert np.all(n[:, 2] > 0)
