"""The reference-signature C++ classes (fornos_amd/csrc/fornos.hpp), driven by
a C++ host program built against the public headers (tests/cxx/shim_bake.cpp,
compiled by __graft_entry__.build()), give the C-ABI path's results bit for
bit on the sphere workload: MeshMapping::init / runStep / progress
(Src/meshmapping.h:42-57) and every solver's init / runStep / getResults
(Src/solver_ao.h:36-96 and analogues), with getResults() returning the
reference's new[] array of workCount (texel count rounded up to 64) entries
(Src/solver_ao.cpp:133-138, 157-163) and one chunk of at most 131072 work items
per runStep (Src/meshmapping.cpp:186-218, Src/solver_ao.cpp:83-87)."""
import os
import subprocess

import numpy as np
import pytest

import fornos_amd as fa
from bakes import same_bits

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cxx", "shim_bake")
STEP = 131072


@pytest.fixture(scope="module")
def shim(gpu, tmp_path_factory):
    if not os.path.exists(BIN):
        pytest.fail("tests/cxx/shim_bake is missing: __graft_entry__.build() compiles it")
    d = tmp_path_factory.mktemp("shim")
    p = subprocess.run([BIN, str(d)], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    meta = dict(line.split() for line in open(d / "meta.txt"))
    return d, {k: int(v) for k, v in meta.items()}


@pytest.fixture(scope="module")
def abi(gpu):
    prep = fa.prepare("sphere512")
    mm = fa.MeshMapping().init(prep["texels"], prep["scene"], True)
    mm.run_all()
    return prep, mm


def test_meshmapping_class_equals_c_abi(shim, abi):
    d, meta = shim
    prep, mm = abi
    coords, tidx = mm.read()
    n = len(tidx)
    assert meta["texels"] == n and meta["work"] == (n + 63) // 64 * 64
    assert meta["mapping_steps"] == -(-meta["work"] // STEP)
    assert np.array_equal(np.fromfile(d / "tidx.bin", np.uint32), tidx)
    assert same_bits(np.fromfile(d / "coords.bin", np.float32).reshape(-1, 4), coords)


SOLVERS = [("ao", lambda: fa.AmbientOcclusionSolver(32, 0.0, 0.5), 1, 32),
           ("bentnormals", lambda: fa.BentNormalsSolver(32, 0.0, 0.5, False), 3, 32),
           ("thickness", lambda: fa.ThicknessSolver(16, 0.0, 0.5), 1, 16),
           ("height", lambda: fa.HeightSolver(True, 0.0), 1, 1),
           ("normals", lambda: fa.NormalsSolver(False), 3, 1),
           ("position", lambda: fa.PositionSolver(), 3, 1)]


@pytest.mark.parametrize("name,make,comps,S", SOLVERS, ids=[s[0] for s in SOLVERS])
def test_solver_class_equals_c_abi(shim, abi, name, make, comps, S):
    d, meta = shim
    prep, mm = abi
    n, work = meta["texels"], meta["work"]
    s = make().init(prep["texels"], mm)
    s.run_all()
    r = np.fromfile(d / (name + ".bin"), np.float32)
    assert r.size == work * comps  # getResults(): workCount entries
    r = r.reshape(work, comps) if comps > 1 else r
    assert same_bits(r[:n], s.getResults())
    assert not r[n:].any()  # the padding past the texel count is zero
    per = max(1, STEP // S) * S  # whole texels per step, at most 131072 work items
    assert meta[name + "_steps"] == -(-(work * S) // per)
