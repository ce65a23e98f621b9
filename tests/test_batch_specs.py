"""Batched-job workload: the seeded job table and the batch entry points'
argument checks (no GPU needed)."""
import ctypes as C

from fornos_amd import _abi, batch
from fornos_amd._abi import lib


def test_job_table_is_seeded_and_in_range():
    a = batch.job_specs(64)
    b = batch.job_specs(64)
    assert a == b
    fams = {j["family"] for j in a}
    assert fams == set(batch.FAMILIES)
    assert {j["ao"]["sampleCount"] for j in a} == set(batch.SAMPLES)
    assert {j["size"] for j in a} == set(batch.SIZES)
    for j in a:
        gen, _, _, p, q = j["high"]
        tris = {_abi.GEN_ICOSPHERE_WAVES: 20 * 4 ** p, _abi.GEN_HEIGHTFIELD: 2 * p * p,
                _abi.GEN_TORUS_KNOT: 2 * p * q}[gen]
        assert 20_000 <= tris <= 5_300_000, j
        gen, _, _, p, q = j["low"]
        tris = {_abi.GEN_UV_SPHERE: 2 * p * (q - 1), _abi.GEN_HEIGHTFIELD_LOW: 2 * p * p,
                _abi.GEN_TORUS_KNOT: 2 * p * q}[gen]
        assert 100 <= tris <= 25_000, j


def test_scaled_jobs_shrink():
    full, small = batch.job_specs(8), batch.job_specs(8, scale=0.05)
    for f, s in zip(full, small):
        assert s["size"] <= f["size"] and s["family"] == f["family"]
        assert s["ao"] == f["ao"]


def test_batch_arguments_are_checked():
    L = lib()
    out = C.c_void_p()
    assert L.fornos_batch_create(None, 0, C.byref(out)) == _abi.INVALID_ARG
    arr = (C.c_void_p * 1)(None)
    assert L.fornos_batch_create(arr, 1, C.byref(out)) == _abi.UNSUPPORTED
    assert b"AO solver" in L.fornos_last_error()
    assert L.fornos_batch_set_launch(None, 4, 16, 24) == _abi.INVALID_ARG
    assert L.fornos_batch_run(None, None) == _abi.INVALID_ARG
    assert L.fornos_batch_autotune(None, C.c_float(0.1), None, None, None, None) == _abi.INVALID_ARG
    assert L.fornos_solver_set_launch(None, 8, 16, 24) == _abi.INVALID_ARG
