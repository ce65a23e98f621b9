"""fornos_solver_run_all_noalloc, the entry point that allocates nothing:
through the C ABI, its results and counts equal fornos_solver_run_all's bit
for bit for AO, bent normals (tangent space) and thickness, on two calls in a
row on the same device, the second with a smaller texel set.  Thickness at
1024 samples needs several launches through the 64 MiB scratch array."""
import ctypes as C

import numpy as np
import pytest

import fornos_amd as fa
from fornos_amd._abi import check, lib
from bakes import Bake, same_bits

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return _hip


class DeviceBuffer:
    """Caller-owned device memory (the caller allocates; the entry point does not)."""

    def __init__(self, nbytes):
        self.nbytes = max(4, nbytes)
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), self.nbytes) == 0
        assert hip().hipMemset(self.p, 0xA5, self.nbytes) == 0  # garbage: the pass must overwrite it
        assert hip().hipDeviceSynchronize() == 0

    def read(self, dtype, count):
        out = np.empty(count, dtype)
        if count:
            assert hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, out.nbytes, 2) == 0  # device to host
        return out

    def __del__(self):
        hip().hipFree(self.p)


def run_noalloc(s):
    n = s.meshMapping.count
    out, cnt = DeviceBuffer(4 * n * s.comps), DeviceBuffer(4 * n)
    check(lib().fornos_solver_run_all_noalloc(s.handle, out.p, cnt.p))
    res = out.read(np.float32, n * s.comps)
    return (res if s.comps == 1 else res.reshape(n, s.comps)), cnt.read(np.uint32, n)


@pytest.fixture(scope="module")
def bakes(gpu):
    big = Bake(fa.configs.scaled("heightfield2048", size=256, high_quads=160))
    small = Bake(fa.configs.scaled("sphere512", size=128))
    assert small.mm.count < big.mm.count
    return big, small


KINDS = [("ao", {"sampleCount": 64, "minDistance": 0.01, "maxDistance": 0.2}),
         ("bentnormals", {"sampleCount": 64, "minDistance": 0.01, "maxDistance": 0.2, "tangentSpace": True}),
         ("thickness", {"sampleCount": 1024, "minDistance": 0.0, "maxDistance": 0.1})]


@pytest.mark.parametrize("kind,params", KINDS, ids=[k for k, _ in KINDS])
def test_noalloc_equals_run_all_twice(bakes, kind, params):
    for b in bakes:  # the second call has fewer texels
        p = dict(params)
        if b is bakes[1] and kind == "bentnormals":
            p["tangentSpace"] = False  # the sphere's texel set has no tangent frame
        s = b.solver(kind, **p)
        s.run_all()
        ref, ref_cnt = s.getResults().copy(), s.counts().copy()
        res, cnt = run_noalloc(s)
        assert same_bits(res, ref), kind
        assert np.array_equal(cnt, ref_cnt), kind
        # the solver's own results are left as they were
        assert same_bits(s.getResults(), ref)


def test_noalloc_rejects_bad_arguments(bakes):
    big, _ = bakes
    h = fa.HeightSolver(True, 0.0).init(big.prep["texels"], big.mm)
    buf = DeviceBuffer(4 * big.mm.count)
    with pytest.raises(fa.FornosError):
        check(lib().fornos_solver_run_all_noalloc(h.handle, buf.p, buf.p))
    ao = big.solver("ao", sampleCount=8, minDistance=0.01, maxDistance=0.2)
    with pytest.raises(fa.FornosError):
        check(lib().fornos_solver_run_all_noalloc(ao.handle, None, buf.p))
