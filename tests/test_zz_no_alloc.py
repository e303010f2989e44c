"""cd_nds_noalloc: the non-dominated sort over device buffers with its scratch
in a fixed device array, called through the C ABI, gives exactly what cd_nds
gives -- on two calls in a row in one stream, the second on a smaller pool
(so the second must clear what the first left in the shared scratch)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip(cdp):
    cdp.lib()  # the engine loads libamdhip64.so.7; this is the same runtime
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


def dev_alloc(h, nbytes):
    p = C.c_void_p()
    assert h.hipMalloc(C.byref(p), nbytes) == 0
    return p


def quantised_pool(rng, P):
    obj = rng.normal(size=(P, 6)).astype(np.float32)
    obj[:, 2] = np.round(obj[:, 2] * 3) / 3
    obj[:, 5] = np.round(obj[:, 5] * 2) / 2
    dup = rng.integers(0, P, size=P // 8)
    obj[dup] = obj[rng.integers(0, P, size=len(dup))]
    obj[rng.integers(0, P, size=P // 8), 0] += np.float32(5e-7)
    return obj


def test_noalloc_equals_nds_twice(cdp):
    print("no-alloc nds: 8192 then 200 designs", flush=True)
    h = hip(cdp)
    lib = cdp.lib()
    rng = np.random.default_rng(77)
    cases = [(8192, 4096), (200, 100)]
    pools = [quantised_pool(rng, P) for P, _ in cases]
    bufs = []
    for (P, N), obj in zip(cases, pools):  # both calls enqueued on the null stream back to back
        d_obj, d_fo, d_nf = dev_alloc(h, P * 24), dev_alloc(h, P * 4), dev_alloc(h, 4)
        assert h.hipMemcpy(d_obj, obj.ctypes.data, P * 24, H2D) == 0
        assert lib.cd_nds_noalloc(d_obj, P, N, d_fo, d_nf, None) == 0, lib.cd_last_error()
        bufs.append((d_obj, d_fo, d_nf))
    assert h.hipDeviceSynchronize() == 0
    for (P, N), obj, (d_obj, d_fo, d_nf) in zip(cases, pools, bufs):
        fo = np.zeros(P, np.int32)
        nf = np.zeros(1, np.int32)
        assert h.hipMemcpy(fo.ctypes.data, d_fo, P * 4, D2H) == 0
        assert h.hipMemcpy(nf.ctypes.data, d_nf, 4, D2H) == 0
        want_fo, want_nf = cdp.nds(obj, N)
        assert np.array_equal(fo, want_fo) and int(nf[0]) == want_nf, (P, N)
        for p in (d_obj, d_fo, d_nf):
            h.hipFree(p)


def test_noalloc_rejects_a_pool_over_its_maximum(cdp):
    lib = cdp.lib()
    p = C.c_void_p(16)  # never dereferenced: the size check comes first
    assert lib.cd_nds_noalloc(p, 8194, 4097, p, p, None) == -1
    assert b"maximum" in lib.cd_last_error()
    assert lib.cd_nds_noalloc(p, 100, 101, p, p, None) == -1
