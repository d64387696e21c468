"""vl_event_sort_noalloc: the event sort for callers that cannot allocate
device memory, whose scratch is a static device array.  Called through the
C ABI on device buffers the test owns, it must give exactly what
vl_event_sort gives, on two calls in a row (the second smaller, so stale
contents of the shared scratch would show), and reject sizes beyond its
bound and host memory."""
import ctypes as C

import numpy as np
import pytest

import vlsi_layout_amd as V

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    from conftest import hip_runtime
    h = hip_runtime()
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMalloc.restype = C.c_int
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemcpy.restype = C.c_int
    h.hipFree.argtypes = [C.c_void_p]
    h.hipFree.restype = C.c_int
    return h


def trackers(rng, n, span):
    t = np.zeros(n, V.TRACKER_DTYPE)
    t["X"] = rng.integers(-span, span, n)
    t["Y"] = rng.integers(-span, span, n)
    t["polygonNumber"] = (np.arange(n) % 251) - 125   # payload: equal keys must keep their order
    return t


def to_device(h, t):
    p = C.c_void_p()
    assert h.hipMalloc(C.byref(p), max(t.nbytes, 1)) == 0
    assert h.hipMemcpy(p, t.ctypes.data_as(C.c_void_p), t.nbytes, H2D) == 0
    return p


def from_device(h, p, n):
    out = np.zeros(n, V.TRACKER_DTYPE)
    assert h.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, D2H) == 0
    return out


def test_noalloc_event_sort_matches_event_sort(gpu_engine):
    h = hip()
    eng = V.Engine(0)
    L = V.lib()
    rng = np.random.default_rng(2024)
    # the bound first, then a smaller input over the same scratch; small
    # coordinate ranges give many equal keys (stability), the wide one signs
    for n, span in ((V.VL_NOALLOC_MAX_TRACKERS, 1 << 31), (123_457, 300)):
        t = trackers(rng, n, span)
        ref = V.device_event_sort(t.copy(), eng)               # the existing entry point
        d1, d2 = to_device(h, t), to_device(h, t)
        try:
            assert L.vl_event_sort_noalloc(eng.handle, d1, n) == V.VL_OK, L.vl_last_error()
            assert L.vl_event_sort(eng.handle, d2, n) == V.VL_OK, L.vl_last_error()
            got, dev_ref = from_device(h, d1, n), from_device(h, d2, n)
        finally:
            h.hipFree(d1)
            h.hipFree(d2)
        # numpy's copy of this padded record type leaves the padding bytes
        # undefined, so the host-memory result is compared field by field;
        # vl_event_sort on the same device bytes must match byte for byte
        for f in ("X", "Y", "polygonNumber"):
            assert np.array_equal(got[f], ref[f]), (n, f)
        assert got.tobytes() == dev_ref.tobytes(), n
    eng.close()


def test_noalloc_event_sort_rejects_what_it_cannot_do(gpu_engine):
    h = hip()
    eng = V.Engine(0)
    L = V.lib()
    t = trackers(np.random.default_rng(1), 1000, 100)
    d = to_device(h, t)
    try:
        assert L.vl_event_sort_noalloc(eng.handle, d, V.VL_NOALLOC_MAX_TRACKERS + 1) == V.VL_EINVAL
        assert b"at most" in L.vl_last_error()
        # host memory is refused (the entry point would need a device copy)
        host = t.copy()
        assert L.vl_event_sort_noalloc(eng.handle, host.ctypes.data_as(C.c_void_p), len(host)) == V.VL_EINVAL
        assert np.array_equal(host, t)
        # n <= 1 is a no-op
        assert L.vl_event_sort_noalloc(eng.handle, d, 1) == V.VL_OK
        assert from_device(h, d, 1000).tobytes() == t.tobytes()
    finally:
        h.hipFree(d)
    eng.close()
