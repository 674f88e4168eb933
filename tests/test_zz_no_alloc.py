"""kpz_roughness_noalloc (fixed per-device workspace, no allocation) against
kpz_roughness on the same lattices, through the C ABI: two calls in a row,
the second on a smaller lattice, bit-exact int64/int128 sums; plus the size
limit and argument errors."""
import ctypes as C

import numpy as np
import pytest

import cudakpz_amd as K

pytestmark = pytest.mark.gpu


class Hip:
    """Device buffers for the test's output words (the test allocates; the
    entry point under test does not)."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        self.h.hipStreamSynchronize.argtypes = [C.c_void_p]

    def alloc(self, n):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), n) == 0
        return p

    def read_i64(self, p, n):
        out = np.zeros(n, dtype=np.int64)
        assert self.h.hipMemcpy(out.ctypes.data, p, 8 * n, 2) == 0  # device to host
        return out


def noalloc_fn():
    L = K.lib()
    f = L.kpz_roughness_noalloc
    f.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    f.restype = C.c_int
    return f


def expected(s):
    out = (C.c_int64 * 3)()
    assert K.lib().kpz_roughness(s._h, out) == 0
    return np.array(list(out), dtype=np.int64)


def evolved(cls, lx, ly, seed, n, **kw):
    s = cls(lx, ly, **kw)
    s.supplySeed(seed)
    s.initHomogeneous()
    if cls is K.SchedulerSCABit:
        s.setPQ(0.95, 0.05)
        s.mcsPQSyncRng(n)
    else:
        s.mcs(n)
    s.sync()
    return s


@pytest.mark.parametrize("cls,big,small", [
    (K.SchedulerDT, (12, 16), (9, 7)),
    (K.SchedulerSCABit, (13, 16), (8, 5)),
])
def test_noalloc_matches_roughness_two_calls(cls, big, small):
    hip = Hip()
    f = noalloc_fn()
    out = hip.alloc(2 * 3 * 8)
    try:
        assert hip.h.hipMemset(out, 0xA5, 48) == 0  # stale words must be overwritten
        res = []
        for k, (lx, ly) in enumerate((big, small)):
            s = evolved(cls, lx, ly, 11 + k, 3)
            want = expected(s)
            dst = C.c_void_p(out.value + 24 * k)
            rc = f(C.c_void_p(s.device_lattice), lx, ly, s.encoding(), dst, C.c_void_p(s.stream))
            assert rc == 0, K.lib().kpz_last_error()
            assert hip.h.hipStreamSynchronize(C.c_void_p(s.stream)) == 0
            got = hip.read_i64(dst, 3)
            assert np.array_equal(got, want), (cls.__name__, lx, ly, got, want)
            res.append(got)
            s.close()
        # the first call's result is untouched by the second
        assert np.array_equal(hip.read_i64(out, 6), np.concatenate(res))
    finally:
        hip.h.hipFree(out)


def test_noalloc_flat_closed_form_and_repeat():
    """Flat lattice: sum h = -N/2, sum h^2 = N/2; three calls in a row on the
    same stream give the same words."""
    hip = Hip()
    f = noalloc_fn()
    out = hip.alloc(24)
    try:
        s = K.SchedulerDT(10, 10)
        s.initHomogeneous()
        s.sync()
        N = 1 << 20
        for _ in range(3):
            assert f(C.c_void_p(s.device_lattice), 10, 10, s.encoding(), out, C.c_void_p(s.stream)) == 0
            assert hip.h.hipStreamSynchronize(C.c_void_p(s.stream)) == 0
            assert list(hip.read_i64(out, 3)) == [-N // 2, N // 2, 0]
        s.close()
    finally:
        hip.h.hipFree(out)


def test_noalloc_rejects_sizes_beyond_the_workspace():
    hip = Hip()
    f = noalloc_fn()
    out = hip.alloc(24)
    try:
        s = K.SchedulerDT(8, 8)
        lat = C.c_void_p(s.device_lattice)
        assert f(lat, 8, 17, K.ENC_LOCAL, out, None) == -1
        assert b"workspace" in K.lib().kpz_last_error()
        assert f(None, 8, 8, K.ENC_LOCAL, out, None) == -1
        assert f(lat, 8, 8, 7, out, None) == -1
        assert f(lat, 1, 8, K.ENC_LOCAL, out, None) == -1
        s.close()
    finally:
        hip.h.hipFree(out)
