"""The TinyMT32 jump ahead of the engine (orbit polynomial, x^n mod m, p(A)v)
on the host, against the oracle's step-by-step generator: the state n steps
on must equal the jumped state for every n, parameter set and seed.  Runs
without a device (kpz_tinymt_jump is host code of libkpz_amd.so)."""
import ctypes as C

import numpy as np
import pytest

import cudakpz_amd as K
from oracle import oracle as O


def jump(gen, n):
    L = K.lib()
    L.kpz_tinymt_jump.argtypes = [C.POINTER(C.c_uint32), C.c_uint64, C.POINTER(C.c_uint32)]
    g = np.ascontiguousarray(gen, dtype=np.uint32)
    out = np.zeros(4, np.uint32)
    deg = L.kpz_tinymt_jump(g.ctypes.data_as(C.POINTER(C.c_uint32)), n, out.ctypes.data_as(C.POINTER(C.c_uint32)))
    assert 0 <= deg <= 128
    return out, deg


PARAMS = [(0x8F7011EE, 0xFC78FF1F, 0x3793FDFF), (0, 0, 0), (0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF),
          (0x12345678, 0x9ABCDEF0, 0x0F1E2D3C)]


@pytest.mark.parametrize("mat", PARAMS)
@pytest.mark.parametrize("seed", [1, 42, 4294967295])
def test_jump_equals_stepping(mat, seed):
    g = O.Tmt(*mat, seed)
    start = g.g.copy()
    stepped = {}
    for n in range(1, 1301):
        g.u32()
        if n in (1, 2, 3, 64, 127, 128, 129, 200, 1000, 1300):
            stepped[n] = g.g[:4].copy()
    for n, want in stepped.items():
        got, deg = jump(start, n)
        assert np.array_equal(got, want), (n, deg)
    got, _ = jump(start, 0)
    assert np.array_equal(got, start[:4])


def test_far_jump_composes():
    # x^(a+b) = x^a x^b along the orbit: jumping a then b equals jumping a + b
    r = np.random.default_rng(3)
    for _ in range(20):
        mat = [int(v) for v in r.integers(0, 1 << 32, 3, dtype=np.uint64)]
        g = O.Tmt(*mat, int(r.integers(0, 1 << 32))).g.copy()
        a, b = int(r.integers(0, 1 << 40)), int(r.integers(0, 1 << 40))
        s1, _ = jump(g, a)
        g2 = g.copy()
        g2[:4] = s1
        s2, _ = jump(g2, b)
        s3, _ = jump(g, a + b)
        assert np.array_equal(s2, s3)


def test_zero_state_stays_zero():
    out, deg = jump([0, 0, 0, 0, 0x8F7011EE, 0xFC78FF1F, 0x3793FDFF], 12345)
    assert deg == 0 and not out.any()
