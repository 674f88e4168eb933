"""ORACLE -- test infrastructure only.

ctypes wrapper around liboracle_kpz.so, the plain-C restatement of the
CudaKpz CPU hot path (kpz_oracle.c).  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg may import this module; the product package
cudakpz_amd never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

ENC_LOCAL, ENC_CBBIT = 0, 1
UPD_KPZ, UPD_PQ, UPD_DIS, UPD_EW = 0, 1, 2, 3


def build(ref=False):
    cmd = ["make", "-s", "-C", _HERE, "liboracle_kpz.so"]
    subprocess.check_call(cmd)
    if ref:
        subprocess.check_call(["make", "-s", "-C", _HERE, "ref"])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle_kpz.so")
        if not os.path.exists(path):
            build()
        _LIB = C.CDLL(path)
        _declare(_LIB)
    return _LIB


_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


def _declare(L):
    vp = C.c_void_p
    L.kor_dsfmt_size.restype = C.c_int
    L.kor_dsfmt_init.argtypes = [vp, C.c_uint32]
    L.kor_dsfmt_close_open.argtypes = [vp]
    L.kor_dsfmt_close_open.restype = C.c_double
    L.kor_reseed.argtypes = [vp, C.c_uint32]
    L.kor_seed32.argtypes = [C.c_uint64]
    L.kor_seed32.restype = C.c_uint32
    L.kor_shuffle4.argtypes = [vp, C.POINTER(C.c_int)]
    L.kor_slcg_skip.argtypes = [C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.kor_slcg_jump.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.kor_default_skip.argtypes = [C.c_uint64]
    L.kor_default_skip.restype = C.c_uint64
    L.kor_streams_randomize.argtypes = [vp, _u64p, C.c_int64]
    L.kor_slcg_draw.argtypes = [C.POINTER(C.c_uint64), C.c_uint64, C.c_uint64]
    L.kor_slcg_profile.argtypes = [C.c_uint64, C.c_uint64, C.c_int, _u64p, _u64p, _u64p]
    L.kor_slcg_draw.restype = C.c_uint32
    for f in ("kor_init_flat_local", "kor_init_flat_cbbit"):
        getattr(L, f).argtypes = [_u32p, C.c_int, C.c_int]
    for f in ("kor_local_to_cbbit", "kor_cbbit_to_local"):
        getattr(L, f).argtypes = [_u32p, _u32p, C.c_int, C.c_int]
    L.kor_get_slope.argtypes = [_u32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    L.kor_heights.argtypes = [_u32p, C.c_int, C.c_int, C.c_int, _i32p]
    L.kor_roughness.argtypes = [_u32p, C.c_int, C.c_int, C.c_int, _i64p]
    L.kor_correlate.argtypes = [_u32p, _u32p, C.c_int, C.c_int, C.c_int, _i64p]
    L.kor_roughness_rows.argtypes = [_u32p, C.c_int, C.c_int, C.c_int, C.c_int, _i64p]
    L.kor_correlate_rows.argtypes = [_u32p, _u32p, C.c_int, C.c_int, C.c_int, C.c_int, _i64p]
    L.kor_consistent.argtypes = [_u32p, C.c_int, C.c_int, C.c_int]
    L.kor_dt_mcs.argtypes = [_u32p, C.c_int, C.c_int, C.c_int, C.c_int, _u64p, C.c_int64,
                             C.c_uint64, C.c_uint64, vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double]
    L.kor_dt_mcs_threaded.argtypes = [_u32p, C.c_int, C.c_int, C.c_int, C.c_int, _u64p, C.c_int64,
                                      C.c_uint64, C.c_uint64, vp, C.c_int, C.c_int, C.c_int, C.c_double,
                                      C.c_double, C.c_int]
    L.kor_db_mcs.argtypes = [_u32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _u64p, C.c_int64,
                             C.c_uint64, C.c_uint64, vp, C.c_int, C.c_int]
    L.kor_dt_block_offset.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.POINTER(C.c_uint), C.c_int, C.POINTER(C.c_int)]
    L.kor_cbbit_mcs.argtypes = [_u32p, C.c_int, C.c_int, _u64p, C.c_int64, C.c_uint64, C.c_uint64,
                                C.c_int, C.c_int, C.c_double, C.c_double]
    L.kor_cbbit_halfsweep_masked.argtypes = [_u32p, C.c_int, C.c_int, C.c_int, _u64p]
    L.kor_cb_mcs.argtypes = [_u32p, C.c_int, C.c_int, _u64p, C.c_int64, C.c_uint64, C.c_uint64,
                             C.c_int, C.c_double, C.c_double]
    L.kor_cbbit_mcs_threaded.argtypes = L.kor_cbbit_mcs.argtypes + [C.c_int]
    L.kor_cb_mcs_threaded.argtypes = L.kor_cb_mcs.argtypes + [C.c_int]
    L.kor_update_local.argtypes = [_u32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float,
                                   C.c_double, C.c_double]
    L.kor_update_cbbit.argtypes = [_u32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float,
                                   C.c_double, C.c_double]
    rp = C.POINTER(Rates)
    L.kor_generate_disorder2.argtypes = [vp, _u32p, C.c_uint64, C.c_double]
    L.kor_dt_mcs_rates.argtypes = [_u32p, C.c_int, C.c_int, C.c_int, C.c_int, _u64p, C.c_int64,
                                   C.c_uint64, C.c_uint64, vp, C.c_int, C.c_int, C.c_int, rp]
    L.kor_cbbit_mcs_rates.argtypes = [_u32p, C.c_int, C.c_int, _u64p, C.c_int64, C.c_uint64, C.c_uint64,
                                      C.c_int, C.c_int, rp]
    L.kor_cb_mcs_rates.argtypes = [_u32p, C.c_int, C.c_int, _u64p, C.c_int64, C.c_uint64, C.c_uint64,
                                   C.c_int, C.c_int, rp]
    L.kor_update_local_rates.argtypes = [_u32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, rp]
    L.kor_update_cbbit_rates.argtypes = [_u32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, rp]
    L.kor_rs_mcs.argtypes = [_u32p, C.c_int, C.c_int, vp, C.c_int, C.c_int, rp]
    L.kor_tmt_init.argtypes = [_u32p, C.c_uint32]
    L.kor_tmt_u32.argtypes = [_u32p]
    L.kor_tmt_u32.restype = C.c_uint32
    L.kor_tmt_float.argtypes = [_u32p]
    L.kor_tmt_float.restype = C.c_float
    L.kor_tmt_randomize.argtypes = [vp, _u32p, C.c_int64]
    L.kor_dtr_mcs_tmt.argtypes = [_u32p, C.c_int, C.c_int, C.c_int, C.c_int, _u32p, C.c_int64, vp,
                                  C.c_int, C.c_int, C.c_double, C.c_double, rp]
    L.kor_cbbit_mcs_tmt.argtypes = [_u32p, C.c_int, C.c_int, _u32p, C.c_int64, C.c_int, C.c_int,
                                    C.c_double, C.c_double, rp]
    L.kor_cb_mcs_tmt.argtypes = [_u32p, C.c_int, C.c_int, _u32p, C.c_int64, C.c_int, C.c_int,
                                 C.c_double, C.c_double, rp]


class Rates(C.Structure):
    """Rates per disorder class and the disorder map (kor_rates)."""
    _fields_ = [("p", C.c_double * 4), ("q", C.c_double * 4), ("disorder", C.c_void_p)]


def rates(p, q, disorder=None):
    """p, q: sequences of up to 4 per-class rates; disorder: uint32 map or None."""
    r = Rates()
    for k, v in enumerate(p):
        r.p[k] = v
    for k, v in enumerate(q):
        r.q[k] = v
    r._keep = disorder
    r.disorder = None if disorder is None else disorder.ctypes.data
    return r


class Dsfmt:
    """dSFMT-19937 master generator (restated)."""

    def __init__(self, seed=None, reseed=True):
        self.buf = C.create_string_buffer(lib().kor_dsfmt_size())
        if seed is not None:
            if reseed:
                lib().kor_reseed(self.buf, seed & 0xFFFFFFFF)
            else:
                lib().kor_dsfmt_init(self.buf, seed & 0xFFFFFFFF)

    @property
    def ptr(self):
        return C.cast(self.buf, C.c_void_p)

    def close_open(self):
        return lib().kor_dsfmt_close_open(self.buf)

    def shuffle4(self):
        out = (C.c_int * 4)()
        lib().kor_shuffle4(self.buf, out)
        return list(out)

    def copy(self):
        d = Dsfmt()
        C.memmove(d.buf, self.buf, len(self.buf))
        return d


def seed32(seed):
    return lib().kor_seed32(seed)


def slcg_skip(S):
    a, c = C.c_uint64(), C.c_uint64()
    lib().kor_slcg_skip(S, C.byref(a), C.byref(c))
    return a.value, c.value


def slcg_jump(a, c, k):
    ak, ck = C.c_uint64(), C.c_uint64()
    lib().kor_slcg_jump(a, c, k, C.byref(ak), C.byref(ck))
    return ak.value, ck.value


def slcg_profile(state, n, nthreads=8):
    """bit counts and the two 12-bucket histograms of n folded draws of the
    base SLCG64 from `state` (inc/slcgProfile.cpp:41-60)"""
    bits = np.zeros(32, np.uint64)
    j = np.zeros(12, np.uint64)
    j2 = np.zeros(12, np.uint64)
    lib().kor_slcg_profile(state, n, nthreads, bits, j, j2)
    return bits, j, j2


def default_skip(T):
    return lib().kor_default_skip(T)


class Streams:
    """T SLCG64 streams with the reference skip form."""

    def __init__(self, T, master, S=None):
        self.T = int(T)
        self.S = default_skip(self.T) if S is None else int(S)
        self.a, self.c = slcg_skip(self.S)
        self.states = np.zeros(self.T, np.uint64)
        lib().kor_streams_randomize(master.ptr, self.states, self.T)

    def draw(self, t):
        s = C.c_uint64(int(self.states[t]))
        r = lib().kor_slcg_draw(C.byref(s), self.a, self.c)
        self.states[t] = s.value
        return r


class Tmt:
    """One TinyMT32 generator (kor_tmt: status[4], mat1, mat2, tmat)."""

    def __init__(self, mat1, mat2, tmat, seed):
        self.g = np.array([0, 0, 0, 0, mat1, mat2, tmat], np.uint32)
        lib().kor_tmt_init(self.g, seed & 0xFFFFFFFF)

    def u32(self):
        return lib().kor_tmt_u32(self.g)

    def float(self):
        return lib().kor_tmt_float(self.g)


class TmtStreams:
    """T TinyMT32 worker generators, parameter set t for worker t, seeded from
    the master like RngIface::TinyMT32::randomize.  g[t] = status[4], mat1,
    mat2, tmat."""

    def __init__(self, T, master, mat1, mat2, tmat):
        self.T = int(T)
        self.g = np.zeros((self.T, 7), np.uint32)
        self.g[:, 4] = mat1
        self.g[:, 5] = mat2
        self.g[:, 6] = tmat
        lib().kor_tmt_randomize(master.ptr, self.g, self.T)

    def u32(self, t):
        row = np.ascontiguousarray(self.g[t])
        v = lib().kor_tmt_u32(row)
        self.g[t] = row
        return v

    def float(self, t):
        row = np.ascontiguousarray(self.g[t])
        v = lib().kor_tmt_float(row)
        self.g[t] = row
        return v


def dtr_mcs_tmt(lat, lx, ly, lbx, lby, streams, master, n, upd=UPD_KPZ, p=1.0, q=0.0, r=None):
    rc = lib().kor_dtr_mcs_tmt(lat, lx, ly, lbx, lby, streams.g, streams.T, master.ptr, n, upd, p, q,
                               None if r is None else C.byref(r))
    if rc:
        raise ValueError("invalid DT layout or missing rates")


def cbbit_mcs_tmt(lat, lx, ly, streams, n, mode, p=0.5, q=0.0, r=None):
    rc = lib().kor_cbbit_mcs_tmt(lat, lx, ly, streams.g, streams.T, n, mode, p, q,
                                 None if r is None else C.byref(r))
    if rc:
        raise ValueError("invalid CBBit size, mode or missing rates")


def cb_mcs_tmt(lat, lx, ly, streams, n, upd=UPD_PQ, p=0.95, q=0.0, r=None):
    rc = lib().kor_cb_mcs_tmt(lat, lx, ly, streams.g, streams.T, n, upd, p, q,
                              None if r is None else C.byref(r))
    if rc:
        raise ValueError("invalid CB size, update kind or missing rates")


def size_words(lx, ly):
    return 1 << (lx + ly - 4)


def flat(lx, ly, enc):
    lat = np.zeros(size_words(lx, ly), np.uint32)
    if enc == ENC_LOCAL:
        lib().kor_init_flat_local(lat, lx, ly)
    else:
        lib().kor_init_flat_cbbit(lat, lx, ly)
    return lat


def local_to_cbbit(lat, lx, ly):
    out = np.zeros_like(lat)
    lib().kor_local_to_cbbit(np.ascontiguousarray(lat), out, lx, ly)
    return out


def cbbit_to_local(lat, lx, ly):
    out = np.zeros_like(lat)
    lib().kor_cbbit_to_local(np.ascontiguousarray(lat), out, lx, ly)
    return out


def slope(lat, enc, lx, ly, x, y, d):
    return lib().kor_get_slope(lat, enc, lx, ly, x, y, d)


def heights(lat, enc, lx, ly):
    h = np.zeros((1 << ly, 1 << lx), np.int32)
    lib().kor_heights(np.ascontiguousarray(lat), enc, lx, ly, h.reshape(-1))
    return h


def _i128(lo, hi):
    return (int(hi) << 64) | (int(lo) & 0xFFFFFFFFFFFFFFFF)


def roughness(lat, enc, lx, ly):
    """(sum h, sum h^2) as exact Python ints."""
    out = np.zeros(3, np.int64)
    lib().kor_roughness(np.ascontiguousarray(lat), enc, lx, ly, out)
    return int(out[0]), _i128(out[1], out[2])


def correlate(sys, snap, enc, lx, ly):
    """(sum h, sum h_snap, sum h^2, sum h*h_snap, tcs) as exact ints."""
    out = np.zeros(7, np.int64)
    lib().kor_correlate(np.ascontiguousarray(sys), np.ascontiguousarray(snap), enc, lx, ly, out)
    return int(out[0]), int(out[1]), _i128(out[2], out[3]), _i128(out[4], out[5]), int(out[6])


def roughness_rows(lat, enc, lx, ly, nthreads=1):
    """roughness() row by row on nthreads threads, no height array"""
    out = np.zeros(3, np.int64)
    lib().kor_roughness_rows(np.ascontiguousarray(lat), enc, lx, ly, nthreads, out)
    return int(out[0]), _i128(out[1], out[2])


def correlate_rows(sys, snap, enc, lx, ly, nthreads=1):
    """correlate() row by row on nthreads threads, no height arrays"""
    out = np.zeros(7, np.int64)
    lib().kor_correlate_rows(np.ascontiguousarray(sys), np.ascontiguousarray(snap), enc, lx, ly, nthreads, out)
    return int(out[0]), int(out[1]), _i128(out[2], out[3]), _i128(out[4], out[5]), int(out[6])


def consistent(lat, enc, lx, ly):
    return bool(lib().kor_consistent(np.ascontiguousarray(lat), enc, lx, ly))


def dt_block_offset(lx, ly, lbx, lby, ds, shift, i):
    sh = (C.c_uint * 2)(*shift)
    off = (C.c_int * 2)()
    lib().kor_dt_block_offset(lx, ly, lbx, lby, ds, sh, i, off)
    return off[0], off[1]


def dt_mcs(lat, lx, ly, lbx, lby, streams, master, n, random_origin=False, upd=UPD_KPZ, p=1.0, q=0.0):
    rc = lib().kor_dt_mcs(lat, lx, ly, lbx, lby, streams.states, streams.T, streams.a, streams.c,
                          master.ptr, n, int(random_origin), upd, p, q)
    if rc:
        raise ValueError("invalid DT layout")


def dt_mcs_threaded(lat, lx, ly, lbx, lby, streams, master, n, nthreads, random_origin=False, upd=UPD_KPZ,
                    p=1.0, q=0.0):
    rc = lib().kor_dt_mcs_threaded(lat, lx, ly, lbx, lby, streams.states, streams.T, streams.a, streams.c,
                                   master.ptr, n, int(random_origin), upd, p, q, int(nthreads))
    if rc:
        raise ValueError("invalid DT layout")


def db_mcs(lat, lx, ly, lbx, lby, streams, master, n, lorest=0, nthreads=1):
    """SchedulerCPUDB::mcs (dead border, dynamic origin) with an explicit layout."""
    rc = lib().kor_db_mcs(lat, lx, ly, lbx, lby, lorest, streams.states, streams.T, streams.a, streams.c,
                          master.ptr, n, int(nthreads))
    if rc:
        raise ValueError("invalid DB layout or stream count")


def cbbit_mcs(lat, lx, ly, streams, n, mode, p=0.5, q=0.0):
    rc = lib().kor_cbbit_mcs(lat, lx, ly, streams.states, streams.T, streams.a, streams.c, n, mode, p, q)
    if rc:
        raise ValueError("invalid CBBit size")


def cbbit_mcs_threaded(lat, lx, ly, streams, n, mode, p=0.5, q=0.0, nthreads=1):
    rc = lib().kor_cbbit_mcs_threaded(lat, lx, ly, streams.states, streams.T, streams.a, streams.c, n, mode, p, q,
                                      int(nthreads))
    if rc:
        raise ValueError("invalid CBBit size or mode")


def cb_mcs_threaded(lat, lx, ly, streams, n, p=0.95, q=0.0, nthreads=1):
    rc = lib().kor_cb_mcs_threaded(lat, lx, ly, streams.states, streams.T, streams.a, streams.c, n, p, q,
                                   int(nthreads))
    if rc:
        raise ValueError("invalid CB size")


def cbbit_halfsweep_masked(lat, lx, ly, signum, masks):
    lib().kor_cbbit_halfsweep_masked(lat, lx, ly, signum, np.ascontiguousarray(masks, np.uint64))


def cb_mcs(lat, lx, ly, streams, n, p=0.95, q=0.0):
    rc = lib().kor_cb_mcs(lat, lx, ly, streams.states, streams.T, streams.a, streams.c, n, p, q)
    if rc:
        raise ValueError("invalid CB size")


def update_local(lat, lx, ly, x, y, upd, rnd, p, q):
    lib().kor_update_local(lat, lx, ly, x, y, upd, rnd, p, q)


def update_cbbit(lat, lx, ly, x, y, rnd, p, q):
    lib().kor_update_cbbit(lat, lx, ly, x, y, rnd, p, q)


def generate_disorder2(master, size_words, fill):
    dis = np.zeros(size_words, np.uint32)
    lib().kor_generate_disorder2(master.ptr, dis, size_words, fill)
    return dis


def dt_mcs_rates(lat, lx, ly, lbx, lby, streams, master, n, random_origin, upd, r):
    rc = lib().kor_dt_mcs_rates(lat, lx, ly, lbx, lby, streams.states, streams.T, streams.a, streams.c,
                                master.ptr, n, int(random_origin), upd, C.byref(r))
    if rc:
        raise ValueError("invalid DT layout")


def cbbit_mcs_rates(lat, lx, ly, streams, n, mode, r):
    rc = lib().kor_cbbit_mcs_rates(lat, lx, ly, streams.states, streams.T, streams.a, streams.c, n, mode,
                                   C.byref(r))
    if rc:
        raise ValueError("invalid CBBit size or mode")


def cb_mcs_rates(lat, lx, ly, streams, n, upd, r):
    rc = lib().kor_cb_mcs_rates(lat, lx, ly, streams.states, streams.T, streams.a, streams.c, n, upd,
                                C.byref(r))
    if rc:
        raise ValueError("invalid CB size or update kind")


RS_MCS, RS_PQ, RS_SYNC, RS_PQ_SYNC, RS_DIS, RS_DIS_SYNC = range(6)


def rs_mcs(lat, lx, ly, master, n, mode, r=None):
    """plain random-sequential SchedulerCPU (master dSFMT draws only)"""
    rc = lib().kor_rs_mcs(lat, lx, ly, master.ptr, n, mode, None if r is None else C.byref(r))
    if rc:
        raise ValueError("invalid size, mode or missing disorder map")


def update_local_rates(lat, lx, ly, x, y, upd, rnd, r):
    lib().kor_update_local_rates(lat, lx, ly, x, y, upd, rnd, C.byref(r))


def update_cbbit_rates(lat, lx, ly, x, y, rnd, r):
    lib().kor_update_cbbit_rates(lat, lx, ly, x, y, rnd, C.byref(r))
