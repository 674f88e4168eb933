"""cudakpz_amd -- MI355X-native engine for the 2+1D octahedron KPZ model.

Python mirror of the CudaKpz scheduler interface (kpz/schedulerService.h:
73-255) over the C ABI of libkpz_amd.so (include/kpz_amd.h).  Every compute
call runs the hand-written gfx950 kernels; there is no CPU fallback: if the
library is missing or no GPU is present, the calls raise.

    from cudakpz_amd import SchedulerDT
    s = SchedulerDT(16, 16)            # 65536 x 65536, ENC_LOCAL
    s.supplySeed(42); s.initHomogeneous()
    s.mcs(100)
    r = s.roughness()                  # exact sums, r.w2
"""
import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkpz_amd.so")
_LIB = None

KPZ_DT, KPZ_DTR, KPZ_SCABIT, KPZ_SCA = 0, 1, 2, 3
ENC_LOCAL, ENC_CBBIT = 0, 1


class KpzError(RuntimeError):
    """Raised when a C ABI call fails (the message is kpz_last_error())."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def build(arch="gfx950", jobs=8):
    """Compile libkpz_amd.so in-tree with hipcc."""
    subprocess.check_call(["make", "-s", "-j", str(jobs), "-C", os.path.join(_HERE, "csrc"), f"ARCH={arch}"])


class _Params(C.Structure):
    _fields_ = [("threads", C.c_int64), ("layout_x", C.c_int), ("layout_y", C.c_int),
                ("skip", C.c_uint64), ("device", C.c_int), ("rng", C.c_int)]


# worker stream generators (kpz_params.rng)
RNG_SLCG64, RNG_TINYMT32 = 0, 1


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise KpzError(-2, f"{LIB_PATH} is not built; run cudakpz_amd.build()")
        L = C.CDLL(LIB_PATH)
        h = C.c_void_p
        i64p = C.POINTER(C.c_int64)
        u64p = C.POINTER(C.c_uint64)
        L.kpz_last_error.restype = C.c_char_p
        L.kpz_create.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(_Params), C.POINTER(h)]
        L.kpz_destroy.argtypes = [h]
        L.kpz_set_seed.argtypes = [h, C.c_uint64]
        L.kpz_set_pq.argtypes = [h, C.c_double, C.c_double]
        for f in ("kpz_init_flat", "kpz_sync", "kpz_corr_tag"):
            getattr(L, f).argtypes = [h]
        L.kpz_upload.argtypes = [h, C.c_void_p]
        L.kpz_download.argtypes = [h, C.c_void_p]
        L.kpz_copy_system.argtypes = [h, h]
        L.kpz_mcs.argtypes = [h, C.c_int]
        L.kpz_mcs_pq.argtypes = [h, C.c_int]
        L.kpz_roughness.argtypes = [h, i64p]
        L.kpz_moments.argtypes = [h, C.POINTER(C.c_double)]
        L.kpz_correlate.argtypes = [h, i64p]
        L.kpz_threads.argtypes = [h]
        L.kpz_threads.restype = C.c_int64
        L.kpz_get_streams.argtypes = [h, u64p]
        L.kpz_set_streams.argtypes = [h, u64p]
        L.kpz_get_master.argtypes = [h, C.c_void_p]
        L.kpz_set_master.argtypes = [h, C.c_void_p]
        L.kpz_master_draw.argtypes = [h]
        L.kpz_master_draw.restype = C.c_double
        L.kpz_randomize.argtypes = [h]
        L.kpz_mcs_count.argtypes = [h]
        L.kpz_mcs_count.restype = C.c_uint32
        L.kpz_mcs_scale.argtypes = [h]
        L.kpz_mcs_scale.restype = C.c_double
        L.kpz_encoding.argtypes = [h]
        L.kpz_size_words.argtypes = [h]
        L.kpz_size_words.restype = C.c_uint64
        L.kpz_skip.argtypes = [h]
        L.kpz_skip.restype = C.c_uint64
        L.kpz_device_lattice.argtypes = [h]
        L.kpz_device_lattice.restype = C.c_void_p
        L.kpz_stream.argtypes = [h]
        L.kpz_stream.restype = C.c_void_p
        u32p = C.POINTER(C.c_uint32)
        L.kpz_rng.argtypes = [h]
        L.kpz_set_tinymt_params.argtypes = [h, u32p, u32p, u32p]
        L.kpz_get_tinymt.argtypes = [h, u32p]
        L.kpz_set_tinymt.argtypes = [h, u32p]
        L.kpz_consistency_check.argtypes = [h, C.c_int, i64p]
        _LIB = L
    return _LIB


def _check(rc):
    if rc != 0:
        raise KpzError(rc, lib().kpz_last_error().decode())


def _i128(lo, hi):
    return (int(hi) << 64) | (int(lo) & 0xFFFFFFFFFFFFFFFF)


@dataclass
class Roughness:
    """Exact height sums (heights as Correlator::set) and the derived w^2
    (kpz/roughness.cpp:38-43)."""
    N: int
    sum_h: int
    sum_h2: int

    @property
    def h(self):
        return self.sum_h / self.N

    @property
    def w2(self):
        from fractions import Fraction
        return float(Fraction(self.sum_h2, self.N) - Fraction(self.sum_h, self.N) ** 2)


@dataclass
class Correlation:
    """Autocorrelation against the tagged snapshot (kpz/correlator.cpp:450-455):
    ch = <h_t h_s> - <h_t><h_s>, cs = slope overlap / N / 2."""
    N: int
    sum_h: int
    sum_hs: int
    sum_h2: int
    sum_hhs: int
    tcs: int

    @property
    def ch(self):
        from fractions import Fraction
        return float(Fraction(self.sum_hhs, self.N) - Fraction(self.sum_h, self.N) * Fraction(self.sum_hs, self.N))

    @property
    def cs(self):
        return self.tcs / self.N / 2.0

    @property
    def w2(self):
        from fractions import Fraction
        return float(Fraction(self.sum_h2, self.N) - Fraction(self.sum_h, self.N) ** 2)


class SchedulerService:
    """One device lattice with its worker streams and master generator."""

    KIND = None
    ENCODING = None

    def __init__(self, lx, ly, threads=0, layout=(4, 4), skip=0, device=-1, kind=None, rng=RNG_SLCG64):
        self.lx, self.ly = int(lx), int(ly)
        kind = self.KIND if kind is None else kind
        p = _Params(int(threads), int(layout[0]), int(layout[1]), int(skip), int(device), int(rng))
        h = C.c_void_p()
        _check(lib().kpz_create(kind, self.lx, self.ly, C.byref(p), C.byref(h)))
        self._h = h
        self.kind = kind

    @classmethod
    def _wrap(cls, h, kind, lx, ly):
        obj = cls.__new__(cls)
        obj.lx, obj.ly, obj._h, obj.kind = int(lx), int(ly), h, kind
        return obj

    # -- checkpoint / resume
    def saveState(self, path, aux=0):
        """Everything needed to continue bit for bit (kind, sizes, MCS count,
        p, q, master dSFMT, stream states, lattice, live tags) in one file."""
        L = lib()
        L.kpz_save_state.argtypes = [C.c_void_p, C.c_char_p, C.c_uint32]
        _check(L.kpz_save_state(self._h, os.fsencode(path), int(aux)))

    def saveH5(self, file, prefix="", prevMeasureOverride=0):
        """SchedulerService::writeH5: the reference's HDF5 datasets
        (<prefix>.system, .disorder, .dsfmt, tags with .corrTag / avgH,
        /custom attributes) plus the SLCG stream states, in
        <file>_0_0_0.h5 (or <file> if it ends in .h5)."""
        L = lib_h5()
        _check(L.kpz_write_h5(self._h, os.fsencode(file), prefix.encode(), int(prevMeasureOverride)))

    # -- lifetime
    def close(self):
        if getattr(self, "_h", None):
            lib().kpz_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- geometry
    @property
    def N(self):
        return 1 << (self.lx + self.ly)

    @property
    def size_words(self):
        return int(lib().kpz_size_words(self._h))

    @property
    def threads(self):
        return int(lib().kpz_threads(self._h))

    @property
    def skip(self):
        return int(lib().kpz_skip(self._h))

    def encoding(self):
        return int(lib().kpz_encoding(self._h))

    # -- reference API
    def supplySeed(self, seed):
        _check(lib().kpz_set_seed(self._h, int(seed) & 0xFFFFFFFFFFFFFFFF))

    def setPQ(self, p, q, n=0):
        """SchedulerService::setPQ(p, q, n): rates of disorder class n."""
        if n == 0:
            _check(lib().kpz_set_pq(self._h, float(p), float(q)))
        else:
            L = lib()
            L.kpz_set_pq_class.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_uint]
            _check(L.kpz_set_pq_class(self._h, float(p), float(q), int(n)))

    def rates(self, n=0):
        """(rateP(n), rateQ(n)): the rates disorder class n uses."""
        L = lib()
        L.kpz_get_pq_class.argtypes = [C.c_void_p, C.c_uint, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        p, q = C.c_double(), C.c_double()
        _check(L.kpz_get_pq_class(self._h, int(n), C.byref(p), C.byref(q)))
        return p.value, q.value

    # -- Disorder2 (kpz/schedulerService.cpp:48-69)
    def generateDisorder2(self, fill):
        L = lib()
        L.kpz_generate_disorder2.argtypes = [C.c_void_p, C.c_double]
        _check(L.kpz_generate_disorder2(self._h, float(fill)))

    def setDisorder(self, m):
        m = np.ascontiguousarray(m, dtype=np.uint32)
        if m.size != self.size_words:
            raise KpzError(-1, "disorder map must have the lattice's size")
        L = lib()
        L.kpz_set_disorder.argtypes = [C.c_void_p, C.c_void_p]
        _check(L.kpz_set_disorder(self._h, m.ctypes.data))

    def disorder(self):
        out = np.empty(self.size_words, np.uint32)
        L = lib()
        L.kpz_get_disorder.argtypes = [C.c_void_p, C.c_void_p]
        _check(L.kpz_get_disorder(self._h, out.ctypes.data))
        return out

    def mcsDisorder2SyncRng(self, n):
        L = lib()
        L.kpz_mcs_disorder2.argtypes = [C.c_void_p, C.c_int]
        _check(L.kpz_mcs_disorder2(self._h, int(n)))

    def mcsEW(self, n):
        """EW fast path (p = q = 1, no acceptance draw); DT schedulers only."""
        L = lib()
        L.kpz_mcs_ew.argtypes = [C.c_void_p, C.c_int]
        _check(L.kpz_mcs_ew(self._h, int(n)))

    def initHomogeneous(self):
        _check(lib().kpz_init_flat(self._h))

    def mcs(self, n):
        _check(lib().kpz_mcs(self._h, int(n)))

    def mcsSyncRng(self, n):
        self.mcs(n)

    def mcsPQSyncRng(self, n):
        _check(lib().kpz_mcs_pq(self._h, int(n)))

    def mcsPQ(self, n):
        self.mcsPQSyncRng(n)

    def mcsScale(self):
        return float(lib().kpz_mcs_scale(self._h))

    def mcsCount(self):
        return int(lib().kpz_mcs_count(self._h))

    def sync(self):
        _check(lib().kpz_sync(self._h))

    def pushSystem(self, lattice):
        a = np.ascontiguousarray(lattice, dtype=np.uint32)
        if a.size != self.size_words:
            raise ValueError(f"lattice has {a.size} words, expected {self.size_words}")
        _check(lib().kpz_upload(self._h, a.ctypes.data))

    def popSystem(self):
        a = np.empty(self.size_words, np.uint32)
        _check(lib().kpz_download(self._h, a.ctypes.data))
        return a

    def copySystem(self, other):
        _check(lib().kpz_copy_system(self._h, other._h))

    def roughness(self):
        out = (C.c_int64 * 3)()
        _check(lib().kpz_roughness(self._h, out))
        return Roughness(self.N, int(out[0]), _i128(out[1], out[2]))

    def moments(self):
        """<h^k>, k = 1..7 (Roughness::h[], kpz/roughness.cpp:38-43)."""
        out = (C.c_double * 7)()
        _check(lib().kpz_moments(self._h, out))
        return list(out)

    def consistencyCheck(self, full=True):
        """SchedulerService::consitencyCheck (kpz/schedulerService.cpp:721-854):
        (ok, (bad rows, bad columns, bad loops)) of the lattice and, with full,
        of every live correlator tag's snapshot as well."""
        out = (C.c_int64 * 3)()
        rc = lib().kpz_consistency_check(self._h, int(bool(full)), out)
        if rc < 0:
            _check(rc)
        return rc == 1, (int(out[0]), int(out[1]), int(out[2]))

    consitencyCheck = consistencyCheck  # the reference's spelling (kpz/schedulerService.h:148)

    def correlateTag(self, stop=None):
        if stop is None:
            _check(lib().kpz_corr_tag(self._h))
        else:
            lib().kpz_corr_tag_stop.argtypes = [C.c_void_p, C.c_uint32]
            _check(lib().kpz_corr_tag_stop(self._h, int(stop)))

    def correlateAll(self, max_tags=64):
        """[(tag MCS, Correlation)] for every live tag, oldest first."""
        L = lib()
        L.kpz_correlate_tags.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.c_int]
        out = (C.c_int64 * (8 * max_tags))()
        n = L.kpz_correlate_tags(self._h, out, int(max_tags))
        if n < 0:
            _check(n)
        return [(int(out[8 * k]), Correlation(self.N, int(out[8 * k + 1]), int(out[8 * k + 2]),
                                              _i128(out[8 * k + 3], out[8 * k + 4]),
                                              _i128(out[8 * k + 5], out[8 * k + 6]), int(out[8 * k + 7])))
                for k in range(n)]

    def correlate(self):
        out = (C.c_int64 * 7)()
        _check(lib().kpz_correlate(self._h, out))
        return Correlation(self.N, int(out[0]), int(out[1]), _i128(out[2], out[3]), _i128(out[4], out[5]),
                           int(out[6]))

    # -- RNG state
    def streams(self):
        a = np.empty(self.threads, np.uint64)
        _check(lib().kpz_get_streams(self._h, a.ctypes.data_as(C.POINTER(C.c_uint64))))
        return a

    def set_streams(self, states):
        a = np.ascontiguousarray(states, dtype=np.uint64)
        if a.size != self.threads:
            raise ValueError("one state per stream expected")
        _check(lib().kpz_set_streams(self._h, a.ctypes.data_as(C.POINTER(C.c_uint64))))

    @property
    def rng(self):
        return int(lib().kpz_rng(self._h))

    def setTinyMTParams(self, mat1, mat2, tmat):
        """TinyMT32 streams: parameter set t of stream t (tinyMTmat1/2/tmat,
        T words each); needed before supplySeed / randomize."""
        arrs = [np.ascontiguousarray(v, dtype=np.uint32) for v in (mat1, mat2, tmat)]
        if any(a.size != self.threads for a in arrs):
            raise ValueError("one parameter set per stream expected")
        u32p = C.POINTER(C.c_uint32)
        _check(lib().kpz_set_tinymt_params(self._h, *(a.ctypes.data_as(u32p) for a in arrs)))

    def tinymt(self):
        """TinyMT32 streams as a (7, T) array: status[0..3], mat1, mat2, tmat."""
        a = np.empty((7, self.threads), np.uint32)
        _check(lib().kpz_get_tinymt(self._h, a.ctypes.data_as(C.POINTER(C.c_uint32))))
        return a

    def set_tinymt(self, planes):
        a = np.ascontiguousarray(planes, dtype=np.uint32)
        if a.size != 7 * self.threads:
            raise ValueError("7 planes of T words expected")
        _check(lib().kpz_set_tinymt(self._h, a.ctypes.data_as(C.POINTER(C.c_uint32))))

    def master_draw(self):
        return float(lib().kpz_master_draw(self._h))

    def randomize(self):
        """Re-seed the T streams from the master (SLCG64::randomize, or
        RngIface::TinyMT32::randomize: one master draw per stream)."""
        _check(lib().kpz_randomize(self._h))

    def master_state(self):
        """master dSFMT state, laid out as the reference's dsfmt_t"""
        b = C.create_string_buffer(lib().kpz_master_bytes())
        _check(lib().kpz_get_master(self._h, b))
        return b.raw

    def set_master_state(self, state):
        n = lib().kpz_master_bytes()
        if len(state) < n:
            raise ValueError(f"master state needs {n} bytes")
        b = C.create_string_buffer(bytes(state[:n]), n)
        _check(lib().kpz_set_master(self._h, b))

    @property
    def device_lattice(self):
        return int(lib().kpz_device_lattice(self._h))

    @property
    def stream(self):
        return int(lib().kpz_stream(self._h))


class SchedulerDT(SchedulerService):
    """Random-sequential double tiling, static origin (SchedulerCPUDT)."""
    KIND = KPZ_DT
    ENCODING = ENC_LOCAL


class SchedulerDTr(SchedulerService):
    """Random-sequential double tiling, random origin (SchedulerCPUDTr)."""
    KIND = KPZ_DTR
    ENCODING = ENC_LOCAL


class SchedulerSCABit(SchedulerService):
    """Checkerboard SCA, bit-vectorised ENC_CBBIT (SchedulerCPUCBBit)."""
    KIND = KPZ_SCABIT
    ENCODING = ENC_CBBIT


class SchedulerSCA(SchedulerService):
    """Checkerboard SCA on ENC_LOCAL (SchedulerCPUCB; mcs() runs p=.95, q=0)."""
    KIND = KPZ_SCA
    ENCODING = ENC_LOCAL


class _Replica(C.Structure):
    _fields_ = [("kind", C.c_int), ("lx", C.c_int), ("ly", C.c_int), ("p", C.c_double), ("q", C.c_double),
                ("seed", C.c_uint64)]


class BatchTuning(C.Structure):
    """Performance knobs of a Batch; none of them changes any result."""
    _fields_ = [("max_lanes_log2", C.c_int), ("pq_words64", C.c_int), ("chunk", C.c_int),
                ("concurrent", C.c_int), ("sort_by_size", C.c_int)]

    @classmethod
    def default(cls):
        return cls(3, 1, 50, 2, 1)

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def _declare_batch(L):
    if getattr(L, "_batch_declared", False):
        return
    h = C.c_void_p
    L.kpz_batch_create.argtypes = [C.POINTER(_Replica), C.c_int, C.POINTER(BatchTuning), C.c_int, C.POINTER(h)]
    L.kpz_batch_destroy.argtypes = [h]
    L.kpz_batch_init_flat.argtypes = [h]
    L.kpz_batch_mcs.argtypes = [h, C.c_int]
    L.kpz_batch_roughness.argtypes = [h, C.POINTER(C.c_int64)]
    L.kpz_batch_download.argtypes = [h, C.c_int, C.c_void_p]
    L.kpz_batch_get_streams.argtypes = [h, C.c_int, C.c_void_p]
    L.kpz_batch_threads.argtypes = [h, C.c_int]
    L.kpz_batch_threads.restype = C.c_int64
    L.kpz_batch_sites.argtypes = [h]
    L.kpz_batch_sites.restype = C.c_uint64
    L.kpz_batch_sync.argtypes = [h]
    L.kpz_batch_stream.argtypes = [h]
    L.kpz_batch_stream.restype = C.c_void_p
    L.kpz_batch_autotune.argtypes = [C.POINTER(_Replica), C.c_int, C.c_int, C.c_int, C.POINTER(BatchTuning),
                                     C.POINTER(C.c_double)]
    L._batch_declared = True


def _replicas(reps):
    arr = (_Replica * len(reps))()
    for i, r in enumerate(reps):
        kind, lx, ly, p, q, seed = r
        arr[i] = _Replica(int(kind), int(lx), int(ly), float(p), float(q), int(seed) & 0xFFFFFFFFFFFFFFFF)
    return arr


class Batch:
    """Many independent replicas advanced together (hipGraph-replayed chunks).

    replicas: sequence of (kind, lx, ly, p, q, seed); kind KPZ_DT, KPZ_DTR or
    KPZ_SCABIT.  Replica i evolves exactly like the stand-alone scheduler of
    that kind, size, seed and (p, q) with default threads: DT/DTr with
    (p, q) = (1, 0) as mcs(n), everything else as mcsPQSyncRng(n)."""

    def __init__(self, replicas, tuning=None, device=-1):
        L = lib()
        _declare_batch(L)
        self.replicas = [tuple(r) for r in replicas]
        arr = _replicas(self.replicas)
        h = C.c_void_p()
        t = C.byref(tuning) if tuning is not None else None
        _check(L.kpz_batch_create(arr, len(arr), t, int(device), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().kpz_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.replicas)

    @property
    def sites(self):
        return int(lib().kpz_batch_sites(self._h))

    @property
    def stream(self):
        return int(lib().kpz_batch_stream(self._h))

    def initHomogeneous(self):
        _check(lib().kpz_batch_init_flat(self._h))

    def mcs(self, n):
        _check(lib().kpz_batch_mcs(self._h, int(n)))

    def sync(self):
        _check(lib().kpz_batch_sync(self._h))

    def roughness(self):
        n = len(self.replicas)
        out = (C.c_int64 * (3 * n))()
        _check(lib().kpz_batch_roughness(self._h, out))
        res = []
        for i, r in enumerate(self.replicas):
            N = 1 << (r[1] + r[2])
            res.append(Roughness(N, int(out[3 * i]), _i128(out[3 * i + 1], out[3 * i + 2])))
        return res

    def popSystem(self, i):
        lx, ly = self.replicas[i][1], self.replicas[i][2]
        a = np.empty(1 << (lx + ly - 4), np.uint32)
        _check(lib().kpz_batch_download(self._h, int(i), a.ctypes.data))
        return a

    def threads(self, i):
        return int(lib().kpz_batch_threads(self._h, int(i)))

    def streams(self, i):
        a = np.empty(self.threads(i), np.uint64)
        _check(lib().kpz_batch_get_streams(self._h, int(i), a.ctypes.data))
        return a

    @staticmethod
    def autotune(replicas, mcs=5, tuning=None, device=-1):
        """Returns (best BatchTuning, ms per MCS)."""
        L = lib()
        _declare_batch(L)
        arr = _replicas(replicas)
        t = BatchTuning.default() if tuning is None else tuning
        ms = C.c_double()
        _check(L.kpz_batch_autotune(arr, len(arr), int(mcs), int(device), C.byref(t), C.byref(ms)))
        return t, ms.value


class _Job(C.Structure):
    _fields_ = [("kind", C.c_int), ("lx", C.c_int), ("ly", C.c_int), ("mcs", C.c_int), ("p", C.c_double),
                ("q", C.c_double), ("seed", C.c_uint64), ("keep_lattice", C.c_int)]


@dataclass
class JobResult:
    """A finished serving job: exact sums at its sample times, timestamps
    (s on the server clock) and optionally the final lattice."""
    id: int
    N: int
    t_submit: float
    t_start: float
    t_end: float
    series: list  # [(mcs, Roughness)]
    lattice: object = None

    @property
    def w2(self):
        return [(m, r.w2) for m, r in self.series]


def _declare_server(L):
    if getattr(L, "_server_declared", False):
        return
    h = C.c_void_p
    L.kpz_server_create.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int, C.POINTER(h)]
    L.kpz_server_destroy.argtypes = [h]
    L.kpz_server_submit.argtypes = [h, C.POINTER(_Job), C.POINTER(C.c_uint64)]
    L.kpz_server_step.argtypes = [h]
    L.kpz_server_step.restype = C.c_int64
    L.kpz_server_peek.argtypes = [h, C.POINTER(C.c_int), C.POINTER(C.c_uint64)]
    L.kpz_server_pop.argtypes = [h, C.POINTER(C.c_uint64), C.POINTER(C.c_double), C.POINTER(C.c_uint32),
                                 C.POINTER(C.c_int64), C.c_void_p]
    L.kpz_server_stats.argtypes = [h, C.POINTER(C.c_uint64)]
    L.kpz_server_now.argtypes = [h]
    L.kpz_server_now.restype = C.c_double
    L._server_declared = True


class Server:
    """Serving loop: submit jobs at any time, call step() repeatedly, pop()
    finished results.  Jobs share batched launches (continuous batching) and
    are admitted as HBM frees up."""

    def __init__(self, hbm_budget=8 << 30, max_slots=256, max_tick_mcs=16, device=-1):
        L = lib()
        _declare_server(L)
        h = C.c_void_p()
        _check(L.kpz_server_create(int(hbm_budget), int(max_slots), int(max_tick_mcs), int(device), C.byref(h)))
        self._h = h
        self._specs = {}

    def close(self):
        if getattr(self, "_h", None):
            lib().kpz_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def now(self):
        return float(lib().kpz_server_now(self._h))

    def submit(self, kind, lx, ly, mcs, p, q, seed, keep_lattice=False):
        j = _Job(int(kind), int(lx), int(ly), int(mcs), float(p), float(q), int(seed) & 0xFFFFFFFFFFFFFFFF,
                 int(bool(keep_lattice)))
        i = C.c_uint64()
        _check(lib().kpz_server_submit(self._h, C.byref(j), C.byref(i)))
        self._specs[i.value] = (kind, lx, ly, mcs, p, q, seed)
        return i.value

    def step(self):
        r = lib().kpz_server_step(self._h)
        if r < 0:
            _check(int(r))
        return int(r)

    def pop(self):
        n = C.c_int()
        words = C.c_uint64()
        if not lib().kpz_server_peek(self._h, C.byref(n), C.byref(words)):
            return None
        i = C.c_uint64()
        times = (C.c_double * 3)()
        mcs = (C.c_uint32 * max(n.value, 1))()
        sums = (C.c_int64 * max(3 * n.value, 1))()
        lat = np.empty(words.value, np.uint32) if words.value else None
        _check(lib().kpz_server_pop(self._h, C.byref(i), times, mcs, sums,
                                    lat.ctypes.data if lat is not None else None))
        spec = self._specs.pop(i.value)
        N = 1 << (spec[1] + spec[2])
        series = [(int(mcs[k]), Roughness(N, int(sums[3 * k]), _i128(sums[3 * k + 1], sums[3 * k + 2])))
                  for k in range(n.value)]
        return JobResult(i.value, N, times[0], times[1], times[2], series, lat)

    def stats(self):
        out = (C.c_uint64 * 5)()
        _check(lib().kpz_server_stats(self._h, out))
        return {"queued": out[0], "active": out[1], "hbm_used": out[2], "hbm_capacity": out[3],
                "site_updates": out[4]}

    def busy(self):
        """(device seconds spent in ticks, seconds from the first tick's start
        to the last finished tick's end)"""
        L = lib()
        L.kpz_server_busy.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        g, w = C.c_double(), C.c_double()
        _check(L.kpz_server_busy(self._h, C.byref(g), C.byref(w)))
        return g.value, w.value


def make_scheduler(name, lx, ly, **kw):
    """SchedulerConfig::mkScheduler by scheduler id (kpz/kpzFrontend.cpp:93-250)."""
    table = {"dtStatic": SchedulerDT, "dt": SchedulerDTr, "scaBit": SchedulerSCABit, "sca": SchedulerSCA,
             "cb": SchedulerSCA}
    if name not in table:
        raise KpzError(-1, f"No such scheduler: {name}")
    return table[name](lx, ly, **kw)


def load_state(path, device=-1):
    """Resume a scheduler written by saveState: returns (scheduler, aux).
    The file is validated (header, sizes, checksum) before any device work."""
    L = lib()
    L.kpz_load_state.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_uint32)]
    L.kpz_kind.argtypes = [C.c_void_p]
    h = C.c_void_p()
    aux = C.c_uint32()
    _check(L.kpz_load_state(os.fsencode(path), int(device), C.byref(h), C.byref(aux)))
    L.kpz_size.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    kind = L.kpz_kind(h)
    lx, ly = C.c_int(), C.c_int()
    _check(L.kpz_size(h, C.byref(lx), C.byref(ly)))
    cls = {KPZ_DT: SchedulerDT, KPZ_DTR: SchedulerDTr, KPZ_SCABIT: SchedulerSCABit, KPZ_SCA: SchedulerSCA}[kind]
    return cls._wrap(h, kind, lx.value, ly.value), int(aux.value)


H5_LIB_PATH = os.path.join(_HERE, "libkpz_amd_h5.so")
_LIB_H5 = None
H5_MASTER, H5_STREAMS, H5_DISORDER, H5_RATES, H5_PMO = 1, 2, 4, 8, 16


def lib_h5():
    """libkpz_amd_h5.so (HDF5 checkpoints, include/kpz_amd_h5.h)."""
    global _LIB_H5
    if _LIB_H5 is None:
        lib()  # the engine library first: the HDF5 library links against it
        if not os.path.exists(H5_LIB_PATH):
            raise KpzError(-2, f"{H5_LIB_PATH} is not built (needs the HDF5 C library); run cudakpz_amd.build()")
        L = C.CDLL(H5_LIB_PATH)
        L.kpz_write_h5.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, C.c_uint32]
        L.kpz_read_h5.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.POINTER(_Params), C.POINTER(C.c_void_p),
                                  C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        _LIB_H5 = L
    return _LIB_H5


def load_h5(file, kind, prefix="", threads=0, layout=(4, 4), skip=0, device=-1, rng=RNG_SLCG64):
    """SchedulerService::readH5 on a new scheduler of `kind` (KPZ_*): the size
    and MCS count come from the file.  Returns (scheduler, prevMeasureOverride,
    restored), restored being H5_* bits (H5_STREAMS clear: the streams were
    re-seeded from the restored master, as the reference does)."""
    L = lib_h5()
    p = _Params(int(threads), int(layout[0]), int(layout[1]), int(skip), int(device), int(rng))
    h = C.c_void_p()
    pmo, restored = C.c_uint32(), C.c_uint32()
    _check(L.kpz_read_h5(os.fsencode(file), prefix.encode(), int(kind), C.byref(p), C.byref(h), C.byref(pmo),
                         C.byref(restored)))
    Lm = lib()
    Lm.kpz_size.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    lx, ly = C.c_int(), C.c_int()
    _check(Lm.kpz_size(h, C.byref(lx), C.byref(ly)))
    cls = {KPZ_DT: SchedulerDT, KPZ_DTR: SchedulerDTr, KPZ_SCABIT: SchedulerSCABit, KPZ_SCA: SchedulerSCA}[kind]
    return cls._wrap(h, kind, lx.value, ly.value), int(pmo.value), int(restored.value)


def valu_peak(device=-1, reps=3):
    """Measured integer VALU issue rates of the device in lane operations per
    second: (slow, fast) = (v_xad_u32 class: shifts, multiplies, add3, bfe,
    perm, alignbit; v_xor_b32 class: xor, add, bitop3), from 8 independent
    chains per lane on 32 blocks of 256 lanes per CU."""
    L = lib()
    L.kpz_valu_peak.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    slow, fast = C.c_double(), C.c_double()
    _check(L.kpz_valu_peak(int(device), int(reps), C.byref(slow), C.byref(fast)))
    return slow.value, fast.value


def valu_mix_rate(slow_of_8, device=-1, reps=3):
    """Measured issue rate (lane operations per second) of blocks of 8
    independent VALU instructions, `slow_of_8` (0, 2, 3, 4, 5 or 8) of them
    v_mul_lo_u32 and the rest v_xor_b32: the VALU roofline of a kernel whose
    instruction stream has that slow-class fraction."""
    L = lib()
    L.kpz_valu_mix_rate.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]
    r = C.c_double()
    _check(L.kpz_valu_mix_rate(int(device), int(reps), int(slow_of_8), C.byref(r)))
    return r.value
