"""The oracle's lattice, update rules and MCS compositions against an
independently written explicit-height model (tests/heightmodel.py) and
closed forms.  Small sizes only: these run on the CPU in seconds."""
import math
import numpy as np
import pytest

from heightmodel import HeightLattice, normalized, thresh23
from oracle import oracle as O

ENCS = [O.ENC_LOCAL, O.ENC_CBBIT]


class Cur:
    """random() and the randomFloat() numerator of either stream set: SLCG64
    (fold & 0x7FFFFF, 23 bits) or TinyMT32 (tempered >> 8, 24 bits)"""

    def __init__(self, st):
        self.st, self.T = st, st.T
        self.tmt = isinstance(st, O.TmtStreams)
        self.bits = 24 if self.tmt else 23

    def draw(self, t):
        return self.st.u32(t) if self.tmt else self.st.draw(t)

    def rf(self, t):
        return self.draw(t) >> 8 if self.tmt else self.draw(t) & 0x7FFFFF

    def thresh(self, p):
        if p <= 0:
            return 0
        if p >= 1:
            return 1 << self.bits
        return math.ceil(p * (1 << self.bits))


def tmt_streams(T, master, seed):
    r = np.random.default_rng(seed)
    m = [r.integers(0, 1 << 32, T, dtype=np.uint64).astype(np.uint32) for _ in range(3)]
    return O.TmtStreams(T, master, *m)


def evolved_local(lx, ly, seed, mcs=3):
    """a rough consistent ENC_LOCAL lattice from the oracle's SCA dynamics"""
    m = O.Dsfmt(seed)
    st = O.Streams(3, m)
    lat = O.flat(lx, ly, O.ENC_LOCAL)
    O.cb_mcs(lat, lx, ly, st, mcs, 0.7, 0.3)
    return lat


@pytest.mark.parametrize("lx,ly", [(6, 6), (7, 6), (8, 8), (6, 9)])
@pytest.mark.parametrize("enc", ENCS)
def test_flat_closed_form(lx, ly, enc):
    lat = O.flat(lx, ly, enc)
    N = 1 << (lx + ly)
    h = O.heights(lat, enc, lx, ly)
    # checkerboard of {-1, 0}: even sites are minima one below their neighbours
    yy, xx = np.mgrid[0:1 << ly, 0:1 << lx]
    assert np.array_equal(h, np.where((xx + yy) % 2 == 0, -1, 0))
    s1, s2 = O.roughness(lat, enc, lx, ly)
    assert s1 == -N // 2 and s2 == N // 2  # <h> = -1/2, w^2 = 1/4
    assert O.consistent(lat, enc, lx, ly)


def test_flat_encodings_agree():
    for lx, ly in [(6, 6), (8, 7)]:
        assert np.array_equal(O.local_to_cbbit(O.flat(lx, ly, 0), lx, ly), O.flat(lx, ly, 1))


def test_one_full_halfsweep_lifts_every_minimum():
    lx, ly = 8, 6
    N = 1 << (lx + ly)
    lat = O.flat(lx, ly, O.ENC_CBBIT)
    masks = np.full(1 << (lx + ly - 7), ~np.uint64(0), np.uint64)
    O.cbbit_halfsweep_masked(lat, lx, ly, 0, masks)
    s1, s2 = O.roughness(lat, O.ENC_CBBIT, lx, ly)
    assert s1 == N // 2 and s2 == N // 2  # heights {0, +1}: w^2 = 1/4
    # the other sublattice now holds the minima; lifting it restores the flat
    # shape two units higher, and because the convention anchors
    # h(0, L-1) = 0 (a sublattice-1 site) the reported heights are {-1, 0}
    O.cbbit_halfsweep_masked(lat, lx, ly, 1, masks)
    s1b, s2b = O.roughness(lat, O.ENC_CBBIT, lx, ly)
    assert s1b == -N // 2 and s2b == N // 2


def test_heights_follow_slope_definition():
    lx, ly = 7, 6
    lat = evolved_local(lx, ly, 5)
    assert O.consistent(lat, 0, lx, ly)
    h = O.heights(lat, 0, lx, ly).astype(np.int64)
    X, Y = 1 << lx, 1 << ly
    for y in range(Y):
        for x in range(X):
            sx = 1 if O.slope(lat, 0, lx, ly, x, y, 0) else -1
            sy = 1 if O.slope(lat, 0, lx, ly, x, y, 1) else -1
            assert h[y, x] - h[y, (x - 1) % X] == sx
            assert h[y, x] - h[(y - 1) % Y, x] == sy
    # the convention anchors h(0, L-1) = 0 (kpz/correlator.cpp:50-64)
    assert h[Y - 1, 0] == 0


def test_encoding_roundtrip_and_reductions_agree():
    lx, ly = 8, 7
    lat = evolved_local(lx, ly, 11)
    cb = O.local_to_cbbit(lat, lx, ly)
    assert np.array_equal(O.cbbit_to_local(cb, lx, ly), lat)
    assert np.array_equal(O.heights(lat, 0, lx, ly), O.heights(cb, 1, lx, ly))
    assert O.roughness(lat, 0, lx, ly) == O.roughness(cb, 1, lx, ly)


def test_correlate_brute_force():
    lx, ly = 7, 7
    a = evolved_local(lx, ly, 1, 2)
    b = evolved_local(lx, ly, 2, 4)
    ha = O.heights(a, 0, lx, ly).astype(np.int64)
    hb = O.heights(b, 0, lx, ly).astype(np.int64)
    s1, t1, s2, s11, tcs = O.correlate(a, b, 0, lx, ly)
    assert s1 == ha.sum() and t1 == hb.sum()
    assert s2 == int((ha * ha).sum()) and s11 == int((ha * hb).sum())
    eq = 0
    for y in range(1 << ly):
        for x in range(1 << lx):
            for d in range(2):
                eq += 1 if O.slope(a, 0, lx, ly, x, y, d) == O.slope(b, 0, lx, ly, x, y, d) else -1
    assert tcs == eq
    # identical lattices: tcs = 2N, and the CBBit path agrees
    assert O.correlate(a, a, 0, lx, ly)[4] == 2 << (lx + ly)
    ca, cb = O.local_to_cbbit(a, lx, ly), O.local_to_cbbit(b, lx, ly)
    assert O.correlate(ca, cb, 1, lx, ly) == (s1, t1, s2, s11, tcs)


@pytest.mark.parametrize("lx,ly,nthreads", [(7, 7, 1), (8, 6, 3), (6, 9, 8), (9, 8, 5)])
def test_row_streamed_reductions_equal_height_field_sums(lx, ly, nthreads):
    # the row-by-row threaded sums (used where a height field would not fit in
    # memory) against the explicit-height ones, on evolved lattices of both
    # encodings and on random words (inconsistent lattices: the row carry
    # then differs from the y slope, and both walks must still agree)
    a = evolved_local(lx, ly, 5, 3)
    b = evolved_local(lx, ly, 6, 5)
    r = np.random.default_rng(lx * 100 + ly)
    ra = r.integers(0, 1 << 32, a.size, dtype=np.uint64).astype(np.uint32)
    rb = r.integers(0, 1 << 32, a.size, dtype=np.uint64).astype(np.uint32)
    for x, y in [(a, b), (ra, rb), (a, a)]:
        for enc in ENCS:
            u = x if enc == O.ENC_LOCAL else O.local_to_cbbit(x, lx, ly)
            v = y if enc == O.ENC_LOCAL else O.local_to_cbbit(y, lx, ly)
            assert O.roughness_rows(u, enc, lx, ly, nthreads) == O.roughness(u, enc, lx, ly)
            assert O.correlate_rows(u, v, enc, lx, ly, nthreads) == O.correlate(u, v, enc, lx, ly)


def test_inconsistent_lattice_detected():
    lx, ly = 6, 6
    lat = O.flat(lx, ly, 0)
    lat[3] ^= 1 << 2
    assert not O.consistent(lat, 0, lx, ly)


@pytest.mark.parametrize("upd", [O.UPD_KPZ, O.UPD_PQ])
def test_local_update_rule_vs_height_model(upd):
    lx, ly = 6, 6
    lat = evolved_local(lx, ly, 3)
    hm = HeightLattice(O.heights(lat, 0, lx, ly))
    rng = np.random.default_rng(7)
    p, q = (1.0, 0.0) if upd == O.UPD_KPZ else (0.6, 0.35)
    tp, tq = thresh23(p), thresh23(q)
    for _ in range(3000):
        x, y = int(rng.integers(64)), int(rng.integers(64))
        r = int(rng.integers(1 << 23))
        O.update_local(lat, lx, ly, x, y, upd, r / 8388608.0, p, q)
        if upd == O.UPD_KPZ:
            hm.update(x, y, True, False)
        else:
            hm.update(x, y, r < tp, r < tq)
    assert np.array_equal(normalized(O.heights(lat, 0, lx, ly)), hm.normalized())
    assert O.consistent(lat, 0, lx, ly)


def test_cbbit_update_rule_vs_height_model():
    lx, ly = 7, 6
    lat = O.local_to_cbbit(evolved_local(lx, ly, 4), lx, ly)
    hm = HeightLattice(O.heights(lat, 1, lx, ly))
    rng = np.random.default_rng(8)
    p, q = 0.8, 0.45
    tp, tq = thresh23(p), thresh23(q)
    for _ in range(3000):
        x, y = int(rng.integers(128)), int(rng.integers(64))
        r = int(rng.integers(1 << 23))
        O.update_cbbit(lat, lx, ly, x, y, r / 8388608.0, p, q)
        hm.update(x, y, r < tp, r < tq)
    assert np.array_equal(normalized(O.heights(lat, 1, lx, ly)), hm.normalized())


def test_threshold_is_exact_float_comparison():
    # randomFloat() < p evaluated as float-vs-double, exactly as the reference
    for p in [0.5, 0.95, 0.05, 1.0 / 3.0, 0.1, 1.0, 0.0, 1e-9, 0.999999999]:
        t = thresh23(p)
        for r in {max(t - 1, 0), t, min(t + 1, (1 << 23) - 1), 0, (1 << 23) - 1}:
            assert (np.float64(np.float32(r / 8388608.0)) < p) == (r < t)


@pytest.mark.parametrize("lx,ly", [(7, 5), (8, 4), (6, 4)])
@pytest.mark.parametrize("signum", [0, 1])
def test_bitparallel_halfsweep_equals_scalar_in_any_order(lx, ly, signum):
    lat0 = O.local_to_cbbit(evolved_local(lx, ly, 9 + lx), lx, ly)
    wbits = 64 if lx >= 7 else 32
    nw = 1 << (lx + ly - (7 if wbits == 64 else 6))
    rng = np.random.default_rng(lx * 10 + signum)
    masks = rng.integers(0, 1 << 63, nw, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, nw, dtype=np.uint64)
    if wbits == 32:
        masks &= np.uint64(0xFFFFFFFF)
    a = lat0.copy()
    O.cbbit_halfsweep_masked(a, lx, ly, signum, masks)
    X = 1 << lx
    W = X // (128 if wbits == 64 else 64)
    sites = []
    for y in range(1 << ly):
        par = signum ^ (y & 1)
        for xw in range(W):
            mw = int(masks[y * W + xw])
            for k in range(wbits):
                if (mw >> k) & 1:
                    sites.append((xw * 2 * wbits + 2 * k + par, y))
    for order in range(2):
        b = lat0.copy()
        idx = np.arange(len(sites))
        if order:
            rng.shuffle(idx)
        for i in idx:
            x, y = sites[i]
            O.update_cbbit(b, lx, ly, x, y, 0.0, 1.0, 0.0)
        assert np.array_equal(a, b)


# ---------------------------------------------------------------- compositions
def _dt_block(lx, ly, lbx, lby, ds, i):
    hx = 1 << (lx - lbx - 1)
    return ((2 * (i % hx) + (ds & 1)) << lbx, (2 * (i // hx) + (ds >> 1)) << lby)


def py_dt(hm, lx, ly, lbx, lby, st, master, n, random_origin, upd, p, q):
    X, Y = 1 << lx, 1 << ly
    nb = 1 << (lx - lbx + ly - lby - 2)
    bs = 1 << (lbx + lby)
    st = Cur(st)
    tp, tq = st.thresh(p), st.thresh(q)
    for _ in range(n):
        seq = master.shuffle4()
        sx = sy = 0
        if random_origin:
            sx = int(master.close_open() * (1 << lbx))
            sy = int(master.close_open() * (1 << lby))
        for ds in seq:
            for t in range(st.T):
                for i in range(t, nb, st.T):
                    ox, oy = _dt_block(lx, ly, lbx, lby, ds, i)
                    for _k in range(bs):
                        if random_origin:
                            x = ((st.rf(t) >> (st.bits - lbx)) + ox + sx) % X
                            y = ((st.rf(t) >> (st.bits - lby)) + oy + sy) % Y
                        else:
                            x = (st.draw(t) >> (32 - lbx)) + ox
                            y = (st.draw(t) >> (32 - lby)) + oy
                        if upd == O.UPD_KPZ:
                            hm.update(x, y, True, False)
                        else:
                            r = st.rf(t)
                            hm.update(x, y, r < tp, r < tq)


@pytest.mark.parametrize("random_origin", [False, True])
@pytest.mark.parametrize("upd,p,q", [(O.UPD_KPZ, 1.0, 0.0), (O.UPD_PQ, 0.5, 0.5), (O.UPD_PQ, 0.9, 0.2)])
@pytest.mark.parametrize("T", [1, 3])
def test_dt_composition(random_origin, upd, p, q, T):
    lx, ly, lbx, lby = 6, 5, 2, 2
    m1, m2 = O.Dsfmt(42), O.Dsfmt(42)
    s1, s2 = O.Streams(T, m1), O.Streams(T, m2)
    lat = O.flat(lx, ly, 0)
    hm = HeightLattice(O.heights(lat, 0, lx, ly))
    O.dt_mcs(lat, lx, ly, lbx, lby, s1, m1, 2, random_origin, upd, p, q)
    py_dt(hm, lx, ly, lbx, lby, s2, m2, 2, random_origin, upd, p, q)
    assert np.array_equal(normalized(O.heights(lat, 0, lx, ly)), hm.normalized())
    assert np.array_equal(s1.states, s2.states)
    assert m1.close_open() == m2.close_open()
    assert O.consistent(lat, 0, lx, ly)


def py_db(hm, lx, ly, lbx, lby, lorest, st, master, n):
    """SchedulerCPUDB::mcs written from kpz/schedulerCPUDB.cpp:65-88,
    kpz/schedulerCPUDB_nested.h:56-80 and inc/dynamicOriginDD.h:51-75 on
    explicit heights: dead border of one site, origin drawn per MCS."""
    X, Y = 1 << lx, 1 << ly
    bx, by = 1 << lbx, 1 << lby
    nbx = X // bx
    nb = nbx * (Y // by)
    for _ in range(n):
        ox = int(master.close_open() * (1 << (lbx - lorest))) << lorest
        oy = int(master.close_open() * (1 << (lby - lorest))) << lorest
        for t in range(st.T):
            for i in range(t, nb, st.T):
                offx = (i % nbx) * bx + ox
                offy = (i // nbx) * by + oy
                for _k in range((bx - 1) * (by - 1)):
                    # (int)(close_open() * m) with close_open() = draw / 2^32
                    x = (math.floor(st.draw(t) / 4294967296.0 * (bx - 1)) + offx) % X
                    y = (math.floor(st.draw(t) / 4294967296.0 * (by - 1)) + offy) % Y
                    hm.update(x, y, True, False)


@pytest.mark.parametrize("lx,ly,lbx,lby,lorest,T,nthreads", [
    (6, 5, 2, 2, 0, 1, 1), (6, 5, 3, 2, 0, 3, 1), (6, 6, 3, 3, 1, 4, 2), (5, 4, 5, 4, 0, 1, 1),
    (6, 5, 2, 3, 2, 16, 3)])
def test_db_composition(lx, ly, lbx, lby, lorest, T, nthreads):
    m1, m2 = O.Dsfmt(7 + T), O.Dsfmt(7 + T)
    s1, s2 = O.Streams(T, m1), O.Streams(T, m2)
    lat = O.flat(lx, ly, 0)
    hm = HeightLattice(O.heights(lat, 0, lx, ly))
    O.db_mcs(lat, lx, ly, lbx, lby, s1, m1, 3, lorest, nthreads)
    py_db(hm, lx, ly, lbx, lby, lorest, s2, m2, 3)
    assert np.array_equal(normalized(O.heights(lat, 0, lx, ly)), hm.normalized())
    assert np.array_equal(s1.states, s2.states)
    assert m1.close_open() == m2.close_open()
    assert O.consistent(lat, 0, lx, ly)
    # the heights moved: (2^lb - 1)^2 updates per block per MCS
    assert np.any(normalized(O.heights(lat, 0, lx, ly)) != normalized(O.heights(O.flat(lx, ly, 0), 0, lx, ly)))


def test_db_rejects_bad_layouts():
    m = O.Dsfmt(1)
    lat = O.flat(5, 4, 0)
    for args in [(1, 2, 0, 1), (2, 2, 3, 1), (6, 2, 0, 1), (3, 3, 0, 9)]:  # lbx, lby, lorest, T (8 blocks)
        st = O.Streams(args[3], m)
        with pytest.raises(ValueError):
            O.db_mcs(lat, 5, 4, args[0], args[1], st, m, 1, args[2])


def py_cbbit(hm, lx, ly, st, n, mode, p, q):
    X, Y = 1 << lx, 1 << ly
    wbits = 64 if lx >= 7 else 32
    st = Cur(st)
    tp, tq = st.thresh(p), st.thresh(q)
    per, rem = Y // st.T, Y % st.T
    for _ in range(n):
        for s in range(2):
            lo = 0
            for t in range(st.T):
                hi = lo + per + (1 if t < rem else 0)
                for y in range(lo, hi):
                    par = s ^ (y & 1)
                    if mode == 0:
                        for xw in range(X // (2 * wbits)):
                            if wbits == 64:
                                a = st.draw(t)
                                b = st.draw(t)
                                rnd = (a << 32) ^ b
                            else:
                                rnd = st.draw(t)
                            for k in range(wbits):
                                if (rnd >> k) & 1:
                                    hm.update(xw * 2 * wbits + 2 * k + par, y, True, False)
                    else:
                        for x in range(par, X, 2):
                            r = st.rf(t)
                            hm.update(x, y, r < tp, r < tq)
                lo = hi


@pytest.mark.parametrize("mode,p,q", [(0, 0.5, 0.0), (1, 0.95, 0.05), (1, 0.5, 0.5), (1, 1.0, 0.0)])
@pytest.mark.parametrize("lx,ly,T", [(7, 4, 1), (8, 3, 3), (6, 4, 5)])
def test_cbbit_composition(mode, p, q, lx, ly, T):
    m1, m2 = O.Dsfmt(7), O.Dsfmt(7)
    s1, s2 = O.Streams(T, m1), O.Streams(T, m2)
    lat = O.flat(lx, ly, 1)
    hm = HeightLattice(O.heights(lat, 1, lx, ly))
    O.cbbit_mcs(lat, lx, ly, s1, 2, mode, p, q)
    py_cbbit(hm, lx, ly, s2, 2, mode, p, q)
    assert np.array_equal(normalized(O.heights(lat, 1, lx, ly)), hm.normalized())
    assert np.array_equal(s1.states, s2.states)
    assert O.consistent(lat, 1, lx, ly)


def py_cb(hm, lx, ly, st, n, p, q):
    X, Y = 1 << lx, 1 << ly
    st = Cur(st)
    tp, tq = st.thresh(p), st.thresh(q)
    dimxw = X // 4
    per, rem = (dimxw // st.T) * 4, dimxw % st.T

    def loop(t, s, x0, x1):
        for yw in range(0, Y, 4):
            for xw in range(x0, x1, 4):
                for y in range(yw, yw + 4):
                    x = xw + ((y & 1) ^ s)
                    for xx in (x, x + 2):
                        r = st.rf(t)
                        hm.update(xx, y, r < tp, r < tq)

    for _ in range(n):
        for s in range(2):
            x1 = 0
            for t in range(st.T):
                x0 = x1
                x1 = x0 + per + (4 if t < rem else 0)
                xc = x0 + (((x1 - x0) // 2) & ~3)
                loop(t, s, x0, xc)
                loop(t, s, xc, x1)


@pytest.mark.parametrize("T", [1, 3, 8])
def test_cb_composition(T):
    lx, ly = 5, 4
    m1, m2 = O.Dsfmt(3), O.Dsfmt(3)
    s1, s2 = O.Streams(T, m1), O.Streams(T, m2)
    lat = O.flat(lx, ly, 0)
    hm = HeightLattice(O.heights(lat, 0, lx, ly))
    O.cb_mcs(lat, lx, ly, s1, 2, 0.8, 0.3)
    py_cb(hm, lx, ly, s2, 2, 0.8, 0.3)
    assert np.array_equal(normalized(O.heights(lat, 0, lx, ly)), hm.normalized())
    assert np.array_equal(s1.states, s2.states)


# ---- the TinyMT32 workers of SchedulerCPUDTr / CBBit / CB (kpz/kpzFrontend.cpp:308-331)
@pytest.mark.parametrize("upd,p,q", [(O.UPD_KPZ, 1.0, 0.0), (O.UPD_PQ, 0.5, 0.5), (O.UPD_PQ, 0.9, 0.2)])
@pytest.mark.parametrize("T", [1, 3])
def test_dtr_tinymt_composition(upd, p, q, T):
    lx, ly, lbx, lby = 6, 5, 2, 2
    m1, m2 = O.Dsfmt(42), O.Dsfmt(42)
    s1, s2 = tmt_streams(T, m1, 5), tmt_streams(T, m2, 5)
    lat = O.flat(lx, ly, 0)
    hm = HeightLattice(O.heights(lat, 0, lx, ly))
    O.dtr_mcs_tmt(lat, lx, ly, lbx, lby, s1, m1, 2, upd, p, q)
    py_dt(hm, lx, ly, lbx, lby, s2, m2, 2, True, upd, p, q)
    assert np.array_equal(normalized(O.heights(lat, 0, lx, ly)), hm.normalized())
    assert np.array_equal(s1.g, s2.g)
    assert O.consistent(lat, 0, lx, ly)


def test_tinymt_dtr_disorder_needs_rates():
    # (the oracle offers TinyMT32 workers for the random-origin DT only:
    # SchedulerCPUDT draws close_open(), which RngIface::TinyMT32 lacks)
    m = O.Dsfmt(1)
    st = tmt_streams(2, m, 1)
    lat = O.flat(6, 6, 0)
    rc = O.lib().kor_dtr_mcs_tmt(lat, 6, 6, 2, 2, st.g, st.T, m.ptr, 1, O.UPD_DIS, 1.0, 0.0, None)
    assert rc == -1  # disorder rule without rates


@pytest.mark.parametrize("mode,p,q", [(0, 0.5, 0.0), (1, 0.95, 0.05), (1, 0.5, 0.5)])
@pytest.mark.parametrize("lx,ly,T", [(7, 4, 1), (8, 3, 3), (6, 4, 5)])
def test_cbbit_tinymt_composition(mode, p, q, lx, ly, T):
    m1, m2 = O.Dsfmt(7), O.Dsfmt(7)
    s1, s2 = tmt_streams(T, m1, 9), tmt_streams(T, m2, 9)
    lat = O.flat(lx, ly, 1)
    hm = HeightLattice(O.heights(lat, 1, lx, ly))
    O.cbbit_mcs_tmt(lat, lx, ly, s1, 2, mode, p, q)
    py_cbbit(hm, lx, ly, s2, 2, mode, p, q)
    assert np.array_equal(normalized(O.heights(lat, 1, lx, ly)), hm.normalized())
    assert np.array_equal(s1.g, s2.g)
    assert O.consistent(lat, 1, lx, ly)


@pytest.mark.parametrize("T", [1, 3, 8])
def test_cb_tinymt_composition(T):
    lx, ly = 5, 4
    m1, m2 = O.Dsfmt(3), O.Dsfmt(3)
    s1, s2 = tmt_streams(T, m1, 2), tmt_streams(T, m2, 2)
    lat = O.flat(lx, ly, 0)
    hm = HeightLattice(O.heights(lat, 0, lx, ly))
    O.cb_mcs_tmt(lat, lx, ly, s1, 2, O.UPD_PQ, 0.8, 0.3)
    py_cb(hm, lx, ly, s2, 2, 0.8, 0.3)
    assert np.array_equal(normalized(O.heights(lat, 0, lx, ly)), hm.normalized())
    assert np.array_equal(s1.g, s2.g)


def test_tinymt_threshold_has_24_bits():
    # randomFloat() of TinyMT32 is (u >> 8) * 2^-24: a rate on the 2^-24 grid
    # between two 2^-23 steps separates the two scales
    p = 0.5 + 2.0 ** -24
    assert Cur.thresh(type("S", (), {"bits": 24})(), p) == (1 << 23) + 1
    assert Cur.thresh(type("S", (), {"bits": 23})(), p) == (1 << 22) + 1
    r24 = 1 << 23  # randomFloat = 0.5 exactly
    assert (np.float64(np.float32(r24 * 2.0 ** -24)) < p) == (r24 < (1 << 23) + 1)


def test_dt_invalid_layout_rejected():
    m = O.Dsfmt(1)
    st = O.Streams(1, m)
    lat = O.flat(6, 6, 0)
    with pytest.raises(ValueError):
        O.dt_mcs(lat, 6, 6, 6, 2, st, m, 1)  # one block per row: no double tiling
    with pytest.raises(ValueError):
        O.dt_mcs(lat, 6, 6, 1, 2, st, m, 1)  # block smaller than the reference minimum


@pytest.mark.parametrize("T,nthreads", [(8, 8), (5, 3), (64, 4)])
def test_threaded_dt_equals_sequential(T, nthreads):
    lx, ly = 8, 8
    m1, m2 = O.Dsfmt(9), O.Dsfmt(9)
    s1, s2 = O.Streams(T, m1), O.Streams(T, m2)
    a, b = O.flat(lx, ly, 0), O.flat(lx, ly, 0)
    O.dt_mcs(a, lx, ly, 3, 3, s1, m1, 3, True, O.UPD_PQ, 0.8, 0.3)
    O.dt_mcs_threaded(b, lx, ly, 3, 3, s2, m2, 3, nthreads, True, O.UPD_PQ, 0.8, 0.3)
    assert np.array_equal(a, b) and np.array_equal(s1.states, s2.states)


@pytest.mark.parametrize("lx,ly,T,nthreads", [(7, 5, 32, 4), (8, 6, 5, 3), (10, 9, 512, 8), (6, 4, 16, 16)])
@pytest.mark.parametrize("mode,p,q", [(0, 0.5, 0.0), (1, 0.9, 0.07)])
def test_threaded_cbbit_equals_sequential(lx, ly, T, nthreads, mode, p, q):
    m1, m2 = O.Dsfmt(4), O.Dsfmt(4)
    s1, s2 = O.Streams(T, m1), O.Streams(T, m2)
    a, b = O.flat(lx, ly, O.ENC_CBBIT), O.flat(lx, ly, O.ENC_CBBIT)
    O.cbbit_mcs(a, lx, ly, s1, 3, mode, p, q)
    O.cbbit_mcs_threaded(b, lx, ly, s2, 3, mode, p, q, nthreads)
    assert np.array_equal(a, b) and np.array_equal(s1.states, s2.states)
    assert not np.array_equal(a, O.flat(lx, ly, O.ENC_CBBIT))


# T = 2 and 3 leave one thread; 256 workers of one cell column each over 8
# threads is the shape of the default SCA engine
@pytest.mark.parametrize("lx,ly,T,nthreads", [(5, 4, 2, 4), (6, 5, 3, 2), (10, 8, 256, 8), (9, 7, 13, 5)])
def test_threaded_cb_equals_sequential(lx, ly, T, nthreads):
    m1, m2 = O.Dsfmt(6), O.Dsfmt(6)
    s1, s2 = O.Streams(T, m1), O.Streams(T, m2)
    a, b = O.flat(lx, ly, O.ENC_LOCAL), O.flat(lx, ly, O.ENC_LOCAL)
    O.cb_mcs(a, lx, ly, s1, 3, 0.8, 0.15)
    O.cb_mcs_threaded(b, lx, ly, s2, 3, 0.8, 0.15, nthreads)
    assert np.array_equal(a, b) and np.array_equal(s1.states, s2.states)
    assert not np.array_equal(a, O.flat(lx, ly, O.ENC_LOCAL))


def test_threaded_db_off_the_word_grid_equals_sequential():
    # lorest = 1: block origins on even sites, so neighbouring blocks of two
    # threads share lattice words; repeated to give a lost update a chance
    lx, ly, lbx, lby, T = 9, 9, 3, 3, 64
    for rep in range(20):
        m1, m2 = O.Dsfmt(rep), O.Dsfmt(rep)
        s1, s2 = O.Streams(T, m1), O.Streams(T, m2)
        a, b = O.flat(lx, ly, 0), O.flat(lx, ly, 0)
        O.db_mcs(a, lx, ly, lbx, lby, s1, m1, 2, 1, 1)
        O.db_mcs(b, lx, ly, lbx, lby, s2, m2, 2, 1, 8)
        assert np.array_equal(a, b) and np.array_equal(s1.states, s2.states)
