"""Disorder2 (two or more rate classes per site, kpz/schedulerService.cpp:54-69,
kpz/schedulerCPUBaseThreadedRng.inl:50-75, kpz/schedulerCPUCBBit_nested.inl:298-331)
and the EW fast path (p = q = 1 without an acceptance draw,
kpz/kpzSimulationPrimitivesTemplates.cu:199-234) in the oracle, against the
explicit-height model with per-site rates and independently written drivers."""
import numpy as np
import pytest

from heightmodel import HeightLattice, normalized, thresh23
from oracle import oracle as O

P4, Q4 = [0.9, 0.3, 0.6, 0.1], [0.05, 0.7, 0.2, 0.5]


def evolved(lx, ly, seed, enc):
    m = O.Dsfmt(seed)
    st = O.Streams(3, m)
    lat = O.flat(lx, ly, O.ENC_LOCAL)
    O.cb_mcs(lat, lx, ly, st, 3, 0.7, 0.3)
    return lat if enc == O.ENC_LOCAL else O.local_to_cbbit(lat, lx, ly)


def class_local(dis, lx, x, y):
    """the 2 bits of site (x, y) in an ENC_LOCAL-laid-out disorder map"""
    w = int(dis[((y >> 2) << (lx - 2)) + (x >> 2)])
    return (w >> ((x & 3) * 2 + (y & 3) * 8)) & 3


def class_cbbit(dis, lx, ly, x, y):
    """the x-slope bit of site (x, y) with the ENC_CBBIT quadrant layout"""
    q = 1 << (lx + ly - 6)
    sub = (x ^ y) & 1
    w = int(dis[2 * sub * q + (y << (lx - 6)) + (x >> 6)])
    return (w >> ((x >> 1) & 31)) & 1


def test_generate_disorder2_restated():
    m1, m2 = O.Dsfmt(17), O.Dsfmt(17)
    dis = O.generate_disorder2(m1, 300, 0.3)
    want = []
    for _ in range(300):
        t = 0
        for s in range(0, 32, 2):
            if m2.close_open() < 0.3:
                t |= 1 << s
        want.append(t)
    assert dis.tolist() == want
    assert m1.close_open() == m2.close_open()
    assert not np.any(dis & 0xAAAAAAAA)
    assert 0.25 < np.mean([bin(int(w)).count("1") for w in dis]) / 16 < 0.35


def test_local_disorder_rule_vs_height_model():
    lx, ly = 6, 6
    lat = evolved(lx, ly, 3, O.ENC_LOCAL)
    dis = np.random.default_rng(1).integers(0, 1 << 32, lat.size, dtype=np.uint64).astype(np.uint32)
    r = O.rates(P4, Q4, dis)
    hm = HeightLattice(O.heights(lat, 0, lx, ly))
    rng = np.random.default_rng(7)
    seen = set()
    for _ in range(4000):
        x, y = int(rng.integers(64)), int(rng.integers(64))
        v = int(rng.integers(1 << 23))
        c = class_local(dis, lx, x, y)
        seen.add(c)
        O.update_local_rates(lat, lx, ly, x, y, O.UPD_DIS, v / 8388608.0, r)
        hm.update(x, y, v < thresh23(P4[c]), v < thresh23(Q4[c]))
    assert seen == {0, 1, 2, 3}
    assert np.array_equal(normalized(O.heights(lat, 0, lx, ly)), hm.normalized())


def test_ew_rule_flips_every_extremum():
    lx, ly = 6, 6
    lat = evolved(lx, ly, 4, O.ENC_LOCAL)
    hm = HeightLattice(O.heights(lat, 0, lx, ly))
    r = O.rates([0.0], [0.0])  # the EW rule ignores the rates
    rng = np.random.default_rng(9)
    for _ in range(3000):
        x, y = int(rng.integers(64)), int(rng.integers(64))
        O.update_local_rates(lat, lx, ly, x, y, O.UPD_EW, 0.99, r)
        hm.update(x, y, True, True)
    assert np.array_equal(normalized(O.heights(lat, 0, lx, ly)), hm.normalized())


def test_cbbit_disorder_rule_vs_height_model():
    lx, ly = 7, 6
    lat = evolved(lx, ly, 5, O.ENC_CBBIT)
    dis = np.random.default_rng(2).integers(0, 1 << 32, lat.size, dtype=np.uint64).astype(np.uint32)
    r = O.rates(P4[:2], Q4[:2], dis)
    hm = HeightLattice(O.heights(lat, 1, lx, ly))
    rng = np.random.default_rng(8)
    for _ in range(3000):
        x, y = int(rng.integers(128)), int(rng.integers(64))
        v = int(rng.integers(1 << 23))
        c = class_cbbit(dis, lx, ly, x, y)
        O.update_cbbit_rates(lat, lx, ly, x, y, v / 8388608.0, r)
        hm.update(x, y, v < thresh23(P4[c]), v < thresh23(Q4[c]))
    assert np.array_equal(normalized(O.heights(lat, 1, lx, ly)), hm.normalized())


# ---------------------------------------------------------------- compositions
def py_dt_rule(hm, lx, ly, lbx, lby, st, master, n, random_origin, rule, draw_accept):
    """DT walk (kpz/schedulerCPUDT.cpp:65-90, kpz/schedulerCPUDTr.cpp:66-98)
    with an arbitrary acceptance rule(x, y, r23) -> (accept_min, accept_max)"""
    X, Y = 1 << lx, 1 << ly
    nb = 1 << (lx - lbx + ly - lby - 2)
    hx = 1 << (lx - lbx - 1)
    for _ in range(n):
        seq = master.shuffle4()
        sx = sy = 0
        if random_origin:
            sx = int(master.close_open() * (1 << lbx))
            sy = int(master.close_open() * (1 << lby))
        for ds in seq:
            for t in range(st.T):
                for i in range(t, nb, st.T):
                    ox = (2 * (i % hx) + (ds & 1)) << lbx
                    oy = (2 * (i // hx) + (ds >> 1)) << lby
                    for _k in range(1 << (lbx + lby)):
                        if random_origin:
                            x = (((st.draw(t) & 0x7FFFFF) >> (23 - lbx)) + ox + sx) % X
                            y = (((st.draw(t) & 0x7FFFFF) >> (23 - lby)) + oy + sy) % Y
                        else:
                            x = (st.draw(t) >> (32 - lbx)) + ox
                            y = (st.draw(t) >> (32 - lby)) + oy
                        r = (st.draw(t) & 0x7FFFFF) if draw_accept else 0
                        hm.update(x, y, *rule(x, y, r))


@pytest.mark.parametrize("random_origin", [False, True])
@pytest.mark.parametrize("T", [1, 3])
def test_dt_disorder_and_ew_compositions(random_origin, T):
    lx, ly, lbx, lby = 6, 5, 2, 2
    m0 = O.Dsfmt(11)
    dis = O.generate_disorder2(m0, 1 << (lx + ly - 4), 0.5)
    r = O.rates(P4, Q4, dis)
    cases = [(O.UPD_DIS, lambda x, y, v: (v < thresh23(P4[class_local(dis, lx, x, y)]),
                                           v < thresh23(Q4[class_local(dis, lx, x, y)])), True),
             (O.UPD_EW, lambda x, y, v: (True, True), False)]
    for upd, rule, draw in cases:
        m1, m2 = O.Dsfmt(42), O.Dsfmt(42)
        s1, s2 = O.Streams(T, m1), O.Streams(T, m2)
        lat = O.flat(lx, ly, 0)
        hm = HeightLattice(O.heights(lat, 0, lx, ly))
        O.dt_mcs_rates(lat, lx, ly, lbx, lby, s1, m1, 2, random_origin, upd, r)
        py_dt_rule(hm, lx, ly, lbx, lby, s2, m2, 2, random_origin, rule, draw)
        assert np.array_equal(normalized(O.heights(lat, 0, lx, ly)), hm.normalized())
        assert np.array_equal(s1.states, s2.states)
        assert m1.close_open() == m2.close_open()
        assert O.consistent(lat, 0, lx, ly)


def test_uniform_rates_equal_pq_in_every_scheduler():
    # disorder with equal rates in every class is exactly the pq dynamics
    lx, ly = 7, 6
    dis = np.random.default_rng(3).integers(0, 1 << 32, 1 << (lx + ly - 4), dtype=np.uint64).astype(np.uint32)
    r = O.rates([0.8] * 4, [0.3] * 4, dis)
    runs = [
        (O.ENC_LOCAL, lambda lat, st, m: O.dt_mcs_rates(lat, lx, ly, 3, 2, st, m, 2, True, O.UPD_DIS, r),
         lambda lat, st, m: O.dt_mcs(lat, lx, ly, 3, 2, st, m, 2, True, O.UPD_PQ, 0.8, 0.3)),
        (O.ENC_LOCAL, lambda lat, st, m: O.cb_mcs_rates(lat, lx, ly, st, 2, O.UPD_DIS, r),
         lambda lat, st, m: O.cb_mcs(lat, lx, ly, st, 2, 0.8, 0.3)),
        (O.ENC_CBBIT, lambda lat, st, m: O.cbbit_mcs_rates(lat, lx, ly, st, 2, 2, r),
         lambda lat, st, m: O.cbbit_mcs(lat, lx, ly, st, 2, 1, 0.8, 0.3)),
    ]
    for enc, a, b in runs:
        out = []
        for f in (a, b):
            m = O.Dsfmt(5)
            st = O.Streams(4, m)
            lat = O.flat(lx, ly, enc)
            f(lat, st, m)
            out.append((lat, st.states.copy()))
        assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("T", [1, 3])
def test_cbbit_and_cb_disorder_compositions(T):
    # per-site classes against drivers of the checkerboard walks with the
    # class read independently from the map
    lx, ly = 7, 4
    dis = O.generate_disorder2(O.Dsfmt(13), 1 << (lx + ly - 4), 0.5)
    r = O.rates(P4[:2], Q4[:2], dis)
    X, Y = 1 << lx, 1 << ly

    # SchedulerCPUCBBit rows per worker, sites of the row's sublattice ascending
    m1, m2 = O.Dsfmt(7), O.Dsfmt(7)
    s1, s2 = O.Streams(T, m1), O.Streams(T, m2)
    lat = O.flat(lx, ly, O.ENC_CBBIT)
    hm = HeightLattice(O.heights(lat, 1, lx, ly))
    O.cbbit_mcs_rates(lat, lx, ly, s1, 2, 2, r)
    per, rem = Y // T, Y % T
    for _ in range(2):
        for s in range(2):
            lo = 0
            for t in range(T):
                hi = lo + per + (1 if t < rem else 0)
                for y in range(lo, hi):
                    for x in range(s ^ (y & 1), X, 2):
                        v = s2.draw(t) & 0x7FFFFF
                        c = class_cbbit(dis, lx, ly, x, y)
                        hm.update(x, y, v < thresh23(P4[c]), v < thresh23(Q4[c]))
                lo = hi
    assert np.array_equal(normalized(O.heights(lat, 1, lx, ly)), hm.normalized())
    assert np.array_equal(s1.states, s2.states)

    # SchedulerCPUCB strips split at their centre, 4x4 cells, 2 sites per cell row
    m1, m2 = O.Dsfmt(9), O.Dsfmt(9)
    s1, s2 = O.Streams(T, m1), O.Streams(T, m2)
    lat = O.flat(lx, ly, O.ENC_LOCAL)
    hm = HeightLattice(O.heights(lat, 0, lx, ly))
    O.cb_mcs_rates(lat, lx, ly, s1, 2, O.UPD_DIS, r)
    dimxw = X // 4
    per, rem = (dimxw // T) * 4, dimxw % T

    def loop(t, s, x0, x1):
        for yw in range(0, Y, 4):
            for xw in range(x0, x1, 4):
                for y in range(yw, yw + 4):
                    x = xw + ((y & 1) ^ s)
                    for xx in (x, x + 2):
                        v = s2.draw(t) & 0x7FFFFF
                        c = class_local(dis, lx, xx, y)
                        hm.update(xx, y, v < thresh23(P4[c]), v < thresh23(Q4[c]))

    for _ in range(2):
        for s in range(2):
            x1 = 0
            for t in range(T):
                x0 = x1
                x1 = x0 + per + (4 if t < rem else 0)
                xc = x0 + (((x1 - x0) // 2) & ~3)
                loop(t, s, x0, xc)
                loop(t, s, xc, x1)
    assert np.array_equal(normalized(O.heights(lat, 0, lx, ly)), hm.normalized())
    assert np.array_equal(s1.states, s2.states)


# ---------------------------------------------------------------- plain RS
def py_rs(hm, lx, ly, master, n, mode, p, q, dis=None):
    """SchedulerCPU (kpz/schedulerCPU.cpp:35-179): x, y from the master per
    update; the pq variants draw their acceptance number at a minimum or
    maximum only (mcsPQ, mcsDisorder2) or always (the SyncRng variants)"""
    X, Y = 1 << lx, 1 << ly
    sync = mode in (O.RS_SYNC, O.RS_PQ_SYNC, O.RS_DIS_SYNC)
    for _ in range(n):
        for _a in range(X * Y):
            x = int(master.close_open() * X)
            y = int(master.close_open() * Y)
            z = master.close_open() if sync else None
            k = hm.kind(x, y)
            if mode in (O.RS_MCS, O.RS_SYNC):
                hm.update(x, y, True, False)
                continue
            c = class_local(dis, lx, x, y) if dis is not None else 0
            if k == 0:
                continue
            if mode == O.RS_PQ and p[0] == 1.0 and q[0] == 1.0:
                hm.update(x, y, True, True)
                continue
            v = z if sync else master.close_open()
            hm.update(x, y, k == 1 and v < p[c], k == -1 and v < q[c])


@pytest.mark.parametrize("mode,p,q", [(O.RS_MCS, [1.0], [0.0]), (O.RS_PQ, [0.7], [0.4]), (O.RS_PQ, [1.0], [1.0]),
                                      (O.RS_SYNC, [1.0], [0.0]), (O.RS_PQ_SYNC, [0.6], [0.3]),
                                      (O.RS_DIS, P4, Q4), (O.RS_DIS_SYNC, P4, Q4)])
def test_plain_rs_scheduler(mode, p, q):
    lx, ly = 5, 4
    dis = None
    if mode in (O.RS_DIS, O.RS_DIS_SYNC):
        dis = np.random.default_rng(5).integers(0, 1 << 32, 1 << (lx + ly - 4), dtype=np.uint64).astype(np.uint32)
    r = O.rates(p, q, dis)
    m1, m2 = O.Dsfmt(23), O.Dsfmt(23)
    lat = O.flat(lx, ly, O.ENC_LOCAL)
    hm = HeightLattice(O.heights(lat, 0, lx, ly))
    O.rs_mcs(lat, lx, ly, m1, 3, mode, r)
    py_rs(hm, lx, ly, m2, 3, mode, p, q, dis)
    assert np.array_equal(normalized(O.heights(lat, 0, lx, ly)), hm.normalized())
    assert m1.close_open() == m2.close_open()
    assert O.consistent(lat, 0, lx, ly)
