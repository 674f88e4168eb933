"""Bit-exact parity of the HIP engine (through the C ABI) with the oracle's
restatement of the CPU counterparts: same seed, same T worker streams, same
decomposition -> identical lattice words, stream states and master state."""
import numpy as np
import pytest

import cudakpz_amd as K
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def run_oracle_dt(lx, ly, lbx, lby, T, seed, n, ro, upd, p, q):
    m = O.Dsfmt(O.seed32(seed))
    st = O.Streams(T, m)
    lat = O.flat(lx, ly, O.ENC_LOCAL)
    O.dt_mcs(lat, lx, ly, lbx, lby, st, m, n, ro, upd, p, q)
    return lat, st.states, m


def run_gpu(cls, lx, ly, seed, n, pq=None, **kw):
    s = cls(lx, ly, **kw)
    s.supplySeed(seed)
    s.initHomogeneous()
    if pq is None:
        s.mcs(n)
    else:
        s.setPQ(*pq)
        s.mcsPQSyncRng(n)
    return s


def assert_same(s, lat, states, master):
    got = s.popSystem()
    if not np.array_equal(got, lat):
        bad = np.flatnonzero(got != lat)
        pytest.fail(f"{bad.size} lattice words differ, first at {bad[:8]}")
    assert np.array_equal(s.streams(), states)
    assert s.master_draw() == master.close_open()


DT_CASES = [
    # lx, ly, lbx, lby, T
    (6, 5, 2, 2, 1),
    (6, 5, 2, 2, 3),
    (6, 5, 2, 2, 0),
    (9, 9, 4, 4, 1),
    (9, 9, 4, 4, 7),
    (9, 9, 4, 4, 0),
    (10, 8, 3, 2, 5),
    (8, 10, 2, 4, 0),
    (7, 7, 3, 3, 64),
]


@pytest.mark.parametrize("lx,ly,lbx,lby,T", DT_CASES)
@pytest.mark.parametrize("ro", [False, True])
def test_dt_kpz_parity(lx, ly, lbx, lby, T, ro):
    cls = K.SchedulerDTr if ro else K.SchedulerDT
    s = run_gpu(cls, lx, ly, 42, 3, threads=T, layout=(lbx, lby))
    lat, st, m = run_oracle_dt(lx, ly, lbx, lby, s.threads, 42, 3, ro, O.UPD_KPZ, 1.0, 0.0)
    assert_same(s, lat, st, m)
    assert s.mcsCount() == 3


@pytest.mark.parametrize("lx,ly,lbx,lby,T", DT_CASES[::2])
@pytest.mark.parametrize("ro", [False, True])
@pytest.mark.parametrize("p,q", [(0.5, 0.5), (0.95, 0.05), (0.7, 0.0)])
def test_dt_pq_parity(lx, ly, lbx, lby, T, ro, p, q):
    cls = K.SchedulerDTr if ro else K.SchedulerDT
    s = run_gpu(cls, lx, ly, 7, 2, pq=(p, q), threads=T, layout=(lbx, lby))
    lat, st, m = run_oracle_dt(lx, ly, lbx, lby, s.threads, 7, 2, ro, O.UPD_PQ, p, q)
    assert_same(s, lat, st, m)


def run_oracle_cbbit(lx, ly, T, seed, n, mode, p, q):
    m = O.Dsfmt(O.seed32(seed))
    st = O.Streams(T, m)
    lat = O.flat(lx, ly, O.ENC_CBBIT)
    O.cbbit_mcs(lat, lx, ly, st, n, mode, p, q)
    return lat, st.states, m


CB_CASES = [
    # lx, ly, T
    (7, 4, 1),
    (8, 6, 3),
    (6, 5, 5),
    (10, 10, 0),
    (12, 6, 7),
    (14, 8, 0),
    (13, 9, 100),
]


@pytest.mark.parametrize("lx,ly,T", CB_CASES)
def test_scabit_bitparallel_parity(lx, ly, T):
    s = run_gpu(K.SchedulerSCABit, lx, ly, 42, 3, threads=T)
    assert s.mcsScale() == 0.5
    lat, st, m = run_oracle_cbbit(lx, ly, s.threads, 42, 3, 0, 0.5, 0.0)
    assert_same(s, lat, st, m)


@pytest.mark.parametrize("lx,ly,T", CB_CASES)
@pytest.mark.parametrize("p,q", [(0.95, 0.05), (0.5, 0.5), (1.0, 0.0), (0.3, 0.6)])
def test_scabit_pq_parity(lx, ly, T, p, q):
    s = run_gpu(K.SchedulerSCABit, lx, ly, 11, 2, pq=(p, q), threads=T)
    lat, st, m = run_oracle_cbbit(lx, ly, s.threads, 11, 2, 1, p, q)
    assert_same(s, lat, st, m)
    assert s.mcsScale() == p + q


def test_scabit_pq_half_zero_uses_bitparallel_path():
    # mcsPQSyncRng with (p, q) = (.5, 0) is the VL3 path (kpz/schedulerCPUCBBit.cpp:95-102)
    s = run_gpu(K.SchedulerSCABit, 9, 7, 3, 2, pq=(0.5, 0.0), threads=4)
    lat, st, m = run_oracle_cbbit(9, 7, 4, 3, 2, 0, 0.5, 0.0)
    assert_same(s, lat, st, m)


# (13, 6) and (15, 5): rows long enough that the correlator's lanes take
# several rounds of words each
@pytest.mark.parametrize("cls,enc,lx,ly", [(K.SchedulerDT, 0, 9, 8), (K.SchedulerSCABit, 1, 9, 8),
                                           (K.SchedulerDT, 0, 5, 4), (K.SchedulerSCABit, 1, 6, 3),
                                           (K.SchedulerDT, 0, 13, 6), (K.SchedulerSCABit, 1, 15, 5)])
def test_roughness_and_correlate_parity(cls, enc, lx, ly):
    kw = {"layout": (2, 2)} if cls is K.SchedulerDT else {}
    s = cls(lx, ly, **kw)
    s.supplySeed(5)
    s.initHomogeneous()
    N = 1 << (lx + ly)
    r0 = s.roughness()
    assert (r0.sum_h, r0.sum_h2) == (-N // 2, N // 2)
    if cls is K.SchedulerDT:
        s.mcs(2)
    else:
        s.setPQ(0.9, 0.1)
        s.mcsPQSyncRng(2)
    s.correlateTag()
    snap = s.popSystem()
    if cls is K.SchedulerDT:
        s.setPQ(0.8, 0.3)
        s.mcsPQSyncRng(3)
    else:
        s.mcsPQSyncRng(3)
    cur = s.popSystem()
    r = s.roughness()
    assert (r.sum_h, r.sum_h2) == O.roughness(cur, enc, lx, ly)
    c = s.correlate()
    assert (c.sum_h, c.sum_hs, c.sum_h2, c.sum_hhs, c.tcs) == O.correlate(cur, snap, enc, lx, ly)
    assert O.consistent(cur, enc, lx, ly)


@pytest.mark.parametrize("cls,enc,lx,ly", [(K.SchedulerDT, 0, 9, 8), (K.SchedulerSCABit, 1, 10, 7),
                                           (K.SchedulerDT, 0, 5, 4), (K.SchedulerSCABit, 1, 6, 3),
                                           (K.SchedulerSCA, 0, 8, 9)])
def test_height_moments(cls, enc, lx, ly):
    # <h^k>, k = 1..7, against exact integer power sums of the oracle's heights
    s = cls(lx, ly, layout=(2, 2)) if cls is K.SchedulerDT else cls(lx, ly)
    s.supplySeed(13)
    s.initHomogeneous()
    s.setPQ(0.9, 0.1)
    s.mcsPQSyncRng(25)
    got = s.moments()
    h = O.heights(s.popSystem(), enc, lx, ly).astype(np.int64).ravel()
    N = h.size
    for k in range(1, 8):
        exact = sum(int(v) ** k * int(c) for v, c in zip(*np.unique(h, return_counts=True)))
        want = exact / N
        assert abs(got[k - 1] - want) <= 1e-12 * max(1.0, abs(want)), (k, got[k - 1], want)
    r = s.roughness()
    assert got[0] * N == r.sum_h and got[1] * N == r.sum_h2


def test_upload_download_roundtrip():
    s = K.SchedulerSCABit(8, 6)
    lat = O.flat(8, 6, O.ENC_CBBIT)
    rng = np.random.default_rng(0)
    lat ^= rng.integers(0, 1 << 32, lat.size, dtype=np.uint64).astype(np.uint32)
    s.pushSystem(lat)
    assert np.array_equal(s.popSystem(), lat)


@pytest.mark.parametrize("cls,enc", [(K.SchedulerSCABit, 1), (K.SchedulerDT, 0), (K.SchedulerSCA, 0)])
def test_several_correlator_tags(cls, enc):
    lx, ly = 8, 8
    s = cls(lx, ly)
    s.supplySeed(77)
    s.initHomogeneous()
    s.setPQ(0.9, 0.1)
    snaps = []
    s.correlateTag(stop=5)            # expires before the evaluation below
    snaps.append((0, s.popSystem()))
    s.mcsPQSyncRng(2)
    s.correlateTag()
    snaps.append((2, s.popSystem()))
    s.mcsPQSyncRng(2)
    s.correlateTag(stop=100)
    snaps.append((4, s.popSystem()))
    s.mcsPQSyncRng(3)
    cur = s.popSystem()
    res = s.correlateAll()
    assert [m for m, _ in res] == [2, 4]
    for (m, c), (m2, snap) in zip(res, snaps[1:]):
        assert m == m2
        assert (c.sum_h, c.sum_hs, c.sum_h2, c.sum_hhs, c.tcs) == O.correlate(cur, snap, enc, lx, ly)
    latest = s.correlate()
    assert latest.sum_hhs == res[-1][1].sum_hhs


def test_master_handover_and_randomize():
    # the reference's dsfmt_t bytes go in, randomize() seeds the streams from them
    m = O.Dsfmt(1234)
    st = O.Streams(37, m.copy())
    s = K.SchedulerSCABit(8, 7, threads=37)
    s.set_master_state(bytes(m.buf))
    s.randomize()
    assert np.array_equal(s.streams(), st.states)
    m.close_open(), m.close_open()
    assert s.master_draw() == m.close_open()
