"""Disorder2 (per-site rate classes, mcsDisorder2SyncRng) and the EW fast path
of the HIP engine, bit-exact against the oracle: the engine's
generateDisorder2 map equals the oracle's for the same master, and the
lattice, streams and master after n MCS are identical."""
import numpy as np
import pytest

import cudakpz_amd as K
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RATES = {0: (0.9, 0.05), 1: (0.3, 0.6), 2: (0.7, 0.2), 3: (0.05, 0.95)}


def engine(cls, lx, ly, seed, classes, **kw):
    s = cls(lx, ly, **kw)
    s.supplySeed(seed)
    s.initHomogeneous()
    for n in classes:
        s.setPQ(*RATES[n], n)
    return s


def oracle_rates(classes, dis):
    p = [RATES[n][0] if n in classes else RATES[0][0] for n in range(4)]
    q = [RATES[n][1] if n in classes else RATES[0][1] for n in range(4)]
    return O.rates(p, q, dis)


def check(s, lat, st, m):
    got = s.popSystem()
    if not np.array_equal(got, lat):
        pytest.fail(f"{np.count_nonzero(got != lat)} words differ")
    assert np.array_equal(s.streams(), st.states)
    assert s.master_draw() == m.close_open()


DT_CASES = [(K.SchedulerDT, False, 9, 8, (4, 4), 0), (K.SchedulerDT, False, 7, 7, (2, 3), 5),
            (K.SchedulerDTr, True, 9, 8, (4, 4), 0), (K.SchedulerDTr, True, 8, 9, (3, 4), 7)]


@pytest.mark.parametrize("cls,ro,lx,ly,layout,T", DT_CASES)
def test_dt_disorder2_generated_map(cls, ro, lx, ly, layout, T):
    s = engine(cls, lx, ly, 31, (0, 1), threads=T, layout=layout)
    s.generateDisorder2(0.4)
    m = O.Dsfmt(O.seed32(31))
    st = O.Streams(s.threads, m)
    dis = O.generate_disorder2(m, s.size_words, 0.4)
    assert np.array_equal(s.disorder(), dis)
    s.mcsDisorder2SyncRng(3)
    lat = O.flat(lx, ly, O.ENC_LOCAL)
    O.dt_mcs_rates(lat, lx, ly, layout[0], layout[1], st, m, 3, ro, O.UPD_DIS, oracle_rates((0, 1), dis))
    check(s, lat, st, m)


@pytest.mark.parametrize("cls,ro,lx,ly,layout,T", DT_CASES[::2])
def test_dt_disorder2_four_classes(cls, ro, lx, ly, layout, T):
    s = engine(cls, lx, ly, 5, (0, 1, 2, 3), threads=T, layout=layout)
    dis = np.random.default_rng(4).integers(0, 1 << 32, s.size_words, dtype=np.uint64).astype(np.uint32)
    s.setDisorder(dis)
    s.mcsDisorder2SyncRng(2)
    m = O.Dsfmt(O.seed32(5))
    st = O.Streams(s.threads, m)
    lat = O.flat(lx, ly, O.ENC_LOCAL)
    O.dt_mcs_rates(lat, lx, ly, layout[0], layout[1], st, m, 2, ro, O.UPD_DIS, oracle_rates((0, 1, 2, 3), dis))
    check(s, lat, st, m)


@pytest.mark.parametrize("cls,ro,lx,ly,layout,T", DT_CASES)
def test_dt_ew_fast_path(cls, ro, lx, ly, layout, T):
    s = engine(cls, lx, ly, 8, (), threads=T, layout=layout)
    s.mcsEW(3)
    m = O.Dsfmt(O.seed32(8))
    st = O.Streams(s.threads, m)
    lat = O.flat(lx, ly, O.ENC_LOCAL)
    O.dt_mcs_rates(lat, lx, ly, layout[0], layout[1], st, m, 3, ro, O.UPD_EW, O.rates([0.0], [0.0]))
    check(s, lat, st, m)


@pytest.mark.parametrize("lx,ly,T", [(7, 5, 1), (8, 6, 3), (6, 5, 4), (10, 8, 0)])
def test_scabit_disorder2(lx, ly, T):
    s = engine(K.SchedulerSCABit, lx, ly, 12, (0, 1), threads=T)
    s.generateDisorder2(0.5)
    m = O.Dsfmt(O.seed32(12))
    st = O.Streams(s.threads, m)
    dis = O.generate_disorder2(m, s.size_words, 0.5)
    assert np.array_equal(s.disorder(), dis)
    s.mcsDisorder2SyncRng(2)
    lat = O.flat(lx, ly, O.ENC_CBBIT)
    O.cbbit_mcs_rates(lat, lx, ly, st, 2, 2, oracle_rates((0, 1), dis))
    check(s, lat, st, m)


@pytest.mark.parametrize("lx,ly,T", [(5, 4, 3), (8, 8, 0), (10, 9, 5), (12, 10, 37)])
def test_sca_disorder2(lx, ly, T):
    s = engine(K.SchedulerSCA, lx, ly, 14, (0, 1, 2, 3), threads=T)
    dis = np.random.default_rng(6).integers(0, 1 << 32, s.size_words, dtype=np.uint64).astype(np.uint32)
    s.setDisorder(dis)
    s.mcsDisorder2SyncRng(2)
    m = O.Dsfmt(O.seed32(14))
    st = O.Streams(s.threads, m)
    lat = O.flat(lx, ly, O.ENC_LOCAL)
    O.cb_mcs_rates(lat, lx, ly, st, 2, O.UPD_DIS, oracle_rates((0, 1, 2, 3), dis))
    check(s, lat, st, m)


def test_disorder_errors_and_state_roundtrip(tmp_path):
    s = K.SchedulerSCABit(7, 6)
    with pytest.raises(K.KpzError):
        s.mcsDisorder2SyncRng(1)      # no map yet
    with pytest.raises(K.KpzError):
        s.mcsEW(1)                    # DT schedulers only
    with pytest.raises(K.KpzError):
        s.setPQ(0.5, 0.5, 4)          # classes 0..3
    a = engine(K.SchedulerDTr, 8, 8, 3, (0, 1))
    b = engine(K.SchedulerDTr, 8, 8, 3, (0, 1))
    for x in (a, b):
        x.generateDisorder2(0.3)
        x.mcsDisorder2SyncRng(2)
    path = tmp_path / "dis.kpzstate"
    b.saveState(str(path))
    b.close()
    c, _ = K.load_state(str(path))
    assert np.array_equal(c.disorder(), a.disorder())
    for x in (a, c):
        x.mcsDisorder2SyncRng(3)
    assert np.array_equal(c.popSystem(), a.popSystem())
    assert np.array_equal(c.streams(), a.streams())
