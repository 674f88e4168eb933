"""TinyMT32 worker streams (kpz_params.rng = KPZ_RNG_TINYMT32): the schedulers
the reference's frontend builds with TinyMT32 workers (SchedulerCPUDTr, CB,
CBBit; kpz/kpzFrontend.cpp:308-331) on the GPU, bit-exact against the
oracle: same master seed, same per-stream parameter sets, same T ->
identical lattice words, TinyMT32 states and master state.  The parameter
sets are random words: the recurrence and the parity do not depend on them
(the reference's own table, tinymt32_param.h, is not in its tree)."""
import numpy as np
import pytest

import cudakpz_amd as K
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RATES = {0: (0.9, 0.05), 1: (0.3, 0.6), 2: (0.7, 0.2), 3: (0.05, 0.95)}


def mats(T, seed):
    r = np.random.default_rng(seed)
    return [r.integers(0, 1 << 32, T, dtype=np.uint64).astype(np.uint32) for _ in range(3)]


def engine(cls, lx, ly, seed, T=0, layout=(4, 4), pseed=1):
    s = cls(lx, ly, threads=T, layout=layout, rng=K.RNG_TINYMT32)
    assert s.rng == K.RNG_TINYMT32
    m = mats(s.threads, pseed)
    s.setTinyMTParams(*m)
    s.supplySeed(seed)
    s.initHomogeneous()
    return s, m


def oracle(s, seed, m):
    master = O.Dsfmt(O.seed32(seed))
    return master, O.TmtStreams(s.threads, master, *m)


def check(s, lat, st, master):
    got = s.popSystem()
    if not np.array_equal(got, lat):
        bad = np.flatnonzero(got != lat)
        pytest.fail(f"{bad.size} lattice words differ, first at {bad[:8]}")
    assert np.array_equal(s.tinymt(), st.g.T)
    assert s.master_draw() == master.close_open()


def rates(classes, dis):
    p = [RATES[n][0] if n in classes else RATES[0][0] for n in range(4)]
    q = [RATES[n][1] if n in classes else RATES[0][1] for n in range(4)]
    return O.rates(p, q, dis)


def test_seeding_matches_randomize():
    s, m = engine(K.SchedulerSCABit, 8, 6, 42, T=5)
    master, st = oracle(s, 42, m)
    assert np.array_equal(s.tinymt(), st.g.T)
    assert s.master_draw() == master.close_open()


# ---- SchedulerCPUDTr<TinyMT32>
DT_CASES = [
    # lx, ly, lbx, lby, T
    (6, 5, 2, 2, 1),
    (6, 5, 2, 2, 3),
    (9, 9, 4, 4, 0),
    (9, 9, 4, 4, 7),
    (10, 8, 3, 2, 5),
    (7, 7, 3, 3, 64),
    (8, 8, 4, 4, 20),  # more workers than blocks: the surplus keep their streams
]


@pytest.mark.parametrize("lx,ly,lbx,lby,T", DT_CASES)
def test_dtr_kpz(lx, ly, lbx, lby, T):
    s, m = engine(K.SchedulerDTr, lx, ly, 42, T, (lbx, lby))
    s.mcs(3)
    master, st = oracle(s, 42, m)
    lat = O.flat(lx, ly, O.ENC_LOCAL)
    O.dtr_mcs_tmt(lat, lx, ly, lbx, lby, st, master, 3)
    check(s, lat, st, master)


@pytest.mark.parametrize("lx,ly,lbx,lby,T", DT_CASES[::2])
@pytest.mark.parametrize("p,q", [(0.5, 0.5), (0.95, 0.05), (0.7, 0.0)])
def test_dtr_pq(lx, ly, lbx, lby, T, p, q):
    s, m = engine(K.SchedulerDTr, lx, ly, 7, T, (lbx, lby))
    s.setPQ(p, q)
    s.mcsPQSyncRng(2)
    master, st = oracle(s, 7, m)
    lat = O.flat(lx, ly, O.ENC_LOCAL)
    O.dtr_mcs_tmt(lat, lx, ly, lbx, lby, st, master, 2, O.UPD_PQ, p, q)
    check(s, lat, st, master)


@pytest.mark.parametrize("lx,ly,lbx,lby,T", [(9, 8, 4, 4, 0), (8, 9, 3, 4, 7)])
def test_dtr_disorder2_and_ew(lx, ly, lbx, lby, T):
    s, m = engine(K.SchedulerDTr, lx, ly, 31, T, (lbx, lby))
    for n in (0, 1, 2):
        s.setPQ(*RATES[n], n)
    s.generateDisorder2(0.4)
    s.mcsDisorder2SyncRng(2)
    s.mcsEW(1)
    master, st = oracle(s, 31, m)
    dis = O.generate_disorder2(master, s.size_words, 0.4)
    assert np.array_equal(s.disorder(), dis)
    lat = O.flat(lx, ly, O.ENC_LOCAL)
    O.dtr_mcs_tmt(lat, lx, ly, lbx, lby, st, master, 2, O.UPD_DIS, r=rates((0, 1, 2), dis))
    O.dtr_mcs_tmt(lat, lx, ly, lbx, lby, st, master, 1, O.UPD_EW)
    check(s, lat, st, master)


# ---- SchedulerCPUCBBit<TinyMT32>
CB_CASES = [
    # lx, ly, T
    (7, 4, 1),
    (8, 6, 3),
    (6, 5, 5),     # L = 64: one draw per u32 word in the bit-parallel mode
    (10, 10, 0),
    (12, 6, 7),
]


@pytest.mark.parametrize("lx,ly,T", CB_CASES + [(11, 8, 0), (12, 9, 5)])  # >= 256 rows: staged stores
def test_scabit_bitparallel(lx, ly, T):
    s, m = engine(K.SchedulerSCABit, lx, ly, 42, T)
    s.mcs(3)
    master, st = oracle(s, 42, m)
    lat = O.flat(lx, ly, O.ENC_CBBIT)
    O.cbbit_mcs_tmt(lat, lx, ly, st, 3, 0)
    check(s, lat, st, master)


@pytest.mark.parametrize("lx,ly,T", CB_CASES + [(14, 4, 3), (15, 3, 0)])  # long rows: split into chunks
@pytest.mark.parametrize("p,q", [(0.95, 0.05), (0.3, 0.6)])
def test_scabit_pq(lx, ly, T, p, q):
    s, m = engine(K.SchedulerSCABit, lx, ly, 11, T)
    s.setPQ(p, q)
    s.mcsPQSyncRng(2)
    master, st = oracle(s, 11, m)
    lat = O.flat(lx, ly, O.ENC_CBBIT)
    O.cbbit_mcs_tmt(lat, lx, ly, st, 2, 1, p, q)
    check(s, lat, st, master)


@pytest.mark.parametrize("lx,ly,T", [(8, 6, 3), (7, 7, 0), (14, 3, 5), (10, 9, 3)])
def test_scabit_disorder2(lx, ly, T):
    s, m = engine(K.SchedulerSCABit, lx, ly, 5, T)
    for n in (0, 1):
        s.setPQ(*RATES[n], n)
    s.generateDisorder2(0.5)
    s.mcsDisorder2SyncRng(2)
    master, st = oracle(s, 5, m)
    dis = O.generate_disorder2(master, s.size_words, 0.5)
    lat = O.flat(lx, ly, O.ENC_CBBIT)
    O.cbbit_mcs_tmt(lat, lx, ly, st, 2, 2, r=rates((0, 1), dis))
    check(s, lat, st, master)


# ---- SchedulerCPUCB<TinyMT32>
SCA_CASES = [
    # lx, ly, T
    (5, 4, 1),
    (5, 4, 3),
    (8, 6, 8),
    (8, 6, 0),     # one cell column per worker: every word is shared, all atomics
    (9, 7, 13),
    (6, 9, 2),
    # at least 16 cells per half strip, so the generate kernel takes its
    # 16-cell path (four aligned 16-byte stores where a walk runs along a half
    # row): strips of 64 cells starting on the 16-cell grid, and strips of
    # 86/85 cells whose halves start off it (scalar and vector stores mix)
    (9, 5, 2),
    (10, 4, 3),
]


@pytest.mark.parametrize("lx,ly,T", SCA_CASES)
def test_sca_mcs(lx, ly, T):
    s, m = engine(K.SchedulerSCA, lx, ly, 3, T)
    s.mcs(2)
    master, st = oracle(s, 3, m)
    lat = O.flat(lx, ly, O.ENC_LOCAL)
    O.cb_mcs_tmt(lat, lx, ly, st, 2, O.UPD_PQ, 0.95, 0.0)
    check(s, lat, st, master)


@pytest.mark.parametrize("lx,ly,T", SCA_CASES[1::2])
def test_sca_pq_and_disorder2(lx, ly, T):
    s, m = engine(K.SchedulerSCA, lx, ly, 9, T)
    s.setPQ(0.6, 0.35)
    s.mcsPQSyncRng(1)
    for n in (1, 3):
        s.setPQ(*RATES[n], n)
    s.generateDisorder2(0.3)
    s.mcsDisorder2SyncRng(2)
    master, st = oracle(s, 9, m)
    lat = O.flat(lx, ly, O.ENC_LOCAL)
    O.cb_mcs_tmt(lat, lx, ly, st, 1, O.UPD_PQ, 0.6, 0.35)
    dis = O.generate_disorder2(master, s.size_words, 0.3)
    r = O.rates([0.6, RATES[1][0], 0.6, RATES[3][0]], [0.35, RATES[1][1], 0.35, RATES[3][1]], dis)
    O.cb_mcs_tmt(lat, lx, ly, st, 2, O.UPD_DIS, r=r)
    check(s, lat, st, master)


# ---- boundary behaviour
def test_errors():
    with pytest.raises(K.KpzError):
        K.SchedulerDT(8, 8, rng=K.RNG_TINYMT32)  # SchedulerCPUDT draws close_open()
    with pytest.raises(K.KpzError):
        K.SchedulerDTr(8, 8, rng=7)
    s = K.SchedulerSCABit(8, 6, threads=4, rng=K.RNG_TINYMT32)
    with pytest.raises(K.KpzError) as e:
        s.supplySeed(1)  # no parameter sets yet
    assert e.value.code == -4
    with pytest.raises(K.KpzError):
        s.streams()
    s.close()
    t = K.SchedulerSCABit(8, 6, threads=4)
    assert t.rng == K.RNG_SLCG64
    with pytest.raises(K.KpzError):
        t.setTinyMTParams(*mats(4, 0))
    with pytest.raises(K.KpzError):
        t.tinymt()


def test_state_file_resumes_bit_exact(tmp_path):
    a, m = engine(K.SchedulerSCA, 8, 6, 21, 5)
    a.mcs(1)
    path = str(tmp_path / "tmt.kst")
    a.saveState(path, aux=9)
    a.mcs(2)
    b, aux = K.load_state(path)
    assert aux == 9 and b.rng == K.RNG_TINYMT32
    b.mcs(2)
    assert np.array_equal(a.popSystem(), b.popSystem())
    assert np.array_equal(a.tinymt(), b.tinymt())
    assert a.master_draw() == b.master_draw()


@pytest.mark.parametrize("T", [32, 96])
def test_dtr_chains_of_blocks_against_oracle(T):
    # 16384 x 8192 in 4x4 blocks: 2^21 blocks per domain set, so a lane runs 4
    # consecutive blocks of one worker from a single jump; with 96 workers the
    # blocks do not divide evenly and the last chain of a worker is short
    s, m = engine(K.SchedulerDTr, 14, 13, 91, T, layout=(2, 2))
    s.mcs(1)
    master, st = oracle(s, 91, m)
    lat = O.flat(14, 13, O.ENC_LOCAL)
    O.dtr_mcs_tmt(lat, 14, 13, 2, 2, st, master, 1)
    check(s, lat, st, master)


@pytest.mark.parametrize("T", [1024, 300])
def test_medium_dtr_against_oracle(T):
    # 1024^2, 16x16 blocks, 1024 blocks per domain set: one block per worker,
    # or up to 4 blocks per worker (streams jumped to each block)
    s, m = engine(K.SchedulerDTr, 10, 10, 77, T)
    s.mcs(2)
    master, st = oracle(s, 77, m)
    lat = O.flat(10, 10, O.ENC_LOCAL)
    O.dtr_mcs_tmt(lat, 10, 10, 4, 4, st, master, 2)
    check(s, lat, st, master)
    assert O.consistent(lat, O.ENC_LOCAL, 10, 10)
