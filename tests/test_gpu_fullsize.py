"""Bit-exact parity at the benchmark size itself, 65536 x 65536 sites (1 GiB
of lattice, 2^32 sites), against the oracle: one SchedulerDT MCS with one
SLCG64 stream per 16 x 16 block (2^22 streams, the configuration the
site-update rate is measured on), one SchedulerDTr MCS, and two SchedulerSCABit
VL3 MCS and one per-site p/q MCS with one stream per row (65536 streams), and
one SchedulerSCA MCS with one stream per cell column (16384 streams).  The
oracle runs them on several threads (DT: a domain set's blocks are disjoint;
checkerboard: a half-sweep's updates own disjoint bits; threaded == sequential
is pinned by tests/test_oracle_model.py).  Beyond word equality,
the device consistency check and the exact roughness sums must agree with the
oracle's."""
import os

import numpy as np
import pytest

import cudakpz_amd as K
from oracle import oracle as O

pytestmark = pytest.mark.gpu

LX = LY = 16
THREADS = max(1, min(16, os.cpu_count() or 1))


def compare(s, lat, states, master, enc):
    got = s.popSystem()
    if not np.array_equal(got, lat):
        bad = np.flatnonzero(got != lat)
        pytest.fail(f"{bad.size} lattice words differ, first at {bad[:8]}")
    del got
    assert np.array_equal(s.streams(), states)
    assert s.master_draw() == master.close_open()
    ok, counts = s.consistencyCheck()
    assert ok and counts == (0, 0, 0)
    r = s.roughness()
    sum_h, sum_h2 = O.roughness(lat, enc, LX, LY)
    assert (r.sum_h, r.sum_h2) == (sum_h, sum_h2)


@pytest.mark.parametrize("cls,ro", [(K.SchedulerDT, False), (K.SchedulerDTr, True)])
def test_dt_one_mcs(cls, ro):
    s = cls(LX, LY, layout=(4, 4))
    assert s.threads == 1 << 22  # one stream per block of a domain set (a quarter of the 2^24 blocks)
    s.supplySeed(42)
    s.initHomogeneous()
    s.mcs(1)
    m = O.Dsfmt(O.seed32(42))
    st = O.Streams(s.threads, m)
    lat = O.flat(LX, LY, O.ENC_LOCAL)
    O.dt_mcs_threaded(lat, LX, LY, 4, 4, st, m, 1, THREADS, ro)
    compare(s, lat, st.states, m, O.ENC_LOCAL)
    s.close()


def test_scabit_vl3_two_mcs():
    s = K.SchedulerSCABit(LX, LY)
    assert s.threads == 1 << LY  # one stream per row
    s.supplySeed(7)
    s.initHomogeneous()
    s.mcs(2)
    m = O.Dsfmt(O.seed32(7))
    st = O.Streams(s.threads, m)
    lat = O.flat(LX, LY, O.ENC_CBBIT)
    O.cbbit_mcs(lat, LX, LY, st, 2, 0)
    compare(s, lat, st.states, m, O.ENC_CBBIT)
    s.close()


def test_scabit_pq_one_mcs():
    s = K.SchedulerSCABit(LX, LY)
    s.supplySeed(11)
    s.initHomogeneous()
    s.setPQ(0.95, 0.05)
    s.mcsPQSyncRng(1)
    m = O.Dsfmt(O.seed32(11))
    st = O.Streams(s.threads, m)
    lat = O.flat(LX, LY, O.ENC_CBBIT)
    O.cbbit_mcs_threaded(lat, LX, LY, st, 1, 1, 0.95, 0.05, THREADS)
    compare(s, lat, st.states, m, O.ENC_CBBIT)
    s.close()


def test_sca_one_mcs():
    s = K.SchedulerSCA(LX, LY)
    assert s.threads == 1 << (LX - 2)  # one stream per 4-site cell column
    s.supplySeed(3)
    s.initHomogeneous()
    s.mcs(1)  # p = .95, q = 0
    m = O.Dsfmt(O.seed32(3))
    st = O.Streams(s.threads, m)
    lat = O.flat(LX, LY, O.ENC_LOCAL)
    O.cb_mcs_threaded(lat, LX, LY, st, 1, 0.95, 0.0, THREADS)
    compare(s, lat, st.states, m, O.ENC_LOCAL)
    s.close()
