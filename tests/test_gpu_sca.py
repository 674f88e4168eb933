"""Checkerboard SCA on ENC_LOCAL (SchedulerSCA / SchedulerCPUCB): bit-exact
against the oracle for several worker counts, including uneven strips."""
import numpy as np
import pytest

import cudakpz_amd as K
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def oracle_cb(lx, ly, T, seed, n, p, q):
    m = O.Dsfmt(O.seed32(seed))
    st = O.Streams(T, m)
    lat = O.flat(lx, ly, O.ENC_LOCAL)
    O.cb_mcs(lat, lx, ly, st, n, p, q)
    return lat, st.states, m


@pytest.mark.parametrize("lx,ly,T", [(5, 4, 1), (5, 4, 3), (5, 4, 8), (8, 8, 0), (10, 7, 5), (9, 9, 7),
                                     (12, 6, 1), (11, 8, 100), (3, 3, 2), (3, 6, 1), (10, 10, 0),
                                     (12, 10, 37), (7, 11, 3)])
@pytest.mark.parametrize("p,q", [(0.95, 0.0), (0.5, 0.5), (0.8, 0.3)])
def test_sca_pq_parity(lx, ly, T, p, q):
    s = K.SchedulerSCA(lx, ly, threads=T)
    s.supplySeed(21)
    s.initHomogeneous()
    s.setPQ(p, q)
    s.mcsPQSyncRng(3)
    lat, st, m = oracle_cb(lx, ly, s.threads, 21, 3, p, q)
    got = s.popSystem()
    assert np.array_equal(got, lat), f"{np.count_nonzero(got != lat)} words differ"
    assert np.array_equal(s.streams(), st)
    assert s.master_draw() == m.close_open()
    assert O.consistent(got, O.ENC_LOCAL, lx, ly)


def test_sca_mcs_forces_p095_q0():
    # SchedulerCPUCB::mcs sets p=.95, q=0 (kpz/schedulerCPUCB.cpp:82-86)
    s = K.SchedulerSCA(7, 7, threads=4)
    s.supplySeed(3)
    s.initHomogeneous()
    s.setPQ(0.2, 0.7)
    s.mcs(2)
    lat, st, _ = oracle_cb(7, 7, 4, 3, 2, 0.95, 0.0)
    assert np.array_equal(s.popSystem(), lat)


def test_sca_roughness_matches_oracle():
    s = K.SchedulerSCA(9, 8)
    s.supplySeed(8)
    s.initHomogeneous()
    s.mcs(4)
    cur = s.popSystem()
    r = s.roughness()
    assert (r.sum_h, r.sum_h2) == O.roughness(cur, O.ENC_LOCAL, 9, 8)
