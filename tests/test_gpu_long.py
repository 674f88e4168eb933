"""Long-horizon bit-exact parity with the threaded oracle, where the
benchmark lattices actually run: rough lattices, streams and master draws
carried over many MCS.

* 4096^2, 100 MCS (120 for the static DT with fewer streams than blocks):
  SchedulerDT, SchedulerDTr, SchedulerSCABit VL3 (p = .5), SchedulerSCABit
  per-site p = .95 / q = .05 with two correlator tags, and SchedulerSCA
  (p = .95); lattice, every stream state and the next master draw compared at
  the end, roughness and the correlators compared exactly.
* 65536^2 (1 GiB of lattice, 2^32 sites), 10 MCS of SchedulerDTr (2^22
  streams) and of SchedulerSCABit p = .95 / q = .05 (65536 streams, the
  autocorrelator configuration), with correlator tags taken at MCS 2 and 5
  (correlateTag, kpz/schedulerService.cpp:266-271) and the exact
  autocorrelator sums (kpz/correlator.cpp:127-164, 450-455) of both tags
  compared at MCS 10.  Each stage is its own test so that progress shows.

The oracle's reductions here are its row-streamed threaded forms (no
lattice-sized height arrays), which tests/test_oracle_model.py pins to the
explicit-height sums."""
import os

import numpy as np
import pytest

import cudakpz_amd as K
from oracle import oracle as O

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def corr_tuple(c):
    return (c.sum_h, c.sum_hs, c.sum_h2, c.sum_hhs, c.tcs)


class Run:
    """One scheduler on the device and its oracle twin, advanced together."""

    def __init__(self, kind, L, seed, threads=0, pq=None):
        self.kind, self.L, self.pq = kind, L, pq
        cls = {"dt": K.SchedulerDT, "dtr": K.SchedulerDTr, "vl3": K.SchedulerSCABit,
               "pq": K.SchedulerSCABit, "sca": K.SchedulerSCA}[kind]
        kw = {"layout": (4, 4)} if kind in ("dt", "dtr") else {}
        self.s = cls(L, L, threads=threads, **kw)
        self.s.supplySeed(seed)
        self.s.initHomogeneous()
        if pq is not None:
            self.s.setPQ(*pq)
        self.enc = O.ENC_CBBIT if kind in ("vl3", "pq") else O.ENC_LOCAL
        self.m = O.Dsfmt(O.seed32(seed))
        self.st = O.Streams(self.s.threads, self.m)
        self.lat = O.flat(L, L, self.enc)
        self.t = 0
        self.snaps = []

    def advance(self, n):
        L, s = self.L, self.s
        if self.kind in ("dt", "dtr"):
            s.mcs(n)
            O.dt_mcs_threaded(self.lat, L, L, 4, 4, self.st, self.m, n, THREADS, self.kind == "dtr")
        elif self.kind == "vl3":
            s.mcs(n)
            O.cbbit_mcs_threaded(self.lat, L, L, self.st, n, 0, 0.5, 0.0, THREADS)
        elif self.kind == "pq":
            s.mcsPQSyncRng(n)
            O.cbbit_mcs_threaded(self.lat, L, L, self.st, n, 1, self.pq[0], self.pq[1], THREADS)
        else:
            s.mcs(n)  # SchedulerCPUCB::mcs runs p = .95, q = 0 (kpz/schedulerCPUCB.cpp:82-86)
            O.cb_mcs_threaded(self.lat, L, L, self.st, n, 0.95, 0.0, THREADS)
        self.t += n
        assert s.mcsCount() == self.t

    def tag(self):
        self.s.correlateTag()
        self.snaps.append((self.t, self.lat.copy()))

    def check_reductions(self):
        L = self.L
        ok, counts = self.s.consistencyCheck()
        assert ok and counts == (0, 0, 0)
        r = self.s.roughness()
        assert (r.sum_h, r.sum_h2) == O.roughness_rows(self.lat, self.enc, L, L, THREADS)
        if self.snaps:
            got = self.s.correlateAll()
            assert [t for t, _ in got] == [t for t, _ in self.snaps]
            for (t, c), (_, snap) in zip(got, self.snaps):
                exp = O.correlate_rows(self.lat, snap, self.enc, L, L, THREADS)
                assert corr_tuple(c) == exp, f"tag at MCS {t}"
            assert corr_tuple(self.s.correlate()) == corr_tuple(got[-1][1])

    def check_state(self):
        got = self.s.popSystem()
        if not np.array_equal(got, self.lat):
            bad = np.flatnonzero(got != self.lat)
            pytest.fail(f"{self.kind} {1 << self.L}^2 after {self.t} MCS: {bad.size} lattice words differ, "
                        f"first at {bad[:8]}")
        del got
        assert np.array_equal(self.s.streams(), self.st.states)

    def check_master(self):
        assert self.s.master_draw() == self.m.close_open()

    def close(self):
        self.s.close()
        self.snaps = []
        self.lat = None


@pytest.mark.timeout(300)
@pytest.mark.parametrize("kind,threads,n", [("dt", 0, 100), ("dt", 1000, 120), ("dtr", 0, 100),
                                            ("vl3", 0, 100), ("sca", 0, 100)])
def test_4096_long(kind, threads, n):
    r = Run(kind, 12, 1234 + n + threads, threads=threads)
    r.advance(n // 2)
    r.check_state()
    r.advance(n - n // 2)
    r.check_state()
    r.check_reductions()
    r.check_master()
    r.close()


@pytest.mark.timeout(300)
def test_4096_long_scabit_pq_with_tags():
    r = Run("pq", 12, 99, pq=(0.95, 0.05))
    r.advance(10)
    r.tag()
    r.advance(40)
    r.tag()
    r.check_reductions()
    r.advance(50)
    r.check_state()
    r.check_reductions()
    r.check_master()
    r.close()


@pytest.fixture(scope="module")
def full_dtr():
    r = Run("dtr", 16, 2024)
    assert r.s.threads == 1 << 22
    yield r
    r.close()


@pytest.fixture(scope="module")
def full_pq():
    r = Run("pq", 16, 4242, pq=(0.95, 0.05))
    assert r.s.threads == 1 << 16
    yield r
    r.close()


@pytest.mark.timeout(600)
@pytest.mark.parametrize("fix", ["full_dtr", "full_pq"])
def test_65536_stage1_tag_at_mcs2(fix, request):
    r = request.getfixturevalue(fix)
    r.advance(2)
    r.tag()
    r.check_reductions()


@pytest.mark.timeout(600)
@pytest.mark.parametrize("fix", ["full_dtr", "full_pq"])
def test_65536_stage2_tag_at_mcs5(fix, request):
    r = request.getfixturevalue(fix)
    assert r.t == 2
    r.advance(3)
    r.tag()
    r.check_state()
    r.check_reductions()


@pytest.mark.timeout(600)
@pytest.mark.parametrize("fix", ["full_dtr", "full_pq"])
def test_65536_stage3_mcs10(fix, request):
    r = request.getfixturevalue(fix)
    assert r.t == 5
    r.advance(5)
    r.check_state()
    r.check_reductions()
    r.check_master()
    r.close()
