"""SchedulerService::consitencyCheck (kpz/schedulerService.cpp:721-854) on the
device, through kpz_consistency_check: counts of rows whose x slopes and
columns whose y slopes do not sum to zero and of loops
sX(x,y) + sY(x-1,y) != sY(x,y) + sX(x,y-1).  Pinned by closed forms (one
flipped x slope breaks its row and the two loops that hold it, one flipped y
slope its column and two loops, a balanced pair of x flips in one row breaks
four loops and no row), by a numpy count on random lattices and by the
oracle's yes/no on lattices the dynamics produced."""
import numpy as np
import pytest

import cudakpz_amd as K
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def slopes(local, lx, ly):
    """sX[y, x], sY[y, x] of a 4x4-cell lattice (SystemSize::shift: bit
    2*(x&3) + 8*(y&3) of word (y>>2)*(L/4) + (x>>2); y slope one bit up)."""
    L, Y = 1 << lx, 1 << ly
    w = np.asarray(local, np.uint32).reshape(Y // 4, L // 4)
    sX = np.zeros((Y, L), np.int64)
    sY = np.zeros((Y, L), np.int64)
    for ry in range(4):
        for rx in range(4):
            sh = np.uint32(2 * rx + 8 * ry)
            sX[ry::4, rx::4] = (w >> sh) & 1
            sY[ry::4, rx::4] = (w >> (sh + np.uint32(1))) & 1
    return sX, sY


def counts(local, lx, ly):
    sX, sY = slopes(local, lx, ly)
    L, Y = 1 << lx, 1 << ly
    rows = int(np.count_nonzero(2 * sX.sum(1) != L))
    cols = int(np.count_nonzero(2 * sY.sum(0) != Y))
    loops = int(np.count_nonzero(sX + np.roll(sY, 1, axis=1) != sY + np.roll(sX, 1, axis=0)))
    return rows, cols, loops


def flip(local, lx, x, y, d):
    out = np.array(local, np.uint32)
    out[(y >> 2) * ((1 << lx) >> 2) + (x >> 2)] ^= np.uint32(1 << (2 * (x & 3) + 8 * (y & 3) + d))
    return out


class Engine:
    """an engine and the conversion between 4x4-cell words and its own layout"""

    def __init__(self, cls, lx, ly):
        self.s = cls(lx, ly)
        self.lx, self.ly = lx, ly
        self.cb = self.s.encoding() == O.ENC_CBBIT

    def push(self, local):
        self.s.pushSystem(O.local_to_cbbit(local, self.lx, self.ly) if self.cb else local)

    def pop(self):
        got = self.s.popSystem()
        return O.cbbit_to_local(got, self.lx, self.ly) if self.cb else got


SIZES = [
    (K.SchedulerSCA, 3, 3),       # 8 x 8: one chunk of 8 sites per row
    (K.SchedulerSCA, 5, 4),
    (K.SchedulerSCA, 6, 9),       # two bands of 256 rows
    (K.SchedulerDT, 9, 8),
    (K.SchedulerDT, 12, 10),      # 64 chunks per row
    (K.SchedulerDT, 13, 6),       # 128 chunks per row: two waves per row
    (K.SchedulerSCABit, 6, 2),
    (K.SchedulerSCABit, 7, 5),
    (K.SchedulerSCABit, 10, 9),
]


@pytest.mark.parametrize("cls,lx,ly", SIZES)
def test_closed_forms(cls, lx, ly):
    e = Engine(cls, lx, ly)
    e.s.initHomogeneous()
    flat = e.pop()
    assert e.s.consistencyCheck() == (True, (0, 0, 0))
    assert e.s.consitencyCheck(False) == (True, (0, 0, 0))
    L, Y = 1 << lx, 1 << ly
    rng = np.random.default_rng(lx * 100 + ly)
    for _ in range(3):
        x, y = int(rng.integers(L)), int(rng.integers(Y))
        e.push(flip(flat, lx, x, y, 0))
        assert e.s.consistencyCheck() == (False, (1, 0, 2)), (x, y)
        e.push(flip(flat, lx, x, y, 1))
        assert e.s.consistencyCheck() == (False, (0, 1, 2)), (x, y)
    # a 1 -> 0 and a 0 -> 1 x flip at two sites of one row: the row
    # still sums to zero, four loops break
    sX, _ = slopes(flat, lx, ly)
    y = int(rng.integers(Y))
    x1 = int(np.flatnonzero(sX[y] == 1)[0])
    x0 = int(np.flatnonzero(sX[y] == 0)[-1])
    e.push(flip(flip(flat, lx, x1, y, 0), lx, x0, y, 0))
    assert e.s.consistencyCheck() == (False, (0, 0, 4))


@pytest.mark.parametrize("cls,lx,ly", SIZES)
def test_random_lattices_against_numpy(cls, lx, ly):
    e = Engine(cls, lx, ly)
    rng = np.random.default_rng(ly * 100 + lx)
    for _ in range(2):
        local = rng.integers(0, 1 << 32, O.size_words(lx, ly), dtype=np.uint64).astype(np.uint32)
        e.push(local)
        ok, c = e.s.consistencyCheck(full=False)
        assert c == counts(local, lx, ly)
        assert ok == (c == (0, 0, 0))


@pytest.mark.parametrize("cls,lx,ly", [(K.SchedulerDT, 9, 9), (K.SchedulerSCA, 8, 7), (K.SchedulerSCABit, 9, 8)])
def test_dynamics_stay_consistent(cls, lx, ly):
    e = Engine(cls, lx, ly)
    e.s.supplySeed(5)
    e.s.initHomogeneous()
    e.s.setPQ(0.8, 0.15)
    e.s.mcsPQSyncRng(6)
    ok, c = e.s.consistencyCheck()
    local = e.pop()
    assert ok and c == (0, 0, 0)
    assert counts(local, lx, ly) == (0, 0, 0)
    assert O.consistent(e.s.popSystem(), e.s.encoding(), lx, ly)


def test_full_checks_the_tags():
    lx, ly = 8, 6
    e = Engine(K.SchedulerDT, lx, ly)
    e.s.initHomogeneous()
    flat = e.pop()
    bad = flip(flat, lx, 3, 5, 1)
    e.push(bad)
    e.s.correlateTag()           # the tag holds a broken lattice
    e.push(flat)
    assert e.s.consistencyCheck(full=False) == (True, (0, 0, 0))
    assert e.s.consistencyCheck(full=True) == (False, (0, 1, 2))
    e.push(bad)                  # both broken: the counts add
    assert e.s.consistencyCheck(full=True) == (False, (0, 2, 4))
