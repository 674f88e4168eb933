"""Batched replicas: every replica is bit-identical to the stand-alone
scheduler with the same kind, size, seed and (p, q) -- alone vs in the mix --
whatever the tuning knobs."""
import numpy as np
import pytest

import cudakpz_amd as K
from oracle import oracle as O

pytestmark = pytest.mark.gpu

MIX = [
    # kind, lx, ly, p, q, seed
    (K.KPZ_DT, 6, 6, 1.0, 0.0, 101),
    (K.KPZ_DT, 8, 7, 0.5, 0.5, 102),
    (K.KPZ_DTR, 7, 7, 1.0, 0.0, 103),
    (K.KPZ_DTR, 6, 8, 0.7, 0.2, 104),
    (K.KPZ_SCABIT, 6, 6, 1.0, 0.0, 105),
    (K.KPZ_SCABIT, 7, 6, 0.5, 0.0, 106),
    (K.KPZ_SCABIT, 9, 7, 0.95, 0.05, 107),
    (K.KPZ_SCABIT, 11, 6, 0.5, 0.5, 108),
    (K.KPZ_DT, 9, 9, 1.0, 0.0, 109),
    (K.KPZ_SCABIT, 8, 8, 0.5, 0.0, 110),
    (K.KPZ_DT, 5, 5, 0.9, 0.1, 111),
    (K.KPZ_SCABIT, 10, 10, 0.6, 0.4, 112),
    # rows of 64 and 128 chunks: one or two passes of a full wave per row in the row statistics
    (K.KPZ_DT, 12, 6, 1.0, 0.0, 113),
    (K.KPZ_SCABIT, 13, 6, 0.5, 0.5, 114),
]

CLS = {K.KPZ_DT: K.SchedulerDT, K.KPZ_DTR: K.SchedulerDTr, K.KPZ_SCABIT: K.SchedulerSCABit}


def alone(rep, n):
    kind, lx, ly, p, q, seed = rep
    s = CLS[kind](lx, ly)
    s.supplySeed(seed)
    s.initHomogeneous()
    if kind != K.KPZ_SCABIT and (p, q) == (1.0, 0.0):
        s.mcs(n)
    else:
        s.setPQ(p, q)
        s.mcsPQSyncRng(n)
    return s


def run_batch(tuning, n, chunks):
    b = K.Batch(MIX, tuning)
    b.initHomogeneous()
    for c in chunks:
        b.mcs(c)
    assert sum(chunks) == n
    return b


@pytest.mark.parametrize("tuning", [None, K.BatchTuning(2, 0, 3, 0, 0), K.BatchTuning(6, 1, 2, 1, 1),
                                    K.BatchTuning(4, 0, 7, 1, 0), K.BatchTuning(6, 1, 2, 2, 1),
                                    K.BatchTuning(3, 0, 5, 2, 0)])
def test_batch_equals_standalone(tuning):
    n = 7
    b = run_batch(tuning, n, [3, 4])
    rough = b.roughness()
    for i, rep in enumerate(MIX):
        s = alone(rep, n)
        assert np.array_equal(b.popSystem(i), s.popSystem()), f"replica {i} {rep}"
        assert np.array_equal(b.streams(i), s.streams()), f"replica {i} streams"
        r = s.roughness()
        assert (rough[i].sum_h, rough[i].sum_h2) == (r.sum_h, r.sum_h2)
        s.close()


@pytest.mark.parametrize("concurrent", [0, 2])
def test_dt_only_batch_in_groups(concurrent):
    """DT/DTr replicas only: one stream, or two lane-balanced groups on two streams."""
    reps = [r for r in MIX if r[0] != K.KPZ_SCABIT]
    b = K.Batch(reps, K.BatchTuning(6, 1, 4, concurrent, 1))
    b.initHomogeneous()
    b.mcs(3)
    b.mcs(6)
    for i, rep in enumerate(reps):
        s = alone(rep, 9)
        assert np.array_equal(b.popSystem(i), s.popSystem()), f"replica {i} {rep}"
        assert np.array_equal(b.streams(i), s.streams()), f"replica {i} streams"
        s.close()
    b.close()


def test_batch_matches_oracle_directly():
    b = run_batch(None, 4, [4])
    for i, (kind, lx, ly, p, q, seed) in enumerate(MIX):
        m = O.Dsfmt(O.seed32(seed))
        st = O.Streams(b.threads(i), m)
        if kind == K.KPZ_SCABIT:
            lat = O.flat(lx, ly, O.ENC_CBBIT)
            O.cbbit_mcs(lat, lx, ly, st, 4, 0 if (p, q) == (0.5, 0.0) else 1, p, q)
        else:
            lat = O.flat(lx, ly, O.ENC_LOCAL)
            kpz = (p, q) == (1.0, 0.0)
            O.dt_mcs(lat, lx, ly, 4, 4, st, m, 4, kind == K.KPZ_DTR, O.UPD_KPZ if kpz else O.UPD_PQ, p, q)
        assert np.array_equal(b.popSystem(i), lat), f"replica {i}"


def test_autotune_returns_valid_tuning():
    reps = MIX[:6]
    t, ms = K.Batch.autotune(reps, mcs=2)
    assert ms > 0 and 0 <= t.max_lanes_log2 <= 6 and t.chunk >= 1
    b = K.Batch(reps, t)
    b.initHomogeneous()
    b.mcs(3)
    s = alone(reps[3], 3)
    assert np.array_equal(b.popSystem(3), s.popSystem())
