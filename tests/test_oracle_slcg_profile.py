"""The oracle's SLCG64 output fold random() = lo ^ hi and its 23-bit
randomFloat() mask (kpz/kpzRandom.cu:62-68), pinned by the reference's own
profiling program inc/slcgProfile.cpp, built unchanged into oracle/_ref: its
printed output (tests/golden/slcg_profile.txt, written by
tests/golden/make_slcg_profile.py) holds the seed, the low state bit of the
first 20 steps, the count of each of the 32 fold bits over the next 2^30
draws and the 12-bucket histograms of the fold's top byte and of
randomFloat().  The oracle replays the same 2^30 draws and must reproduce
every number.  A dropped high half (lo only), a shifted fold (lo ^ hi >> 1) or
a 24-bit mask changes bit counts or the second histogram by far more than
the counts' statistical spread (about 2^15)."""
import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
A, Cc, M64 = 2862933555777941757, 1442695040888963407, (1 << 64) - 1


def parse():
    with open(os.path.join(HERE, "golden", "slcg_profile.txt")) as f:
        lines = [ln.strip() for ln in f if ln.strip()]
    assert lines[0].startswith("seed ")
    seed = int(lines[0].split()[1])
    low = [int(v) for v in lines[1:21]]
    assert lines[21] == "End of bit 0 samples."
    n = int(lines[22].split()[0])
    assert lines[23] == "bits:"
    bits = [int(lines[24 + b].split()[1]) for b in range(32)]
    assert [int(lines[24 + b].split()[0]) for b in range(32)] == list(range(32))
    assert lines[56] == "jumps:"
    j = [int(lines[57 + k].split()[1]) for k in range(12)]
    j2 = [int(lines[57 + k].split()[2]) for k in range(12)]
    return seed, low, n, bits, j, j2


def test_fixture_is_the_reference_program_output():
    seed, low, n, bits, j, j2 = parse()
    assert n == 1 << 30
    assert sum(j) == n and sum(j2) == n


def test_low_state_bits_of_base_generator():
    # SLCGen (inc/kmcRandom.h:122-124): the state's bit 0 alternates
    seed, low, *_ = parse()
    s, got = seed, []
    for _ in range(20):
        s = (s * A + Cc) & M64
        got.append(s & 1)
    assert got == low


@pytest.mark.timeout(300)
def test_fold_and_randomfloat_reproduce_reference_profile():
    seed, low, n, bits, j, j2 = parse()
    s = seed
    for _ in range(20):
        s = (s * A + Cc) & M64
    nthreads = max(1, min(8, os.cpu_count() or 1))
    gb, gj, gj2 = O.slcg_profile(s, n, nthreads)
    assert [int(v) for v in gb] == bits
    assert [int(v) for v in gj] == j
    assert [int(v) for v in gj2] == j2


def test_profile_detects_planted_fold_mistakes():
    # the same statistics from a numpy replay of 2^16 draws: the oracle's
    # counts equal the exact fold's, and differ from the planted variants'
    s0 = 0x0123456789ABCDEF
    n = 1 << 16
    st = np.empty(n, np.uint64)
    s = s0
    for i in range(n):
        s = (s * A + Cc) & M64
        st[i] = s
    lo = (st & np.uint64(0xFFFFFFFF)).astype(np.uint64)
    hi = st >> np.uint64(32)

    def counts(t):
        return [int(((t >> np.uint64(b)) & np.uint64(1)).sum()) for b in range(32)]

    def hist2(t):
        r = (t & np.uint64(0x7FFFFF)).astype(np.float32)
        f = r / np.float32(8388608.0) * np.float32(12.0)
        return np.bincount(f.astype(np.int64), minlength=12).tolist()

    gb, gj, gj2 = O.slcg_profile(s0, n, 3)
    exact = lo ^ hi
    assert [int(v) for v in gb] == counts(exact)
    assert [int(v) for v in gj2] == hist2(exact)
    for planted in [lo, lo ^ (hi >> np.uint64(1)), (lo ^ hi) ^ np.uint64(1) << np.uint64(7)]:
        assert counts(planted) != counts(exact)
    assert hist2(exact) != hist2((exact & np.uint64(0xFFFFFF)) >> np.uint64(1))
