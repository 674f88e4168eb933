"""The oracle's RNG and geometry pieces against golden vectors produced by the
reference's own compiled code (tests/golden/make_golden.py), plus the SLCG
closed forms."""
import json
import os

import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
M64 = (1 << 64) - 1
A, Cc = 2862933555777941757, 1442695040888963407


def load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def test_dsfmt_matches_reference():
    for rec in load("dsfmt.json"):
        d = O.Dsfmt(rec["seed"], reseed=False)
        for _ in range(rec["discard"]):
            d.close_open()
        got = [d.close_open().hex() for _ in rec["draws"]]
        assert got == rec["draws"], (rec["seed"], rec["discard"])


def test_shuffle4_matches_reference():
    for rec in load("shuffle4.json"):
        assert rec["discard"] == 100000
        d = O.Dsfmt(rec["seed"])  # reSeed: init + 100000 discards
        got = [d.shuffle4() for _ in rec["perms"]]
        assert got == rec["perms"]
        for p in got:
            assert sorted(p) == [0, 1, 2, 3]


def test_dt_offsets_match_reference():
    for rec in load("dt_geometry.json")["offsets"]:
        for i, off in enumerate(rec["offsets"]):
            got = O.dt_block_offset(rec["lx"], rec["ly"], rec["lbx"], rec["lby"], rec["ds"], rec["shift"], i)
            assert list(got) == off, (rec["lx"], rec["ly"], rec["ds"], rec["shift"], i)


def test_dt_offsets_tile_the_lattice():
    # the four domain sets of an unshifted DT cover every block exactly once
    for (lx, ly, lbx, lby) in [(9, 9, 4, 4), (6, 5, 2, 2), (10, 8, 4, 3)]:
        nb = 1 << (lx - lbx + ly - lby - 2)
        seen = set()
        for ds in range(4):
            for i in range(nb):
                seen.add(O.dt_block_offset(lx, ly, lbx, lby, ds, (0, 0), i))
        assert len(seen) == 4 * nb
        assert all(x % (1 << lbx) == 0 and y % (1 << lby) == 0 for x, y in seen)


def test_dt_worker_walk_matches_reference():
    # Func_MCSBase::exec visits blocks id, id+T, id+2T, ... (DDIterator::next)
    for rec in load("dt_geometry.json")["walks"]:
        nb = 1 << (rec["lx"] - rec["lbx"] + rec["ly"] - rec["lby"] - 2)
        assert rec["blocks"] == list(range(rec["id"], nb, rec["T"]))


def test_slcg_skip_known_answers():
    g = load("slcg.json")
    for rec in g["table"] + g["template_small"] + g["function"]:
        assert O.slcg_skip(rec["S"]) == (rec["a"], rec["c"]), rec["S"]
    for rec in g["base_step"]:
        assert (rec["in"] * A + Cc) & M64 == rec["out"]


def test_slcg_skip_closed_form():
    # A_S = A^S, C_S = C * sum_{k=1..S} A^k, independent of the oracle's loop
    for S in [1, 2, 3, 5, 30721, 4194305]:
        a, c = O.slcg_skip(S)
        assert a == pow(A, S, 1 << 64)
        # geometric series by the doubling formula on Python ints
        acc, pw = 0, 1
        for _ in range(S if S < 40000 else 0):
            pw = (pw * A) & M64
            acc = (acc + pw) & M64
        if S < 40000:
            assert c == (Cc * acc) & M64
        # the skip increment is A times the true S-step jump increment
        ja, jc = O.slcg_jump(A, Cc, S)
        assert ja == a and c == (A * jc) & M64


def test_slcg_jump_equals_iteration():
    for (a, c) in [(A, Cc), O.slcg_skip(30721), O.slcg_skip(65537)]:
        s = 0x0123456789ABCDEF
        x = s
        for k in range(0, 300):
            ak, ck = O.slcg_jump(a, c, k)
            assert (ak * s + ck) & M64 == x
            x = (a * x + c) & M64
        # large k: composition law J(m+n) = J(m) o J(n)
        for m, n in [(1000, 12345), (1 << 20, 3), (77777, 88888)]:
            am, cm = O.slcg_jump(a, c, m)
            an, cn = O.slcg_jump(a, c, n)
            amn, cmn = O.slcg_jump(a, c, m + n)
            assert (am * an) & M64 == amn and (am * cn + cm) & M64 == cmn


def test_default_skip_is_odd_and_exceeds_T():
    assert O.default_skip(1) == 30721
    assert O.default_skip(30720) == 30721
    for T in [30721, 30722, 1 << 22, (1 << 22) + 1, 10 ** 6 + 3]:
        S = O.default_skip(T)
        assert S > T and S % 2 == 1 and S - T <= 2
        a, c = O.slcg_skip(S)
        assert a % 8 == 5 and c % 2 == 1  # full period 2^64 (Hull-Dobell)


def test_stream_seeding_follows_slcg64_randomize():
    # kpz/GPURandom.cu:42-54: s0 from two master draws, s_t = A s_{t-1} + C
    m = O.Dsfmt(42)
    ref = m.copy()
    st = O.Streams(5, m)
    d1, d2 = ref.close_open(), ref.close_open()
    s0 = int(d1 * 4294967295.0) ^ (int(d2 * 4294967295.0) << 32)
    exp = [s0]
    for _ in range(4):
        exp.append((exp[-1] * A + Cc) & M64)
    assert [int(v) for v in st.states] == exp
    # draw = fold of the skipped state
    s = exp[0]
    s = (st.a * s + st.c) & M64
    assert st.draw(0) == ((s ^ (s >> 32)) & 0xFFFFFFFF)


def test_seed32_identity_for_32bit_seeds():
    for s in [0, 1, 42, 0xFFFFFFFF]:
        assert O.seed32(s) == s
    assert O.seed32((5 << 32) | 7) == 2


def test_tinymt32_matches_reference():
    recs = load("tinymt32.json")["generators"]
    assert len(recs) == 25
    for rec in recs:
        g = O.Tmt(rec["mat1"], rec["mat2"], rec["tmat"], rec["seed"])
        assert [int(v) for v in g.g[:4]] == rec["init"], rec["seed"]
        assert [g.u32() for _ in rec["u32"]] == rec["u32"], (rec["mat1"], rec["seed"])
        h = O.Tmt(rec["mat1"], rec["mat2"], rec["tmat"], rec["seed"])
        assert [float(h.float()).hex() for _ in rec["float"]] == rec["float"]


def test_tinymt32_scheduler_streams_match_reference():
    # RngIface::TinyMT32::randomize from a reSeed'ed master, then random() and
    # randomFloat() alternately, per generator
    for rec in load("tinymt32.json")["interface"]:
        m = O.Dsfmt(rec["seed"])
        assert rec["discard"] == 100000
        T, n = len(rec["mat1"]), rec["n"]
        st = O.TmtStreams(T, m, rec["mat1"], rec["mat2"], rec["tmat"])
        got_u, got_f = [], []
        for t in range(T):
            for _ in range(n):
                got_u.append(st.u32(t))
                got_f.append(float(st.float(t)).hex())
        assert got_u == rec["random"]
        assert got_f == rec["randomFloat"]
