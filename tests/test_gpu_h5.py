"""HDF5 checkpoints on the GPU: a file in the reference's collectH5 layout
(written independently by tests/golden/h5_fixture_writer.c) loads with the
reference's retrieveH5 semantics, and the engine's own files resume bit for
bit for every scheduler kind (lattice, SLCG streams, master, MCS count,
tags, disorder map, class rates); the kpz_scaling driver's --saveH5 /
--loadH5 reproduce the uninterrupted output table."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import cudakpz_amd as K
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIX = os.path.join(ROOT, "tests", "golden", "h5")
EXE = os.path.join(ROOT, "cudakpz_amd", "kpz_scaling")


def test_reference_layout_file_resumes_like_retrieveH5():
    z = np.load(os.path.join(FIX, "ref_expected.npz"))
    s, pmo, restored = K.load_h5(os.path.join(FIX, "ref.h5"), K.KPZ_DT, threads=4)
    assert (s.lx, s.ly) == (6, 6) and s.mcsCount() == 40 and pmo == 37
    assert restored == K.H5_MASTER | K.H5_RATES | K.H5_PMO  # no SLCG states in a reference file
    assert np.array_equal(s.popSystem(), z["lattice"])
    for n in range(4):
        assert s.rates(n) == (z["P"][n], z["Q"][n])
    # the master is the stored dsfmt_t; the streams were re-seeded from it
    m = O.Dsfmt()
    raw = z["dsfmt"].tobytes()[:len(m.buf)]
    import ctypes as C
    C.memmove(m.buf, raw, len(raw))
    st = O.Streams(4, m)
    assert np.array_equal(s.streams(), st.states)
    # tags: 20 (stop 1000) stays live, 10 (legacy name, stop 30) has expired at 40
    tags = s.correlateAll()
    assert [t for t, _ in tags] == [20]
    s1, t1, s2, s11, tcs = O.correlate(z["lattice"], z["tag20"], O.ENC_LOCAL, 6, 6)
    c = tags[0][1]
    assert (c.sum_h, c.sum_hs, c.sum_h2, c.sum_hhs, c.tcs) == (s1, t1, s2, s11, tcs)
    # and the dynamics continue from exactly that state
    s.mcs(5)
    lat = z["lattice"].copy()
    O.dt_mcs(lat, 6, 6, 4, 4, st, m, 5)
    assert np.array_equal(s.popSystem(), lat)
    assert np.array_equal(s.streams(), st.states)
    assert s.master_draw() == m.close_open()


@pytest.mark.parametrize("cls,kw,mode", [(K.SchedulerDT, {"threads": 5, "layout": (3, 2)}, "kpz"),
                                         (K.SchedulerDTr, {}, "pq"),
                                         (K.SchedulerSCABit, {"threads": 7}, "dis"),
                                         (K.SchedulerSCA, {"threads": 3}, "dis"),
                                         (K.SchedulerDT, {}, "dis")])
def test_round_trip_is_bit_exact(tmp_path, cls, kw, mode):
    lx, ly = 8, 7

    def advance(s, n):
        {"kpz": s.mcs, "pq": s.mcsPQSyncRng, "dis": s.mcsDisorder2SyncRng}[mode](n)

    ref, run = cls(lx, ly, **kw), cls(lx, ly, **kw)
    for s in (ref, run):
        s.supplySeed(123)
        s.initHomogeneous()
        if mode == "pq":
            s.setPQ(0.7, 0.2)
        if mode == "dis":
            s.setPQ(0.9, 0.1)
            s.setPQ(0.4, 0.3, 1)
            s.setPQ(0.6, 0.0, 2)
            s.generateDisorder2(0.4)
        advance(s, 3)
        s.correlateTag(stop=60)
        advance(s, 2)
        s.correlateTag(stop=8)
        advance(s, 4)
    base = str(tmp_path / "ck")
    run.saveH5(base, prevMeasureOverride=5)
    kind = run.kind
    run.close()
    res, pmo, restored = K.load_h5(base, kind, **kw)
    want = K.H5_MASTER | K.H5_STREAMS | K.H5_RATES | K.H5_PMO | (K.H5_DISORDER if mode == "dis" else 0)
    assert restored == want and pmo == 5 and res.mcsCount() == 9 and type(res) is cls
    for n in range(4):
        assert res.rates(n) == ref.rates(n)
    if mode == "dis":
        assert np.array_equal(res.disorder(), ref.disorder())
    for s in (ref, res):
        advance(s, 6)
    assert np.array_equal(res.popSystem(), ref.popSystem())
    assert np.array_equal(res.streams(), ref.streams())
    assert res.master_draw() == ref.master_draw()
    a, b = res.correlateAll(), ref.correlateAll()
    assert [m for m, _ in a] == [m for m, _ in b] == [3]
    assert a[0][1] == b[0][1]


def test_written_file_has_the_reference_names(tmp_path):
    h5ls = shutil.which("h5ls") or "/opt/conda/bin/h5ls"
    if not os.path.exists(h5ls):
        pytest.skip("h5ls not installed")
    s = K.SchedulerDT(8, 6, threads=4)
    s.supplySeed(1)
    s.initHomogeneous()
    s.mcs(2)
    s.correlateTag()
    s.mcs(3)
    s.saveH5(str(tmp_path / "f.h5"), prefix="p")
    out = subprocess.run([h5ls, "-r", str(tmp_path / "f.h5")], capture_output=True, text=True, timeout=60).stdout
    for line in ["/custom Group", "/data/2/p.system          Dataset {16, 64}", "/data/5/p.system          Dataset {16, 64}",
                 "/data/5/p.dsfmt           Dataset {770}", "/data/5/p.slcg64          Dataset {4}"]:
        assert " ".join(line.split()) in [" ".join(x.split()) for x in out.splitlines()], (line, out)
    h5dump = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    attrs = subprocess.run([h5dump, "-A", "-g", "/data/2", str(tmp_path / "f.h5")], capture_output=True, text=True,
                           timeout=60).stdout
    assert 'ATTRIBUTE "p.corrTag"' in attrs and "4294967295" in attrs


def test_driver_h5_resume_reproduces_table(tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("kpz_scaling not built")

    def drive(*args):
        p = subprocess.run([EXE, *map(str, args)], capture_output=True, text=True, timeout=120)
        assert p.returncode == 0, p.stderr
        return p.stdout

    common = ["--scheduler", "scaBit", "--pq", 0.95, 0.05, "--corrStart", 20, 300]
    drive(400, 8, *common, "--seed", 11, "-f", tmp_path / "full.txt")
    drive(200, 8, *common, "--seed", 11, "-f", tmp_path / "a.txt", "--saveH5", tmp_path / "ck")
    assert (tmp_path / "ck_0_0_0.h5").exists()
    out = drive(400, 8, *common, "--loadH5", tmp_path / "ck", "-f", tmp_path / "b.txt")
    assert "Loading Hdf5 file at t = 200" in out and "randomizing" not in out
    full = (tmp_path / "full.txt").read_text().splitlines()
    b = (tmp_path / "b.txt").read_text().splitlines()
    t = lambda line: int(line.split("\t")[0])
    assert [x for x in b if t(x) > 200] == [x for x in full if t(x) > 200]
    assert len([x for x in b if t(x) > 200]) > 5


@pytest.mark.parametrize("cls,T", [(K.SchedulerDTr, 9), (K.SchedulerSCABit, 5), (K.SchedulerSCA, 4)])
def test_tinymt32_round_trip(tmp_path, cls, T):
    # TinyMT32::writeH5 / readH5 (kpz/GPUTinyMT.cu:113-144): <prefix>.tinyMT,
    # extents {7, T}, the state planes with the parameter sets
    r = np.random.default_rng(T)
    mats = [r.integers(0, 1 << 32, T, dtype=np.uint64).astype(np.uint32) for _ in range(3)]
    ref, run = (cls(8, 7, threads=T, rng=K.RNG_TINYMT32) for _ in range(2))
    for s in (ref, run):
        s.setTinyMTParams(*mats)
        s.supplySeed(77)
        s.initHomogeneous()
        s.mcs(3)
    base = str(tmp_path / "tm")
    run.saveH5(base, prefix="k")
    run.close()
    res, _, restored = K.load_h5(base, ref.kind, prefix="k", threads=T, rng=K.RNG_TINYMT32)
    assert restored & K.H5_STREAMS and res.rng == K.RNG_TINYMT32
    assert np.array_equal(res.tinymt(), ref.tinymt())
    for s in (ref, res):
        s.mcs(4)
    assert np.array_equal(res.popSystem(), ref.popSystem())
    assert np.array_equal(res.tinymt(), ref.tinymt())
    h5ls = shutil.which("h5ls") or "/opt/conda/bin/h5ls"
    if os.path.exists(h5ls):
        out = subprocess.run([h5ls, "-r", base + "_0_0_0.h5"], capture_output=True, text=True, timeout=60).stdout
        assert f"/data/3/k.tinyMT Dataset {{7, {T}}}" in [" ".join(x.split()) for x in out.splitlines()], out
    with pytest.raises(K.KpzError):
        # an SLCG64 checkpoint has no .tinyMT state to resume TinyMT32 workers from
        a = cls(8, 7, threads=T)
        a.supplySeed(1)
        a.initHomogeneous()
        a.saveH5(str(tmp_path / "sl"))
        K.load_h5(str(tmp_path / "sl"), a.kind, threads=T, rng=K.RNG_TINYMT32)

