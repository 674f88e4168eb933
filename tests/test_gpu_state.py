"""Checkpoint / resume: a scheduler saved after n1 MCS and loaded again
continues bit for bit like an uninterrupted run (lattice, stream states,
master dSFMT, MCS count, live correlator tags); the kpz_scaling driver's
--saveState / --loadState reproduce the uninterrupted output table."""
import os
import subprocess

import numpy as np
import pytest

import cudakpz_amd as K

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cudakpz_amd", "kpz_scaling")


def advance(s, n, pq):
    (s.mcsPQSyncRng if pq else s.mcs)(n)


@pytest.mark.parametrize("cls,kw,pq", [(K.SchedulerDT, {"threads": 5, "layout": (3, 2)}, None),
                                       (K.SchedulerDTr, {}, (0.7, 0.2)),
                                       (K.SchedulerSCABit, {"threads": 7}, (0.95, 0.05)),
                                       (K.SchedulerSCA, {"threads": 3}, None)])
def test_resume_is_bit_exact(tmp_path, cls, kw, pq):
    lx, ly = 8, 7
    ref = cls(lx, ly, **kw)
    run = cls(lx, ly, **kw)
    for s in (ref, run):
        s.supplySeed(99)
        s.initHomogeneous()
        if pq:
            s.setPQ(*pq)
        advance(s, 3, pq)
        s.correlateTag(stop=50)
        advance(s, 4, pq)
    path = tmp_path / "state.kpzstate"
    run.saveState(str(path), aux=1234)
    run.close()
    res, aux = K.load_state(str(path))
    assert aux == 1234 and type(res) is cls and res.mcsCount() == 7
    assert (res.lx, res.ly) == (lx, ly) and res.threads == ref.threads
    for s in (ref, res):
        advance(s, 6, pq)
    assert np.array_equal(res.popSystem(), ref.popSystem())
    assert np.array_equal(res.streams(), ref.streams())
    assert res.master_draw() == ref.master_draw()
    a, b = res.correlateAll(), ref.correlateAll()
    assert [m for m, _ in a] == [m for m, _ in b] == [3]
    assert a[0][1] == b[0][1]


def test_driver_resume_reproduces_table(tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("kpz_scaling not built")

    def drive(*args):
        p = subprocess.run([EXE, *map(str, args)], capture_output=True, text=True, timeout=120)
        assert p.returncode == 0, p.stderr
        return p.stdout

    common = ["--scheduler", "dt", "--corrStart", 20, 300, "--corrStart", 120]
    drive(400, 8, *common, "--seed", 7, "-f", tmp_path / "full.txt")
    drive(200, 8, *common, "--seed", 7, "-f", tmp_path / "a.txt", "--saveState", tmp_path / "s.kpzstate")
    out = drive(400, 8, *common, "--loadState", tmp_path / "s.kpzstate", "-f", tmp_path / "b.txt")
    assert "Loading state at t = 200" in out
    full = (tmp_path / "full.txt").read_text().splitlines()
    a = (tmp_path / "a.txt").read_text().splitlines()
    b = (tmp_path / "b.txt").read_text().splitlines()
    t = lambda line: int(line.split("\t")[0])
    # run a ends with a measurement at t = 200 (its end event), which the
    # uninterrupted run does not make; the resumed run repeats it, then
    # continues the uninterrupted run's schedule and values exactly
    assert [x for x in a if t(x) < 200] == [x for x in full if t(x) < 200]
    assert [x for x in b if t(x) == 200] == [x for x in a if t(x) == 200]
    assert [x for x in b if t(x) > 200] == [x for x in full if t(x) > 200]
    assert len([x for x in b if t(x) > 200]) > 5
