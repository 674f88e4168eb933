"""The kpzConsistencyCheck tool (kpz/kpzConsistencyCheck.cpp:38-86) on the
engine, cudakpz_amd/kpz_consistency_check: argument handling and file-name
matching on the CPU (nothing here touches a device), Good / Bad verdicts on
HDF5 files on the GPU."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cudakpz_amd", "kpz_consistency_check")

pytestmark = pytest.mark.skipif(not os.path.exists(EXE), reason="kpz_consistency_check not built (needs HDF5)")


def run(*args):
    p = subprocess.run([EXE, *map(str, args)], capture_output=True, text=True, timeout=120)
    return p.returncode, p.stdout, p.stderr


def test_ana_threads_needs_a_number():
    for args in (["--anaThreads"], ["--anaThreads", "x"]):
        rc, _, err = run(*args)
        assert rc == 1 and "Expected number of threads for --anaThreads." in err


def test_names_and_unreadable_files_are_skipped(tmp_path):
    missing = tmp_path / "run_0_0_0.h5"
    rc, out, err = run("--anaThreads", 3, "plain.h5", tmp_path / "run_0_0.h5", missing)
    assert rc == 0
    assert "[MM] running 3 threads for analysis (per process)" in out
    assert "[WW] Skipping file with non-matching name: plain.h5" in err
    assert f"[WW] Skipping file with non-matching name: {tmp_path / 'run_0_0.h5'}" in err
    assert f"[EE] Failed to read file: {missing}" in err
    assert "Good" not in out and "Bad" not in out


@pytest.mark.gpu
def test_good_and_bad_files(tmp_path):
    import cudakpz_amd as K
    lx, ly = 8, 7
    s = K.SchedulerSCA(lx, ly)
    s.supplySeed(3)
    s.initHomogeneous()
    s.mcs(4)
    s.correlateTag()
    s.mcs(2)
    good = str(tmp_path / "good_0_0_0.h5")
    s.saveH5(good)
    dis = str(tmp_path / "auto_1_0_0.h5")  # rank 1: the "dis" prefix
    s.saveH5(dis, prefix="dis")
    lat = s.popSystem()
    lat[5] ^= np.uint32(1 << 8)  # the x slope of site (4*5 + 0, 1): its row, two loops
    s.pushSystem(lat)
    bad = str(tmp_path / "bad_0_0_0.h5")
    s.saveH5(bad)
    rc, out, err = run(good, dis, bad)
    assert rc == 0, err
    assert f"[MM] Good: {good}" in out
    assert f"[MM] Good: {dis}" in out
    assert f"[MM] Bad: {bad}" in out
    assert "in m_system: 1 lines, 0 columns, 2 loops" in out
    s.close()
