"""Command-line handling of the kpz_scaling driver (the reference's
kpzScaling front end, kpz/kpzScaling.cpp:47-100, kpz/kpzFrontend.cpp:93-232,
400-733).  These cases fail before any device is touched."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cudakpz_amd", "kpz_scaling")

pytestmark = pytest.mark.skipif(not os.path.exists(EXE), reason="kpz_scaling not built (make -C cudakpz_amd/csrc)")


def run(*args):
    p = subprocess.run([EXE, *map(str, args)], capture_output=True, text=True, timeout=60)
    return p.returncode, p.stdout, p.stderr


def test_usage_without_arguments():
    rc, out, _ = run()
    assert rc == 1 and out.startswith("kpz_scaling mcs [x [y]]")


def test_missing_mcs():
    rc, _, err = run("--scheduler", "dt")
    assert rc == 1 and "Number of mcs was not supplied" in err


@pytest.mark.parametrize("args,msg", [
    (["100", "8", "--scheduler", "db4"], "No such scheduler: db4"),
    (["100", "8", "--bogus"], "Unknown argument: --bogus"),
    (["100", "8", "--pq"], "--pq needs p"),
    (["100", "8", "--pq5", "0.5"], "rate sets are --pq"),
    (["100", "8", "--mcsFct", "nop"], "not supported"),
    (["100", "8", "--saveH5"], "--saveH5 needs an HDF5 file-name prefix"),
    (["100", "8", "--loadH5", "/nonexistent/ckpt"], "cannot open /nonexistent/ckpt_0_0_0.h5"),
    (["100", "8", "--loadH5", "/nonexistent/ckpt.h5"], "cannot open /nonexistent/ckpt.h5"),
    (["100", "8", "--loadH5", "a", "--loadState", "b"], "exclude each other"),
    (["100", "8", "--corrStart"], "--corrStart needs the tag time"),
    (["100", "8", "--linSample", "0"], "--linSample"),
    (["100", "8", "--maxH", "-1"], "--maxH needs a positive number of hours"),
    (["100", "8", "--scheduler", "dt", "--mcsThreads"], "--mcsThreads needs a thread count"),
    (["100", "8", "--scheduler", "dt", "--ddBlocksize"], "--ddBlocksize needs a block size"),
    (["100", "8", "--scheduler", "sca", "--1d"], "--1d"),
    (["100", "8", "-f"], "-f / --file needs a file name"),
    (["100", "8", "--rng", "mt19937"], "--rng takes slcg64 or tinymt32"),
    (["100", "8", "--tinymtParams"], "--tinymtParams needs a file name"),
    (["100", "8", "--scheduler", "dt", "--rng", "tinymt32"], "--rng tinymt32 needs --tinymtParams"),
])
def test_argument_errors(args, msg):
    rc, _, err = run(*args)
    assert rc == 1
    assert msg in err
