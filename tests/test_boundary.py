"""Host-side checks of the C ABI library that need no GPU: it loads, exports
every symbol include/kpz_amd.h declares, and rejects bad input cleanly."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "kpz_amd.h")
LIB = os.path.join(ROOT, "cudakpz_amd", "libkpz_amd.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kpz_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        import cudakpz_amd
        cudakpz_amd.build()
    return C.CDLL(LIB)


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ["kpz_create", "kpz_destroy", "kpz_set_seed", "kpz_set_pq", "kpz_init_flat",
                 "kpz_upload", "kpz_download", "kpz_mcs", "kpz_mcs_pq", "kpz_sync",
                 "kpz_roughness", "kpz_corr_tag", "kpz_correlate", "kpz_last_error"]:
        assert must in names


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_null_arguments_are_rejected(lib):
    lib.kpz_last_error.restype = C.c_char_p
    assert lib.kpz_mcs(None, 1) == -1
    assert b"null" in lib.kpz_last_error()
    assert lib.kpz_create(0, 8, 8, None, None) == -1
    assert lib.kpz_destroy(None) == 0
    assert lib.kpz_master_bytes() > 3000
    out = (C.c_int64 * 3)()
    assert lib.kpz_consistency_check(None, 1, out) == -1  # no scheduler: an error, not "inconsistent"
    assert b"null" in lib.kpz_last_error()


def test_measurement_utilities_reject_bad_arguments(lib):
    lib.kpz_last_error.restype = C.c_char_p
    r = C.c_double()
    for bad in (1, 6, 7, 9, -1):
        assert lib.kpz_valu_mix_rate(-1, 1, bad, C.byref(r)) == -1  # rejected before any device call
        assert b"slow_of_8" in lib.kpz_last_error()
    assert lib.kpz_valu_mix_rate(-1, 0, 3, C.byref(r)) == -1
    assert lib.kpz_valu_mix_rate(-1, 1, 3, None) == -1
    assert lib.kpz_valu_peak(-1, 1, None, None) == -1


def test_python_mirror_fails_loudly_without_library(monkeypatch, tmp_path):
    import cudakpz_amd as K
    monkeypatch.setattr(K, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(K, "_LIB", None)
    with pytest.raises(K.KpzError):
        K.SchedulerDT(8, 8)


def test_unknown_scheduler_id():
    import cudakpz_amd as K
    with pytest.raises(K.KpzError):
        K.make_scheduler("db4", 8, 8)
