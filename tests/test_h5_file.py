"""HDF5 checkpoints (include/kpz_amd_h5.h), host side: the library exports
what its header declares, and the reader validates a file completely before
it touches a device.  The fixtures in tests/golden/h5/ follow the layout of
the reference's collectH5 (kpz/schedulerService.cpp:583-612) and were written
by tests/golden/h5_fixture_writer.c, independently of the engine's writer."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "kpz_amd_h5.h")
H5LIB = os.path.join(ROOT, "cudakpz_amd", "libkpz_amd_h5.so")
FIX = os.path.join(ROOT, "tests", "golden", "h5")

pytestmark = pytest.mark.skipif(not os.path.exists(H5LIB), reason="libkpz_amd_h5.so not built (no HDF5 C library)")


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    return sorted(set(re.findall(r"\b(kpz_[a-z0-9_]+)\s*\(", text)))


def test_h5_library_exports_every_declared_symbol():
    import cudakpz_amd as K
    L = K.lib_h5()
    names = declared_functions()
    assert names == ["kpz_read_h5", "kpz_write_h5"]
    assert all(hasattr(L, n) for n in names)


def test_null_arguments_are_rejected():
    import cudakpz_amd as K
    L = K.lib_h5()
    assert L.kpz_write_h5(None, b"x", b"", 0) == -1
    assert L.kpz_read_h5(None, b"", 0, None, None, None, None) == -1
    assert b"null" in K.lib().kpz_last_error()


@pytest.mark.parametrize("variant,msg", [
    ("rank3", "is not a 2-d power-of-two lattice"),
    ("npot", "is not a 2-d power-of-two lattice"),
    ("tagdims", "tag at 20 has no matching .system"),
    ("nodata", "no /data group"),
])
def test_malformed_files_are_rejected_before_device_work(variant, msg):
    import cudakpz_amd as K
    with pytest.raises(K.KpzError) as e:
        K.load_h5(os.path.join(FIX, variant + ".h5"), K.KPZ_DT)
    assert msg in str(e.value)


def test_reference_layout_parses_up_to_the_device():
    """Everything in ref.h5 is read and accepted; without a GPU the first
    failure is the scheduler's device allocation, not the file."""
    import cudakpz_amd as K
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present: tests/test_gpu_h5.py covers this file")
    except ImportError:
        pass
    with pytest.raises(K.KpzError) as e:
        K.load_h5(os.path.join(FIX, "ref.h5"), K.KPZ_DT)
    assert "HDF5" not in str(e.value) and e.value.code == -2


def test_fixture_matches_its_record():
    """The stored expectations describe a consistent 64x64 lattice at t = 40
    and a master state of 3080 bytes (dsfmt_t without SSE2 alignment)."""
    from oracle import oracle as O
    z = np.load(os.path.join(FIX, "ref_expected.npz"))
    assert z["lattice"].size == 256 and z["dsfmt"].size == 3080
    for k in ("lattice", "tag20", "tag10"):
        assert O.consistent(z[k], O.ENC_LOCAL, 6, 6)
    assert not np.array_equal(z["lattice"], z["tag20"])
