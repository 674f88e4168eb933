"""Write the HDF5 checkpoint fixtures of tests/golden/h5/.

The files follow the reference's collectH5 layout (kpz/schedulerService.cpp:
583-612 through libSplash's serial collector) and are written by
h5_fixture_writer.c with the HDF5 C API, independently of the engine's own
writer.  Content: a 64x64 ENC_LOCAL lattice evolved with the oracle's
SchedulerCPUDT restatement to t = 40 MCS, correlator tags at 20 (".corrTag",
stop 1000) and 10 (legacy "corrTag", stop 30), the master dSFMT state at 40
as 770 u32 words, the class rates and prevMeasureOverride = 37.  Malformed
variants (rank3, npot, tagdims, nodata) check the reader's validation.

    python tests/golden/make_h5_fixture.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))
from oracle import oracle as O  # noqa: E402

HDF5 = "/opt/conda"
LX = LY = 6


def main():
    out = os.path.join(HERE, "h5")
    os.makedirs(out, exist_ok=True)
    m = O.Dsfmt(2024)
    st = O.Streams(4, m)
    lat = O.flat(LX, LY, O.ENC_LOCAL)
    O.dt_mcs(lat, LX, LY, 4, 4, st, m, 10)
    tag2 = lat.copy()
    O.dt_mcs(lat, LX, LY, 4, 4, st, m, 10)
    tag1 = lat.copy()
    O.dt_mcs(lat, LX, LY, 4, 4, st, m, 20)
    raw = bytes(m.buf)[:3076] + b"\0" * 4  # dsfmt_t with 8-byte alignment: 3080 bytes
    N = 1 << (LX + LY)
    avg = [O.roughness(t, O.ENC_LOCAL, LX, LY)[0] / N for t in (tag1, tag2)]
    np.savez(os.path.join(out, "ref_expected.npz"), lattice=lat, tag20=tag1, tag10=tag2,
             dsfmt=np.frombuffer(raw, np.uint8), mcs=np.uint32(40), pmo=np.uint32(37),
             P=np.array([0.9, 0.3, 0.6, 0.1]), Q=np.array([0.05, 0.7, 0.2, 0.5]))
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "w")
        subprocess.check_call(["gcc", "-O1", "-o", exe, os.path.join(HERE, "h5_fixture_writer.c"),
                               f"-I{HDF5}/include", f"-L{HDF5}/lib", "-lhdf5", f"-Wl,-rpath,{HDF5}/lib"])
        files = []
        for name, arr in (("lat", lat), ("t1", tag1), ("t2", tag2)):
            p = os.path.join(tmp, name + ".bin")
            arr.astype(np.uint32).tofile(p)
            files.append(p)
        dp = os.path.join(tmp, "dsfmt.bin")
        open(dp, "wb").write(raw)
        files.append(dp)
        for variant in ("ref", "rank3", "npot", "tagdims", "nodata"):
            subprocess.check_call([exe, os.path.join(out, variant + ".h5"), variant, str(1 << (LX - 2)),
                                   str(1 << (LY - 2)), "40", *files, repr(avg[0]), repr(avg[1])])
    print("wrote", out)


if __name__ == "__main__":
    main()
