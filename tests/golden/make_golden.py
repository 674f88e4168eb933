"""Write the golden fixtures of tests/golden/*.json.

Runs in the development container only: it builds oracle/_ref/libkpzref.so
from the reference's own sources (inc/dSFMT/dSFMT.cpp plus the header-only
inc/shuffle.h, inc/doubleTilingDD.h, inc/blockDD.h, inc/ddIterators.h,
inc/kmcRandom.h, inc/helper/slcgInst.h, inc/tinyMT/tinymt32.c and the
header part of kpz/CPURngIface.h) via oracle/Makefile and records
their outputs.  The fixtures are data (inputs and expected outputs) and are
what the CPU tests compare the oracle against on any machine.

    python tests/golden/make_golden.py
"""
import ctypes as C
import json
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.join(HERE, "..", "..", "oracle")


def main():
    subprocess.check_call(["make", "-s", "-C", ORACLE, "ref"])
    R = C.CDLL(os.path.join(ORACLE, "_ref", "libkpzref.so"))
    R.ref_dsfmt_draws.argtypes = [C.c_uint32, C.c_int, C.c_int, C.POINTER(C.c_double)]
    R.ref_shuffle4.argtypes = [C.c_uint32, C.c_int, C.c_int, C.POINTER(C.c_int)]
    R.ref_dt_offsets.argtypes = [C.c_int] * 5 + [C.c_uint, C.c_uint, C.POINTER(C.c_int)]
    R.ref_dt_offsets.restype = C.c_int
    R.ref_dt_walk.argtypes = [C.c_int] * 6 + [C.POINTER(C.c_int)]
    R.ref_dt_walk.restype = C.c_int
    R.ref_slcg_skip_fn.argtypes = [C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    R.ref_slcg_skip_table.argtypes = [C.POINTER(C.c_uint64)]
    R.ref_slcg_skip_tmpl_small.argtypes = [C.POINTER(C.c_uint64)]
    R.ref_slcg_gen.argtypes = [C.c_uint64]
    R.ref_slcg_gen.restype = C.c_uint64

    # dSFMT: draws after init_gen_rand(seed) and a number of discards
    dsfmt = []
    buf = (C.c_double * 16)()
    for seed in [42, 1, 0, 12345, 4294967295, 2718281828]:
        for disc in [0, 1, 381, 382, 383, 1000, 100000]:
            R.ref_dsfmt_draws(seed, disc, 16, buf)
            dsfmt.append({"seed": seed, "discard": disc, "draws": [float(v).hex() for v in buf]})

    # Shuffle<4> sequences from a reSeed'ed master
    shuffles = []
    out = (C.c_int * 4 * 200)()
    for seed in [42, 7, 123456789]:
        R.ref_shuffle4(seed, 100000, 200, C.cast(out, C.POINTER(C.c_int)))
        flat = [out[i][j] for i in range(200) for j in range(4)]
        shuffles.append({"seed": seed, "discard": 100000, "perms": [flat[4 * i:4 * i + 4] for i in range(200)]})

    # DoubleTilingDD<2> block offsets
    dt = []
    obuf = (C.c_int * (2 * 4096))()
    for (lx, ly, lbx, lby) in [(9, 9, 4, 4), (6, 5, 2, 2), (8, 6, 3, 2), (10, 8, 4, 3), (7, 9, 5, 2)]:
        for ds in range(4):
            for sh in [(0, 0), (3, 1), (15, 15), (1, 2)]:
                nb = R.ref_dt_offsets(lx, ly, lbx, lby, ds, sh[0], sh[1], obuf)
                dt.append({"lx": lx, "ly": ly, "lbx": lbx, "lby": lby, "ds": ds, "shift": list(sh),
                           "offsets": [[obuf[2 * i], obuf[2 * i + 1]] for i in range(nb)]})
    walks = []
    wbuf = (C.c_int * 4096)()
    for (lx, ly, lbx, lby, T) in [(9, 9, 4, 4, 3), (9, 9, 4, 4, 7), (8, 6, 3, 2, 5)]:
        for wid in range(T):
            n = R.ref_dt_walk(lx, ly, lbx, lby, wid, T, wbuf)
            walks.append({"lx": lx, "ly": ly, "lbx": lbx, "lby": lby, "T": T, "id": wid,
                          "blocks": [wbuf[i] for i in range(n)]})

    # SLCG constants
    tab = (C.c_uint64 * 12)()
    R.ref_slcg_skip_table(tab)
    small = (C.c_uint64 * 16)()
    R.ref_slcg_skip_tmpl_small(small)
    fn = []
    for n in [1, 2, 3, 7, 100, 1023, 30721]:
        a, c = C.c_uint64(), C.c_uint64()
        R.ref_slcg_skip_fn(n, C.byref(a), C.byref(c))
        fn.append({"S": n, "a": a.value, "c": c.value})
    gen = []
    s = 42
    for _ in range(8):
        t = R.ref_slcg_gen(s)
        gen.append({"in": s, "out": t})
        s = t
    slcg = {
        "table": [{"S": S, "a": tab[2 * k], "c": tab[2 * k + 1]}
                  for k, S in enumerate([3840, 7680, 15360, 30720, 61440, 30721])],
        "template_small": [{"S": S, "a": small[2 * k], "c": small[2 * k + 1]}
                           for k, S in enumerate([1, 2, 3, 5, 8, 17, 64, 101])],
        "function": fn,
        "base_step": gen,
    }

    # TinyMT32: the vendored generator, and the CPU schedulers' stream
    # interface (randomize from a reSeed'ed master, random(), randomFloat())
    R.ref_tmt.argtypes = [C.c_uint32] * 4 + [C.c_int, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                             C.POINTER(C.c_float)]
    R.ref_tmt_iface.argtypes = [C.c_uint32, C.c_int, C.c_int] + [C.POINTER(C.c_uint32)] * 3 + \
        [C.c_int, C.POINTER(C.c_uint32), C.POINTER(C.c_float)]
    params = [(0x8F7011EE, 0xFC78FF1F, 0x3793FDFF), (0, 0, 0), (0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF),
              (0x12345678, 0x9ABCDEF0, 0x0F1E2D3C), (0xDEADBEEF, 0x01234567, 0x89ABCDEF)]
    gens = []
    n = 64
    st, u, f = (C.c_uint32 * 4)(), (C.c_uint32 * n)(), (C.c_float * n)()
    for (m1, m2, tm) in params:
        for seed in [1, 0, 42, 4294967295, 2718281828]:
            R.ref_tmt(m1, m2, tm, seed, n, st, u, f)
            gens.append({"mat1": m1, "mat2": m2, "tmat": tm, "seed": seed, "init": list(st),
                         "u32": list(u), "float": [float(v).hex() for v in f]})
    iface = []
    T, n2 = 6, 8
    import random
    rng = random.Random(7)
    for seed in [42, 3]:
        m1 = [rng.getrandbits(32) for _ in range(T)]
        m2 = [rng.getrandbits(32) for _ in range(T)]
        tm = [rng.getrandbits(32) for _ in range(T)]
        a1, a2, a3 = ((C.c_uint32 * T)(*v) for v in (m1, m2, tm))
        u2, f2 = (C.c_uint32 * (T * n2))(), (C.c_float * (T * n2))()
        R.ref_tmt_iface(seed, 100000, T, a1, a2, a3, n2, u2, f2)
        iface.append({"seed": seed, "discard": 100000, "mat1": m1, "mat2": m2, "tmat": tm, "n": n2,
                      "random": list(u2), "randomFloat": [float(v).hex() for v in f2]})

    def dump(name, obj):
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(obj, f, separators=(",", ":"))

    dump("dsfmt.json", dsfmt)
    dump("shuffle4.json", shuffles)
    dump("dt_geometry.json", {"offsets": dt, "walks": walks})
    dump("slcg.json", slcg)
    dump("tinymt32.json", {"generators": gens, "interface": iface})


if __name__ == "__main__":
    main()
