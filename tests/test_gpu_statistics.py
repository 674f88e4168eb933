"""Statistical checks of the HIP engine at sizes the oracle cannot reach in a
test: the interface width of the 2+1d KPZ octahedron model must grow with the
KPZ growth exponent, and the p = q (Edwards-Wilkinson) point must not.

Bit-exact parity with the oracle is covered by the parity tests; these tests
pin the physics the reference reports (kpz/roughness.cpp, w^2 against MCS):
beta = 0.2415(15) for 2+1d KPZ from large-scale DT runs (Kelling & Odor,
PRE 84, 061150 (2011)); at t = 10^2..10^3 MCS the effective exponent is still
approaching it from below.  2+1d EW growth is logarithmic (beta = 0)."""
import math
import os

import numpy as np
import pytest

import cudakpz_amd as K
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TIMES = [100, 200, 400, 800, 1600]


def width_series(cls, l, seeds, pq=None, **kw):
    out = []
    for seed in seeds:
        s = cls(l, l, **kw)
        s.supplySeed(seed)
        s.initHomogeneous()
        if pq is not None:
            s.setPQ(*pq)
        done, w2 = 0, []
        for t in TIMES:
            if pq is None:
                s.mcs(t - done)
            else:
                s.mcsPQSyncRng(t - done)
            done = t
            w2.append(s.roughness().w2)
        out.append(w2)
    return np.array(out)


def beta_fit(w2):
    m = w2.mean(0)
    return 0.5 * np.polyfit(np.log(TIMES), np.log(m), 1)[0]


@pytest.mark.parametrize("cls,kw", [(K.SchedulerDT, {}), (K.SchedulerDTr, {}),
                                    (K.SchedulerSCABit, {})])
def test_kpz_growth_exponent(cls, kw):
    w2 = width_series(cls, 12, [1, 2, 3, 4], **kw)
    beta = beta_fit(w2)
    print(f"{cls.__name__}: w2(t) {w2.mean(0)} beta_eff {beta:.4f}")
    assert np.all(np.diff(w2.mean(0)) > 0)
    assert 0.20 < beta < 0.26


def test_kpz_growth_exponent_sca_pq():
    # p = 0.95, q = 0.05 is still KPZ (the --pq option, kpz/kpzFrontend.cpp:536)
    w2 = width_series(K.SchedulerSCABit, 12, [5, 6], pq=(0.95, 0.05))
    beta = beta_fit(w2)
    print(f"SCABit pq .95/.05: beta_eff {beta:.4f}")
    assert 0.18 < beta < 0.26


def minimum_density(lat, lx, ly):
    """Density of local height minima; with p = 1 every attempt at one
    succeeds and raises h by 2, so the growth velocity is 2 * density."""
    h = O.heights(lat, O.ENC_LOCAL, lx, ly).astype(np.int64)
    lower = np.ones_like(h, dtype=bool)
    for ax in (0, 1):
        for d in (1, -1):
            lower &= h < np.roll(h, d, axis=ax)
    return lower.mean()


def test_gpu_and_oracle_ensembles_agree():
    # 16 GPU runs and 16 oracle runs on disjoint seeds: mean w^2(t), the
    # effective beta and the growth velocity agree within seed-to-seed error
    l, T, times = 9, 256, [25, 50, 100, 200, 400]
    nthreads = max(1, min(16, os.cpu_count() or 1))

    def gpu(seed):
        s = K.SchedulerDTr(l, l, threads=T)
        s.supplySeed(seed)
        s.initHomogeneous()
        done, w2 = 0, []
        for t in times:
            s.mcs(t - done)
            done = t
            w2.append(s.roughness().w2)
        return w2, minimum_density(s.popSystem(), l, l)

    def cpu(seed):
        m = O.Dsfmt(O.seed32(seed))
        st = O.Streams(T, m)
        lat = O.flat(l, l, O.ENC_LOCAL)
        done, w2 = 0, []
        for t in times:
            O.dt_mcs_threaded(lat, l, l, 4, 4, st, m, t - done, nthreads, True)
            done = t
            s1, s2 = O.roughness(lat, O.ENC_LOCAL, l, l)
            N = 1 << (2 * l)
            w2.append(s2 / N - (s1 / N) ** 2)
        return w2, minimum_density(lat, l, l)

    g = [gpu(s) for s in range(1, 17)]
    c = [cpu(s) for s in range(1001, 1017)]
    for name, a, b in (("w2", np.array([x[0] for x in g]), np.array([x[0] for x in c])),
                       ("rho_min", np.array([x[1] for x in g]), np.array([x[1] for x in c]))):
        err = np.sqrt(a.var(0, ddof=1) / len(a) + b.var(0, ddof=1) / len(b))
        z = np.abs(a.mean(0) - b.mean(0)) / err
        print(f"{name}: gpu {a.mean(0)} oracle {b.mean(0)} z {z}")
        assert np.all(z < 5)
    bg = 0.5 * np.polyfit(np.log(times), np.log(np.mean([x[0] for x in g], 0)), 1)[0]
    bc = 0.5 * np.polyfit(np.log(times), np.log(np.mean([x[0] for x in c], 0)), 1)[0]
    print(f"beta_eff gpu {bg:.4f} oracle {bc:.4f}, velocity {2 * np.mean([x[1] for x in g]):.4f}")
    assert abs(bg - bc) < 0.02
    assert 0.17 < bg < 0.26


def test_edwards_wilkinson_point_is_logarithmic():
    w2 = width_series(K.SchedulerDT, 12, [7, 8], pq=(0.5, 0.5))
    m = w2.mean(0)
    beta = beta_fit(w2)
    # logarithmic growth: w^2 rises by a near-constant amount per doubling of t
    inc = np.diff(m)
    print(f"DT pq .5/.5: w2 {m} beta_eff {beta:.4f}")
    assert beta < 0.12
    assert np.all(inc > 0) and inc.max() < 2.5 * inc.min()
    assert math.isfinite(beta)
