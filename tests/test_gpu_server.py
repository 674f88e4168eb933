"""Serving loop: jobs that arrive while others run are admitted as memory
frees up, share launches, and each result (full exact w^2(t) series and the
final lattice) equals the same job run alone on a stand-alone scheduler."""
import numpy as np
import pytest

import cudakpz_amd as K

pytestmark = pytest.mark.gpu

CLS = {K.KPZ_DT: K.SchedulerDT, K.KPZ_DTR: K.SchedulerDTr, K.KPZ_SCABIT: K.SchedulerSCABit}


def alone_series(spec, times):
    kind, lx, ly, mcs, p, q, seed = spec
    s = CLS[kind](lx, ly)
    s.supplySeed(seed)
    s.initHomogeneous()
    kpz_dt = kind != K.KPZ_SCABIT and (p, q) == (1.0, 0.0)
    if not kpz_dt:
        s.setPQ(p, q)
    out, t = [], 0
    for m in times:
        if kpz_dt:
            s.mcs(m - t)
        else:
            s.mcsPQSyncRng(m - t)
        t = m
        r = s.roughness()
        out.append((m, r.sum_h, r.sum_h2))
    lat = s.popSystem()
    s.close()
    return out, lat


JOBS = [
    (K.KPZ_SCABIT, 8, 8, 40, 0.8, 0.2, 11),
    (K.KPZ_DT, 7, 7, 25, 1.0, 0.0, 12),
    (K.KPZ_DTR, 8, 6, 33, 0.6, 0.4, 13),
    (K.KPZ_SCABIT, 9, 7, 17, 0.5, 0.0, 14),
    (K.KPZ_DT, 9, 9, 60, 0.75, 0.25, 15),
    (K.KPZ_SCABIT, 6, 6, 90, 1.0, 0.0, 16),
    (K.KPZ_SCABIT, 10, 8, 12, 0.55, 0.45, 17),
    (K.KPZ_DTR, 6, 6, 50, 1.0, 0.0, 18),
    (K.KPZ_SCABIT, 13, 6, 20, 0.7, 0.3, 19),
    (K.KPZ_DT, 12, 5, 15, 1.0, 0.0, 20),
]


@pytest.mark.parametrize("budget,slots,tick", [(64 << 20, 16, 16), (1 << 20, 3, 4)])
def test_server_results_equal_standalone(budget, slots, tick):
    srv = K.Server(hbm_budget=budget, max_slots=slots, max_tick_mcs=tick)
    ids = {}
    pending = list(JOBS)
    results = {}
    steps = 0
    # a few jobs up front, the rest arrive while others are running
    for spec in pending[:3]:
        ids[srv.submit(*spec, keep_lattice=True)] = spec
    pending = pending[3:]
    while True:
        left = srv.step()
        steps += 1
        if pending and steps % 2 == 0:
            spec = pending.pop(0)
            ids[srv.submit(*spec, keep_lattice=True)] = spec
            left += 1
        while True:
            r = srv.pop()
            if r is None:
                break
            results[r.id] = r
        if left == 0 and not pending:
            break
        assert steps < 10000
    assert len(results) == len(JOBS)
    for i, res in results.items():
        spec = ids[i]
        times = [m for m, _ in res.series]
        assert times[-1] == spec[3]
        exp, lat = alone_series(spec, times)
        got = [(m, r.sum_h, r.sum_h2) for m, r in res.series]
        assert got == exp, spec
        assert np.array_equal(res.lattice, lat), spec
        assert res.t_submit <= res.t_start <= res.t_end
    st = srv.stats()
    assert st["queued"] == 0 and st["active"] == 0 and st["hbm_used"] == 0
    gpu_s, span_s = srv.busy()
    assert 0 < gpu_s <= span_s * 1.001 + 1e-4


def test_server_rejects_oversized_job():
    srv = K.Server(hbm_budget=1 << 20, max_slots=2)
    with pytest.raises(K.KpzError):
        srv.submit(K.KPZ_SCABIT, 13, 13, 10, 0.5, 0.5, 1)
