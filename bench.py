"""Benchmark of the MI355X KPZ engine on its headline workload.

Workload (one "step" = one Monte-Carlo step, N = 65536^2 site updates):
random-sequential double tiling (SchedulerDT semantics, static origin,
16x16-site blocks, one SLCG stream per block = 2^22 streams), flat initial
condition, p=1, q=0, with exact w^2 sampled at the log-spaced times
ceil(exp(log(t+10)+0.1)) inside the timed region.  The lattice (1 GiB,
ENC_LOCAL) and all RNG state stay resident in HBM.

Prints ONE JSON line (rank 0, N=1).  Timing uses HIP events recorded on the
engine's own stream.  The dominant kernel's average launch duration gives
roofline.achieved; cpu_baseline times the oracle (sequential C restatement of
SchedulerCPUDT) on a bounded sample of the same workload.

    python bench.py [--steps K] [--warmup W] [--lx 16] [--no-cpu] [--secondary]
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "site updates/s (DT, SCABit @65536²); batched replica-MCS/s vs serial"
HBM_PEAK_GBS = 8000.0
VALU_SIMDS = 256 * 4       # MI355X: 256 CUs x 4 SIMDs of 16 lanes
CLOCK_HZ = 2.4e9           # peak engine clock
VALU_PEAK_LANE_OPS = VALU_SIMDS * 16 * CLOCK_HZ


class HipEvents:
    """hipEvent timing on an explicit stream (the engine's, not torch's)."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.h.hipEventSynchronize.argtypes = [C.c_void_p]
        self.h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.h.hipEventDestroy.argtypes = [C.c_void_p]

    def event(self):
        e = C.c_void_p()
        assert self.h.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.h.hipEventRecord(e, C.c_void_p(stream)) == 0

    def ms(self, a, b):
        assert self.h.hipEventSynchronize(b) == 0
        t = C.c_float()
        assert self.h.hipEventElapsedTime(C.byref(t), a, b) == 0
        return float(t.value)


def sample_times(total):
    """kpz/sampling.h:84-100 default: next = ceil(exp(log(t+10)+0.1))"""
    out, t = [], 0
    while True:
        t = int(math.ceil(math.exp(math.log(t + 10) + 0.1)))
        if t > total:
            return out
        out.append(t)


def run_dt(K, args, ev):
    s = K.SchedulerDT(args.lx, args.lx, layout=(4, 4))
    s.supplySeed(args.seed)
    s.initHomogeneous()
    stream = s.stream
    s.mcs(args.warmup)
    s.sync()
    s.initHomogeneous()  # measure from the flat state
    samples = set(sample_times(args.steps))
    e0, e1 = ev.event(), ev.event()
    a, b = ev.event(), ev.event()
    mcs_ms = 0.0
    w2 = []
    ev.record(e0, stream)
    done = 0
    while done < args.steps:
        nxt = min([t for t in samples if t > done] + [args.steps])
        ev.record(a, stream)
        s.mcs(nxt - done)
        ev.record(b, stream)
        mcs_ms += ev.ms(a, b)
        done = nxt
        if done in samples:
            w2.append((done, s.roughness().w2))
    ev.record(e1, stream)
    total_ms = ev.ms(e0, e1)
    return s, total_ms, mcs_ms, w2


def run_reductions(K, lx, seed, reps=5):
    """Exact w2 and autocorrelator samples on the lattice of a run (the C2 DT
    lattice, ENC_LOCAL, and the C3 SCABit lattice, ENC_CBBIT) after a few MCS
    with a tag taken: the best wall time of `reps` calls of roughness() /
    correlate() (each queues the reductions, copies 7 int64 back and waits).
    Algorithmic bytes: for w2 the x slopes, all of an ENC_LOCAL lattice's
    words (N/4 bytes: x and y bits interleave) but only the two x quadrants of
    an ENC_CBBIT one (N/8); for the correlator every bit of lattice and
    snapshot (N/2: its tcs term compares the y slopes too)."""
    N = 1 << (2 * lx)
    out = {}
    for name, cls, pq in (("dt_enc_local", K.SchedulerDT, None), ("scabit_enc_cbbit", K.SchedulerSCABit, (0.95, 0.05))):
        s = cls(lx, lx)
        s.supplySeed(seed)
        s.initHomogeneous()
        run = s.mcs
        if pq:
            s.setPQ(*pq)
            run = s.mcsPQSyncRng
        run(2)
        s.correlateTag()
        run(2)
        s.sync()
        r = {}
        w2bytes = N // 8 if cls is K.SchedulerSCABit else N // 4
        for what, fn, nbytes in (("w2", s.roughness, w2bytes), ("correlate", s.correlate, N // 2)):
            best = float("inf")
            for _ in range(reps):
                t0 = time.perf_counter()
                fn()
                best = min(best, time.perf_counter() - t0)
            r[what] = {"ms": best * 1e3, "hbm_GBs_algorithmic": nbytes / best / 1e9,
                       "hbm_frac": nbytes / best / 1e9 / HBM_PEAK_GBS}
        out[name] = r
        s.close()
    return out


def run_scabit(K, lx, p, q, steps, seed, ev, cls=None, kpz=False, rng=0, threads=0):
    """steps MCS of mcsPQSyncRng (kpz: mcs) after 2 warm-up MCS, hipEvent time.
    rng = TinyMT32: the engine's default worker count, parameter set t of
    stream t = random words (the reference's table is not in its tree)"""
    s = (cls or K.SchedulerSCABit)(lx, lx, rng=rng, threads=threads)
    if rng == K.RNG_TINYMT32:
        r = np.random.default_rng(seed)
        s.setTinyMTParams(*(r.integers(0, 1 << 32, s.threads, dtype=np.uint64).astype(np.uint32)
                            for _ in range(3)))
    s.supplySeed(seed)
    s.initHomogeneous()
    s.setPQ(p, q)
    run = s.mcs if kpz else s.mcsPQSyncRng
    run(2)
    a, b = ev.event(), ev.event()
    ev.record(a, s.stream)
    run(steps)
    ev.record(b, s.stream)
    ms = ev.ms(a, b)
    s.close()
    return ms


def batch_mix(n=4096, base_seed=1000, umin=6, umax=11, rng_seed=4096):
    """4096 replicas, L = 2^u with u uniform in {6..11}, half DT / half SCABit,
    half KPZ (p=1, q=0) / half EW (p=q=.5), seed_i = base + i."""
    import random
    import cudakpz_amd as K
    r = random.Random(rng_seed)
    reps = []
    for i in range(n):
        u = r.randint(umin, umax)
        kind = K.KPZ_DT if i % 2 == 0 else K.KPZ_SCABIT
        p, q = (1.0, 0.0) if (i // 2) % 2 == 0 else (0.5, 0.5)
        reps.append((kind, u, u, p, q, base_seed + i))
    return reps


def run_batch_bench(K, ev, mcs, serial_mcs):
    reps = batch_mix()
    t0 = time.perf_counter()
    tuning, tune_ms = K.Batch.autotune(reps, mcs=4)
    tune_s = time.perf_counter() - t0
    # serial: one stand-alone scheduler per replica, one lattice per launch
    cls = {K.KPZ_DT: K.SchedulerDT, K.KPZ_SCABIT: K.SchedulerSCABit}
    alone = []
    for kind, lx, ly, p, q, seed in reps:
        s = cls[kind](lx, ly)
        s.supplySeed(seed)
        s.initHomogeneous()
        if not (kind == K.KPZ_DT and (p, q) == (1.0, 0.0)):
            s.setPQ(p, q)
        alone.append(s)
    for s in alone:
        s.sync()
    t0 = time.perf_counter()
    for (kind, lx, ly, p, q, seed), s in zip(reps, alone):
        if kind == K.KPZ_DT and (p, q) == (1.0, 0.0):
            s.mcs(serial_mcs)
        else:
            s.mcsPQSyncRng(serial_mcs)
        s.sync()
    serial_s = time.perf_counter() - t0
    # batched: same replicas, identical results after serial_mcs
    b = K.Batch(reps, tuning)
    b.initHomogeneous()
    b.mcs(serial_mcs)
    b.sync()
    same = sum(int(np.array_equal(b.popSystem(i), s.popSystem())) for i, s in enumerate(alone))
    for s in alone:
        s.close()
    del alone
    samples = set(sample_times(mcs))
    e0, e1 = ev.event(), ev.event()
    ev.record(e0, b.stream)
    done, w2 = 0, []
    while done < mcs:
        nxt = min([t for t in samples if t > done] + [mcs])
        b.mcs(nxt - done)
        done = nxt
        if done in samples:
            rs = b.roughness()
            w2.append((done + serial_mcs, float(np.mean([r.w2 for r in rs]))))
    ev.record(e1, b.stream)
    batch_ms = ev.ms(e0, e1)
    sites = b.sites
    b.close()
    n = len(reps)
    batch_rate = n * mcs / (batch_ms * 1e-3)
    serial_rate = n * serial_mcs / serial_s
    return {"replicas": n, "sites": sites, "mcs": mcs, "replica_mcs_per_s": batch_rate,
            "site_updates_per_s": sites * mcs / (batch_ms * 1e-3),
            "serial_replica_mcs_per_s": serial_rate, "serial_mcs": serial_mcs,
            "speedup_vs_serial": batch_rate / serial_rate,
            "identical_to_standalone": f"{same}/{n}", "tuning": tuning.as_dict(),
            "autotune_s": tune_s, "ensemble_w2_final": w2[-1][1] if w2 else None,
            "ensemble_w2_samples": len(w2)}


def run_serve_bench(K, window_s, rate=200.0, seed=5, kind_name="scaBit", budget=64 << 30):
    """Poisson stream of jobs (rate/s for window_s seconds): L = 2^u, u in
    {8..13}; MCS log-uniform 10^2..10^4; p ~ U[.5, 1], q = 1 - p; random u64
    seeds.  Latency = completion - scheduled arrival."""
    rng = np.random.default_rng(seed)
    kind = {"scaBit": K.KPZ_SCABIT, "dt": K.KPZ_DTR, "dtStatic": K.KPZ_DT}[kind_name]
    gaps = rng.exponential(1.0 / rate, int(rate * window_s * 3) + 10)
    arr = np.cumsum(gaps)
    arr = arr[arr < window_s]
    n = len(arr)
    us = rng.integers(8, 14, n)
    mcs = np.round(np.exp(rng.uniform(np.log(100), np.log(10000), n))).astype(int)
    ps = rng.uniform(0.5, 1.0, n)
    seeds = rng.integers(0, 2 ** 63, n, dtype=np.int64).astype(np.uint64) * np.uint64(2) + np.uint64(1)
    srv = K.Server(hbm_budget=budget, max_slots=1024, max_tick_mcs=16)
    base = srv.now() + 0.05
    arr = arr + base
    due = {}
    lat, work, i, first_done, last_done = [], 0, 0, None, None
    while True:
        now = srv.now()
        while i < n and arr[i] <= now:
            jid = srv.submit(kind, int(us[i]), int(us[i]), int(mcs[i]), float(ps[i]), 1.0 - float(ps[i]), int(seeds[i]))
            due[jid] = arr[i]
            work += int(mcs[i]) << (2 * int(us[i]))
            i += 1
        left = srv.step()
        while True:
            r = srv.pop()
            if r is None:
                break
            lat.append(r.t_end - due.pop(r.id))
            last_done = r.t_end
        if i == n and left == 0:
            break
        if left == 0 and i < n:
            time.sleep(max(0.0, arr[i] - srv.now()))
    st = srv.stats()
    gpu_s, gpu_span = srv.busy()
    srv.close()
    lat = np.array(lat)
    span = last_done - base
    return {"jobs": n, "kind": kind_name, "rate_per_s": rate, "window_s": window_s,
            "offered_site_updates": work, "site_updates_per_s": st["site_updates"] / span,
            "site_updates_per_gpu_s": st["site_updates"] / gpu_s if gpu_s > 0 else None,
            "gpu_busy_frac": gpu_s / gpu_span if gpu_span > 0 else None,
            "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
            "latency_max_s": float(lat.max()), "makespan_s": span,
            "offered_load_vs_capacity": (work / window_s) / (st["site_updates"] / span)}


import numpy as np  # noqa: E402


def l1_layout(lx, ly, lmin=2):
    """Kmc::optimizeForL1 (inc/cacheLayout.h:62-94) for SchedulerCPUDB: shrink
    the larger block side until the block, at 2 bits per site, fits half of
    cpu0's L1 data cache (32 KiB assumed where sysfs does not say)."""
    kib = 32
    try:
        txt = open("/sys/devices/system/cpu/cpu0/cache/index0/size").read().strip()
        kib = int("".join(ch for ch in txt if ch.isdigit()) or 32)
    except OSError:
        pass
    use = int(math.log2(kib)) + 12
    lm = [lx, ly]
    while lm[0] + lm[1] + 1 > use:
        i = 0 if (lm[0] >= lm[1] and lm[0] > lmin) else 1
        if lm[i] <= lmin:
            break
        lm[i] -= 1
    return lm[0], lm[1]


def host_cpu():
    """The host CPU as this run sees it: model, sockets, physical cores and
    logical CPUs of the machine (lscpu), and the CPUs this process may run on
    (its affinity mask, capped by a cgroup CPU quota where one is set)."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = {k.strip(): v.strip() for k, v in (ln.split(":", 1) for ln in out.splitlines() if ":" in ln)}
        info["model"] = kv.get("Model name")
        sockets = int(kv.get("Socket(s)", "0") or 0)
        cores = int(kv.get("Core(s) per socket", "0") or 0)
        info["sockets"] = sockets
        info["physical_cores"] = sockets * cores if sockets and cores else None
        info["threads_per_core"] = int(kv.get("Thread(s) per core", "0") or 0) or None
        info["logical_cpus"] = int(kv.get("CPU(s)", "0") or 0) or None
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    avail = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()
        if q != "max":
            quota = max(1, int(int(q) // int(per)))
    except (OSError, ValueError):
        pass
    info["affinity_cpus"] = avail
    info["cgroup_cpu_quota"] = quota
    info["usable_cpus"] = min(avail, quota) if quota else avail
    return info


def cpu_baseline(lx=12):
    """The oracle's restatements of the reference's threaded CPU schedulers,
    timed on this host with T = every CPU this run may use and with T = 1:
    SchedulerCPUDT (pthreads spawned and joined per domain set, 16x16-site
    blocks, p=1 q=0) and SchedulerCPUCBBit (rows split over T workers, per
    half-sweep; the per-site p=.95 q=.05 path and the bit-parallel p=.5 path),
    each on a 4096^2 flat sample for a bounded time with T SLCG64 streams;
    plus, on one core, the reference's CPU-runnable 512^2 configuration, the
    plain random-sequential SchedulerCPU and SchedulerCPUDB."""
    from oracle import oracle as O
    cpu = host_cpu()
    cores = max(1, cpu["usable_cpus"])

    def timed(run, enc, nthreads, budget, L=lx):
        m = O.Dsfmt(42)
        st = O.Streams(nthreads, m)
        lat = O.flat(L, L, enc)
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < budget:
            run(lat, L, st, m, nthreads)
            n += 1
        dt = time.perf_counter() - t0
        return {"site_updates_per_s": n * (1 << (2 * L)) / dt, "mcs": n, "seconds": dt, "threads": nthreads,
                "lattice": f"{1 << L}x{1 << L}"}

    def dt_run(lat, L, st, m, T):
        O.dt_mcs_threaded(lat, L, L, 4, 4, st, m, 1, T)

    def cb_pq(lat, L, st, m, T):
        O.cbbit_mcs_threaded(lat, L, L, st, 1, 1, 0.95, 0.05, T)

    def cb_vl3(lat, L, st, m, T):
        O.cbbit_mcs_threaded(lat, L, L, st, 1, 0, 0.5, 0.0, T)

    dt_all = timed(dt_run, O.ENC_LOCAL, cores, 8.0)
    dt_one = timed(dt_run, O.ENC_LOCAL, 1, 4.0)
    pq_all = timed(cb_pq, O.ENC_CBBIT, cores, 5.0)
    pq_one = timed(cb_pq, O.ENC_CBBIT, 1, 3.0)
    # the bit-parallel path streams the lattice: a 16384^2 sample (64 MiB)
    # keeps it out of the caches, as the 1 GiB benchmark lattice is
    vl3_all = timed(cb_vl3, O.ENC_CBBIT, cores, 2.0, lx + 2)
    vl3_one = timed(cb_vl3, O.ENC_CBBIT, 1, 2.0, lx + 2)

    # the reference's CPU-runnable configuration: SchedulerCPUDT on 512^2,
    # 16x16 blocks, one worker, seed 42, 100 MCS, exact w^2 after every MCS
    m = O.Dsfmt(42)
    st = O.Streams(1, m)
    lat = O.flat(9, 9, O.ENC_LOCAL)
    t0 = time.perf_counter()
    for _ in range(100):
        O.dt_mcs(lat, 9, 9, 4, 4, st, m, 1)
        O.roughness(lat, O.ENC_LOCAL, 9, 9)
    c1 = time.perf_counter() - t0

    # plain random-sequential SchedulerCPU (master draws only) on 512^2
    m = O.Dsfmt(42)
    lat = O.flat(9, 9, O.ENC_LOCAL)
    nrs, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < 2.0:
        O.rs_mcs(lat, 9, 9, m, 1, O.RS_MCS)
        nrs += 1
    rs = time.perf_counter() - t0

    # SchedulerCPUDB (dead border, dynamic origin) on 512^2 with the block
    # layout its layoutFromL1 picks on this host (inc/cacheLayout.h:84-94),
    # as many workers as it has blocks (at most `cores`)
    lbx, lby = l1_layout(9, 9)
    tdb = max(1, min(cores, 1 << (18 - lbx - lby)))
    m = O.Dsfmt(42)
    st = O.Streams(tdb, m)
    lat = O.flat(9, 9, O.ENC_LOCAL)
    ndb, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < 2.0:
        O.db_mcs(lat, 9, 9, lbx, lby, st, m, 1, 0, tdb)
        ndb += 1
    tdb_s = time.perf_counter() - t0
    db_upd = ((1 << lbx) - 1) * ((1 << lby) - 1) << (18 - lbx - lby)
    side = f"{1 << lx}x{1 << lx}"
    return {"value": dt_all["site_updates_per_s"], "unit": "site updates/s", "cores": cores, "kind": "port",
            "value_1core": dt_one["site_updates_per_s"],
            "host": cpu,
            "dt": {"all_cores": dt_all, "one_core": dt_one,
                   "workload": f"SchedulerCPUDT {side}, 16x16 blocks, p=1 q=0, T streams = T threads"},
            "cbbit_pq": {"all_cores": pq_all, "one_core": pq_one,
                         "workload": f"SchedulerCPUCBBit {side} ENC_CBBIT, per-site p=.95 q=.05, T row workers"},
            "cbbit_vl3": {"all_cores": vl3_all, "one_core": vl3_one,
                          "workload": f"SchedulerCPUCBBit {vl3_all['lattice']} ENC_CBBIT, bit-parallel p=.5 q=0, "
                                      f"T row workers"},
            "c1_512_dt_1core": {"site_updates_per_s": 100 * (1 << 18) / c1, "seconds": c1,
                                "workload": "SchedulerCPUDT 512x512, 16x16 blocks, T=1, 100 MCS, w2 every MCS"},
            "plain_rs_512_1core": {"site_updates_per_s": nrs * (1 << 18) / rs,
                                   "workload": "SchedulerCPU random-sequential, 512x512, master dSFMT draws"},
            "db_512": {"site_updates_per_s": ndb * db_upd / tdb_s, "threads": tdb,
                       "workload": f"SchedulerCPUDB 512x512, {1 << lbx}x{1 << lby} blocks (layoutFromL1 on this "
                                   f"host), dead border, dynamic origin, T={tdb}"},
            "sample": f"oracle restatements on flat lattices, T={cores} threads (the CPUs this run may use) "
                      f"and T=1, each for a bounded time: SchedulerCPUDT {side} {dt_all['mcs']} + {dt_one['mcs']} "
                      f"MCS, SchedulerCPUCBBit p=.95/q=.05 {side} {pq_all['mcs']} + {pq_one['mcs']} MCS, VL3 "
                      f"{vl3_all['lattice']} {vl3_all['mcs']} + {vl3_one['mcs']} MCS"}


def kernel_roofline(K, key, avg_launch_ms, algo_bytes, fast, slow, mix=None):
    """Roofline object of one kernel.  The kernels here are bound by integer
    VALU issue, not HBM, so the binding roofline is VALU: achieved = the
    lane operations of one launch (rocprof SQ_INSTS_VALU per launch x 64
    lanes, profiles/traffic.json) / the live average launch time; peak = the
    highest issue rate of any VALU instruction measured on this device in this
    run (v_xor_b32 class, kpz_valu_peak), which no instruction mix can
    exceed.  Beside it: the measured rate of the kernel's own instruction mix
    (kpz_valu_mix_rate; a fast-class instruction next to slow-class ones
    issues at about the slow rate, so this is the bound the kernel can
    reach), the HBM roofline (algorithmic bytes / launch time vs 8 TB/s) and
    the counted HBM traffic (FETCH_SIZE + WRITE_SIZE, calibrated)."""
    prof = {}
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
        prof = json.load(open(tfile)).get(key, {})
    t = avg_launch_ms * 1e-3
    hbm = algo_bytes / t / 1e9
    out = {"bound": "valu", "achieved": None, "peak": fast / 1e12, "unit": "T lane-ops/s", "frac": None,
           "traffic": prof.get("bytes_per_launch"), "kernel": key.split("@")[0], "avg_launch_ms": avg_launch_ms,
           "hbm": {"achieved": hbm, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": hbm / HBM_PEAK_GBS,
                   "bytes_per_launch": algo_bytes},
           "valu_peaks_measured": {"fast_class_T": fast / 1e12, "slow_class_T": slow / 1e12,
                                   "nominal_T": VALU_PEAK_LANE_OPS / 1e12}}
    if prof.get("valu_insts_per_launch"):
        lane_ops = prof["valu_insts_per_launch"] * 64.0
        out["achieved"] = lane_ops / t / 1e12
        out["frac"] = out["achieved"] / out["peak"]
        out["lane_ops_per_launch"] = lane_ops
        if mix:
            # the same 8-instruction mix as the kernel's update loop (slow-class
            # share), measured on this device in this run: a fast-class
            # instruction next to slow-class ones issues at about the slow rate
            out["mix_peak"] = {"slow_of_8": mix[0], "T_lane_ops": mix[1] / 1e12,
                               "frac": out["achieved"] / (mix[1] / 1e12)}
        if prof.get("valu_dual_issue_quads_per_launch") is not None and prof.get("units_per_launch"):
            # VALU issue time of one launch: one quad-cycle (a wave64 instruction
            # on a 16-lane SIMD) per instruction, minus the quad-cycles that
            # issued two (SQ_ACTIVE_INST_VALU2), v_mad_u64_u32 at 1.4 (its
            # measured rate), over the 1024 SIMDs of the 256 CUs at the
            # 2.4 GHz peak engine clock
            quads = (prof["valu_insts_per_launch"] - prof["valu_dual_issue_quads_per_launch"]
                     + 0.4 * prof["units_per_launch"] * prof.get("mad64_per_unit", 0) / 64.0)
            busy = quads * 4.0 / (VALU_SIMDS * CLOCK_HZ)
            out["valu_issue_busy"] = {"quad_cycles_per_launch": quads, "issue_ms": busy * 1e3, "frac": busy / t}
    return out


def run_c3(K, lx, seed, mcs, ev):
    """The autocorrelator configuration end to end: SchedulerSCABit 2^lx x 2^lx,
    flat IC, p = .95, q = .05, `mcs` MCS with correlator tags at MCS 10, 100
    and 1000, and at every sample time (kpz/sampling.h:84-100) and tag time the
    exact w2 and the autocorrelator of every live tag, as the kpzScaling event
    loop does (kpz/kpzScaling.cpp:186-208: mcs to the next event, then
    correlate()), with the device reductions in place of popSystem and the CPU
    correlator.  hipEvent time of the whole loop on the engine's stream."""
    s = K.SchedulerSCABit(lx, lx)
    s.supplySeed(seed)
    s.initHomogeneous()
    s.setPQ(0.95, 0.05)
    tags = [t for t in (10, 100, 1000) if t < mcs]
    events = sorted(set(sample_times(mcs)) | set(tags) | {mcs})
    stream = s.stream
    e0, e1, a, b = ev.event(), ev.event(), ev.event(), ev.event()
    s.sync()
    mcs_ms, ncorr, done, last = 0.0, 0, 0, None
    ev.record(e0, stream)
    for t in events:
        ev.record(a, stream)
        s.mcsPQSyncRng(t - done)
        ev.record(b, stream)
        mcs_ms += ev.ms(a, b)
        done = t
        if t in tags:
            s.correlateTag()
        r = s.roughness()
        res = s.correlateAll()
        ncorr += len(res)
        last = (r, res)
    ev.record(e1, stream)
    total_ms = ev.ms(e0, e1)
    N = 1 << (2 * lx)
    out = {"workload": f"SchedulerSCABit {1 << lx}x{1 << lx} ENC_CBBIT, flat IC, p=.95 q=.05, {mcs} MCS, "
                       f"correlator tags at MCS {', '.join(map(str, tags))}, exact w2 and every live tag's "
                       "autocorrelator at each sample and tag time",
           "events": len(events), "correlator_evaluations": ncorr, "device_ms": total_ms, "mcs_ms": mcs_ms,
           "site_updates_per_s": N * mcs / (total_ms * 1e-3), "analysis_share": 1.0 - mcs_ms / total_ms,
           "w2_final": last[0].w2,
           "c_h_final": {str(m): c.ch for m, c in last[1]}}
    s.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--lx", type=int, default=16)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-secondary", dest="secondary", action="store_false",
                    help="skip the SCABit timings at 65536^2")
    ap.add_argument("--batch", type=int, default=2000, metavar="MCS",
                    help="batched 4096-replica mix for MCS steps vs serial per-replica launches (0: skip)")
    ap.add_argument("--serial-mcs", type=int, default=10, help="MCS of the serial per-replica baseline")
    ap.add_argument("--serve", type=float, default=2.0, metavar="SECONDS",
                    help="serving loop: Poisson arrivals at 200/s for SECONDS (0: skip)")
    ap.add_argument("--c3-mcs", type=int, default=10000, metavar="MCS",
                    help="the autocorrelator configuration end to end for MCS (0: skip)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="K",
                    help="after the run, K lattice copies (1 GiB read + 1 GiB write each) whose "
                         "FETCH_SIZE/WRITE_SIZE calibrate the counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; runs on several GPUs (row stripes of the "
              "lattice with one halo row exchanged per phase, or whole replicas per GPU) are not "
              "part of it", file=sys.stderr)
        sys.exit(2)

    import cudakpz_amd as K
    ev = HipEvents()
    N = 1 << (2 * args.lx)
    s, total_ms, mcs_ms, w2 = run_dt(K, args, ev)
    launches = 4 * args.steps
    avg_launch_ms = mcs_ms / launches
    bytes_per_launch = 0.5 * N / 4  # 0.5 B per site update, N/4 updates per phase
    value = N * args.steps / (total_ms * 1e-3)
    slow, fast = K.valu_peak()
    # both hot loops are 3 of 8 slow-class instructions: DT 8 of 20 per update
    # (v_mad_u64_u32, v_mul_lo_u32, v_add3_u32), per-site SCABit 3 of 8 per
    # site (v_mad_u64_u32, 2 v_mad_u32_u24)
    mix = (3, K.valu_mix_rate(3))
    roof = kernel_roofline(K, f"dt_phase_kernel<4,4,KPZ,static>@{1 << args.lx}", avg_launch_ms,
                           bytes_per_launch, fast, slow, mix)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "site updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"SchedulerDT {1 << args.lx}x{1 << args.lx} ENC_LOCAL, flat IC, p=1 q=0, "
                               f"16x16-site DT blocks, one SLCG64 stream per block ({s.threads} streams), "
                               f"{args.steps} MCS, exact w2 at {len(w2)} log-spaced times",
                   "lattice": f"{1 << args.lx}x{1 << args.lx}", "mcs": args.steps, "parallelism": "single GPU"},
        "roofline": roof,
        "w2_final": w2[-1][1] if w2 else None,
    }
    s.close()
    if args.secondary:
        sec = {}
        for name, p, q, cls in [("scabit_p0.5_q0", 0.5, 0.0, K.SchedulerSCABit),
                                ("scabit_p0.95_q0.05", 0.95, 0.05, K.SchedulerSCABit),
                                ("sca_local_p0.95_q0", 0.95, 0.0, K.SchedulerSCA)]:
            steps = 100 if cls is K.SchedulerSCABit else 20
            ms = run_scabit(K, args.lx, p, q, steps, args.seed, ev, cls)
            ups = N * steps / (ms * 1e-3)
            sec[name] = {"site_updates_per_s": ups, "hbm_GBs_algorithmic": ups * 1.0 / 1e9,
                         "hbm_frac": ups / 1e9 / HBM_PEAK_GBS, "ms_per_mcs": ms / steps}
            if name == "scabit_p0.95_q0.05":
                # the autocorrelator configuration's kernel: one launch per
                # half-sweep, 2 launches per MCS, 2^31 site updates each
                line["roofline_c3"] = kernel_roofline(
                    K, f"cbbit_half_kernel<u64,per-site>@{1 << args.lx}", ms / steps / 2, 1.0 * N / 2, fast, slow,
                    mix)
        # the reference's default "dt" scheduler: DT with a random origin per
        # MCS (SchedulerCPUDTr semantics), p = 1, q = 0, 0.5 B per update
        ms = run_scabit(K, args.lx, 1.0, 0.0, 100, args.seed, ev, K.SchedulerDTr, kpz=True)
        ups = N * 100 / (ms * 1e-3)
        sec["dtr_p1_q0"] = {"site_updates_per_s": ups, "hbm_GBs_algorithmic": ups * 0.5 / 1e9,
                            "hbm_frac": ups * 0.5 / 1e9 / HBM_PEAK_GBS, "ms_per_mcs": ms / 100}
        # TinyMT32 workers, the generator the reference's frontend gives dt, sca
        # and scaBit (kpz/kpzFrontend.cpp:308-331): one lane per worker stream
        # (T = 0: the engine's default worker count; T = 32: a reference CPU
        # run's thread count, whose results the engine reproduces bit for bit
        # with streams jumped ahead to every lane's start)
        for name, cls, kpz, p, q, steps, bpu, T in (
                ("dtr_tinymt32_p1_q0", K.SchedulerDTr, True, 1.0, 0.0, 40, 0.5, 0),
                ("scabit_tinymt32_p0.5_q0", K.SchedulerSCABit, True, 0.5, 0.0, 20, 1.0, 0),
                ("scabit_tinymt32_p0.95_q0.05", K.SchedulerSCABit, False, 0.95, 0.05, 10, 1.0, 0),
                ("sca_tinymt32_p0.95_q0", K.SchedulerSCA, True, 0.95, 0.0, 10, 1.0, 0),
                ("dtr_tinymt32_T32_p1_q0", K.SchedulerDTr, True, 1.0, 0.0, 20, 0.5, 32),
                ("scabit_tinymt32_T32_p0.5_q0", K.SchedulerSCABit, True, 0.5, 0.0, 20, 1.0, 32),
                ("sca_tinymt32_T32_p0.95_q0", K.SchedulerSCA, True, 0.95, 0.0, 10, 1.0, 32)):
            ms = run_scabit(K, args.lx, p, q, steps, args.seed, ev, cls, kpz=kpz, rng=K.RNG_TINYMT32, threads=T)
            ups = N * steps / (ms * 1e-3)
            sec[name] = {"site_updates_per_s": ups, "hbm_GBs_algorithmic": ups * bpu / 1e9,
                         "hbm_frac": ups * bpu / 1e9 / HBM_PEAK_GBS, "ms_per_mcs": ms / steps}
            print(f"# {name}: {ms / steps:.3f} ms/MCS", file=sys.stderr, flush=True)
        line["secondary"] = sec
        line["reductions"] = run_reductions(K, args.lx, args.seed)
    if args.c3_mcs > 0:
        line["c3_end_to_end"] = run_c3(K, args.lx, args.seed, args.c3_mcs, ev)
    if args.batch:
        line["batched"] = run_batch_bench(K, ev, args.batch, args.serial_mcs)
    if args.serve > 0:
        over = run_serve_bench(K, args.serve)
        # a second point below capacity: the same job mix at 0.8x the offered
        # load the engine sustained in the first run
        rate = 200.0 * 0.8 / over["offered_load_vs_capacity"]
        under = run_serve_bench(K, max(args.serve, 4.0), rate=rate, seed=6)
        line["serving"] = dict(over, at_0p8_capacity=under)
    if args.copy_calibration:
        src = K.SchedulerDT(args.lx, args.lx, layout=(4, 4), threads=1)
        dst = K.SchedulerDT(args.lx, args.lx, layout=(4, 4), threads=1)
        src.initHomogeneous()
        for _ in range(args.copy_calibration):
            dst.copySystem(src)
        dst.sync()
        line["copy_calibration"] = {"copies": args.copy_calibration, "bytes_read_per_copy": N // 4,
                                    "bytes_written_per_copy": N // 4, "kernel": "copy_kernel"}
        src.close()
        dst.close()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
