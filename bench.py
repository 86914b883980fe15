#!/usr/bin/env python3
"""Registered points/s of the voxel_plus registration path on one MI355X.

A step is one registered scan, the reference's LIO_MAPPING work per
SyncPackage (lio_builder.cpp:212-247): undistortCloud (IMU propagation of the
filter state over the package's 200 Hz samples and backward compensation of
every point), the IESKF update (<= 5 observe passes of sharedUpdateFunc + the
FP64 filter algebra) and the map update (VoxelMap::update).  The workload is
a seeded synthetic street sequence of 120,000 points per raw scan (0.5 m
voxels, 2,000,000-voxel LRU).  Before the warmup the sequence is registered
from its start until the map is at its capacity and evicting (the steady
state of a long street run); the timed scans follow on.  Raw scans and their
per-point times are copied to HBM before timing; the IMU samples of a package
(about 20) go with each step.  The timed region is the device path with one
host synchronisation per scan (the plain, single-scan form), so the per-scan
wall time is also the latency reported as p50 / p99.

The same JSON line carries the roofline figure of the kernel that takes the
most time per scan (HIP events around every launch on the library's stream,
over scans that follow the timed region) and the FP64 oracle timed on the
host over a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "registered points/s (batched vs plain, same GPU); p50/p99 scan latency"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per unit of work (the per-scan byte model of the path):
# an observe pass moves 76 B per point (xyz 12, hash probe 16, plane record
# 32, residual slot 16); an insertion 64 B per point (xyz 12, hash probe 16,
# retained point + symmetric cov 36); a refit 1,312 B (the voxel's 69 FP64
# moments read and written 2 x 552, its plane record read and written
# 2 x 64, its state 80); a merge() evaluation
# 288 B (6 x (16 B probe + 32 B plane)); an LRU creation or eviction
# decision 32 B (ring entry 16, state and stamp 16).
OBS_BYTES_POINT = 76
INS_BYTES_POINT = 64
REFIT_BYTES = 2 * 552 + 2 * 64 + 80
MERGE_BYTES = 288
LRU_BYTES = 32
# an IESKF iteration reads the observe pass's partial rows (248 B per block
# of 256 x points-per-thread points) and moves P^-1, H_'s factor, the state
# and the prior state (9,500 B); an IMU propagation reads and writes P and
# the state (9,040 B) and per IMU sample reads it (56 B) and writes a pose
# table entry (176 B); a candidate pass reads per touched voxel its record
# (80 B) and its six neighbours' hash entries and fields (6 x 32 B) and
# writes their ids and its dirty time (32 B)
PART_ROW_BYTES = 248
IESKF_ITER_BYTES = 9500
PRED_BYTES = 9040
PRED_SAMPLE_BYTES = 56 + 176
CAND_VOXEL_BYTES = 80 + 6 * 32 + 32
# observe launches timed back to back per event pair in the roofline pass
OBS_TIME_REPS = 8


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity_cores": len(os.sched_getaffinity(0)), "cpu_model": model}


def make_oracle(cfg):
    from oracle import Oracle

    o = Oracle(cfg)
    o.imu_config()
    o.set_scan_resolution(cfg.scan_resolution)
    return o


def cpu_baseline(pkgs, cfg, n_warm, n_timed):
    """FP64 oracle (one host thread), LIOBuilder::process per package, on the
    first packages of the same sequence: IMU_INIT, MAP_INIT, n_warm
    registered scans, then n_timed timed ones."""
    o = make_oracle(cfg)
    t_reg = 0.0
    for k in range(2 + n_warm + n_timed):
        t0 = time.perf_counter()
        o.process(*pkgs[k])
        dt = time.perf_counter() - t0
        if k >= 2 + n_warm:
            t_reg += dt
    n = len(pkgs[0][0])
    return {
        "value": n * n_timed / t_reg,
        "unit": "points/s",
        "cores": 1,
        "kind": "port",
        "sample": f"street {n} pts/scan, packages {2 + n_warm}-{1 + n_warm + n_timed} of the sequence (after IMU_INIT, "
        f"MAP_INIT and {n_warm} registered scans; map {o.map_size()} voxels at the end; {n_timed} timed, "
        f"{t_reg:.1f} s), FP64 oracle incl. undistortion, 1 thread; the timed GPU scans run on the map "
        "after the untimed fill scans (several times larger), so the per-point CPU cost here is a lower bound",
    }


def cpu_baseline_sessions(pkgs, cfg, n_warm, n_timed, threads):
    """FP64 oracle with one independent session per host thread, all threads
    registering at once (the host-CPU form of the batched / serving workload):
    aggregate points/s over the timed scans."""
    import threading

    ready = threading.Barrier(threads + 1)
    done = threading.Barrier(threads + 1)
    errors = []

    def worker():
        try:
            o = make_oracle(cfg)
            for k in range(2 + n_warm):
                o.process(*pkgs[k])
            ready.wait()
            for k in range(2 + n_warm, 2 + n_warm + n_timed):
                o.process(*pkgs[k])
            done.wait()
            o.close()
        except Exception as e:  # reported, never hangs the barrier
            errors.append(repr(e))
            ready.abort()
            done.abort()

    th = [threading.Thread(target=worker) for _ in range(threads)]
    for t in th:
        t.start()
    ready.wait()
    t0 = time.perf_counter()
    done.wait()
    wall = time.perf_counter() - t0
    for t in th:
        t.join()
    if errors:
        raise RuntimeError(errors[0])
    n = len(pkgs[0][0])
    return {
        "value": n * n_timed * threads / wall,
        "unit": "points/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{threads} independent street sessions of {n} pts/scan, one per host thread (all the process's "
        f"cores), packages {2 + n_warm}-{1 + n_warm + n_timed} registered concurrently ({wall:.1f} s), FP64 oracle",
    }


def upload_packages(pkgs, first, dev_dtype="float32"):
    """Points and times of packages [first, len) into one HBM tensor
    [k, 4, n] (x, y, z, t_ms); packages must share n."""
    import torch

    n = len(pkgs[first][0])
    dev = torch.empty((len(pkgs) - first, 4, n), dtype=torch.float32, device="cuda:0")
    for k in range(first, len(pkgs)):
        x, y, z, t = pkgs[k][:4]
        dev[k - first].copy_(torch.from_numpy(np.stack([x, y, z, t])))
    return dev


def run_plain_evicting(ctx, dev, pkgs, N, capacity, F, W, K):
    """The plain form on the same street packages with an LRU small enough
    that the fill leaves it full: every timed scan creates and evicts voxels
    (the LRU-eviction regime of the reference's map, voxel_map.cpp:232-256)."""
    import torch

    from voxelmapplus_fastlio2_amd import LIOConfig, Registrar

    g = Registrar(ctx, LIOConfig(max_scan_points=N, map_capacity=capacity))
    stride = dev.stride(1) * 4

    def bind(k):
        base = dev[k - 1].data_ptr()
        _, _, _, _, imu, cs, ce = pkgs[k]
        g.set_raw_device(base, base + stride, base + 2 * stride, base + 3 * stride, N, imu, cs, ce)

    assert g.imu_init(pkgs[0][4], pkgs[0][6]), "IMU_INIT within the first package"
    bind(1)
    g.build()
    for k in range(2, 2 + F + W):
        bind(k)
        g.register()
    torch.cuda.synchronize()
    size0 = g.map_size()
    lat, evicted, created = [], 0, 0
    t0 = time.perf_counter()
    for k in range(2 + F + W, 2 + F + W + K):
        ts = time.perf_counter()
        bind(k)
        _, m = g.register()
        lat.append(time.perf_counter() - ts)
        evicted += m.evicted
        created += m.created
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    size1 = g.map_size()
    g.close()
    lat_ms = np.array(lat) * 1e3
    out = {
        "value": N * K / elapsed,
        "unit": "points/s",
        "ms_per_step": elapsed / K * 1e3,
        "map_capacity": capacity,
        "map_voxels": [size0, size1],
        "per_scan_mean": {"created": created / K, "evicted": evicted / K},
        "latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99))},
    }
    log(f"plain, LRU evicting ({capacity} voxels): {out['ms_per_step']:.3f} ms/scan, "
        f"{out['per_scan_mean']['evicted']:.0f} evictions per scan")
    return out


def run_batched(ctx, B, W, K, capacity, nmax=120000, groups=1):
    """B independent sequences (room / street alternating), raw scans of
    sizes uniform in [2k, nmax] with their IMU packages, advanced together by
    vmp_register_batch (undistortion, IESKF and map update of every session
    in one set of launches per stage).  With groups > 1 the sessions are
    split round-robin into that many groups, each on its own context (HIP
    stream) and host thread, so one group's latency-bound stages overlap
    another group's bandwidth-bound ones."""
    import threading

    import torch

    from voxelmapplus_fastlio2_amd import (
        Context,
        LIOConfig,
        Registrar,
        register_batch,
        register_batch_submit,
        register_batch_wait,
        set_raw_device_batch,
    )
    from voxelmapplus_fastlio2_amd.synth import ROOM, SEED0, STREET, ImuSequence

    rng = np.random.default_rng(0xBA7C4)
    total = W + K + 3  # IMU_INIT and MAP_INIT packages first, one step for the per-launch timing last
    sizes = rng.integers(2000, nmax + 1, size=(total - 1, B))
    sizes = np.vstack([sizes, rng.integers(2000, nmax + 1, size=(1, B))])
    # the IMU_INIT package is a full sweep, so it holds the imu_init_num
    # samples (a scan of fewer points ends earlier in its sweep)
    sizes[0, :] = nmax
    seqs = [ImuSequence(ROOM if i % 2 == 0 else STREET, SEED0 + 1000 + i, nmax) for i in range(B)]
    t0 = time.perf_counter()
    steps = []
    for k in range(total):
        offs = np.concatenate([[0], np.cumsum(sizes[k])])
        buf = torch.empty((4, int(offs[-1])), dtype=torch.float32)
        meta = []
        for i in range(B):
            x, y, z, t, imu, cs, ce = seqs[i].package(k, sizes[k, i])
            buf[:, offs[i] : offs[i + 1]] = torch.from_numpy(np.stack([x, y, z, t]))
            meta.append((imu, cs, ce))
        steps.append((buf.to("cuda:0"), offs, meta))
    log(f"batched: generated {total} packages x {B} sessions ({int(sizes.sum())} points) in "
        f"{time.perf_counter() - t0:.1f} s")
    cfg = LIOConfig(max_scan_points=nmax, map_capacity=capacity)
    ctxs = [ctx] + [Context(0) for _ in range(groups - 1)]
    # sessions alternate room / street; each group gets both kinds
    gof = [(i // 2) % groups for i in range(B)]
    members = [[i for i in range(B) if gof[i] == gi] for gi in range(groups)]
    regs = [Registrar(ctxs[gof[i]], cfg) for i in range(B)]
    for gi, cx in enumerate(ctxs):  # the largest step's workspace, allocated before timing
        cx.reserve(int(sizes[:, members[gi]].sum(axis=1).max()), len(members[gi]))

    def bind(k, idx):
        dev, offs, meta = steps[k]
        base = dev.data_ptr()
        stride = dev.stride(0) * 4
        idx = np.asarray(list(idx))
        xo = base + offs[idx].astype(np.uint64) * 4
        set_raw_device_batch([regs[i] for i in idx], xo, xo + stride, xo + 2 * stride, xo + 3 * stride,
                             sizes[k, idx], [meta[i][0] for i in idx], [meta[i][1] for i in idx],
                             [meta[i][2] for i in idx])

    for i, g in enumerate(regs):  # IMU_INIT
        imu, cs, ce = steps[0][2][i]
        assert g.imu_init(imu, ce), "one package holds imu_init_num samples"
    for gi in range(groups):  # a batch binding takes the sessions of one context
        bind(1, members[gi])
    for g in regs:  # MAP_INIT: undistortion + build
        g.build()

    def step(k, idx, wait=True):
        rg = [regs[i] for i in idx]
        bind(k, idx)
        if wait:
            return register_batch(rg)
        register_batch_submit(rg)  # launched without waiting: the host prepares the next step meanwhile
        return None

    from voxelmapplus_fastlio2_amd import autotune_observe_batch

    tuned = {}
    for gi in range(groups):
        table, ms = autotune_observe_batch([regs[i] for i in members[gi]], reps=5)
        tuned[f"group{gi}"] = {"points_per_thread": table, "candidates_ms": ms}
    log(f"batched autotune: {tuned}")
    for k in range(2, W + 2):
        for idx in members:
            step(k, idx)
    torch.cuda.synchronize()
    lat = [[] for _ in members]
    last_mst = [None] * groups

    def run_group(gi):
        for k in range(W + 2, W + K + 2):
            ts = time.perf_counter()
            step(k, members[gi], wait=False)
            lat[gi].append(time.perf_counter() - ts)
        _, last_mst[gi] = register_batch_wait(ctxs[gi], len(members[gi]))

    t_start = time.perf_counter()
    if groups == 1:
        run_group(0)
    else:
        th = [threading.Thread(target=run_group, args=(gi,)) for gi in range(groups)]
        for t in th:
            t.start()
        for t in th:
            t.join()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    pts = int(sizes[W + 2 : W + K + 2].sum())
    ev = np.zeros(B, np.int64)
    mg = np.zeros(B, np.int64)
    wn = np.zeros(B, np.int64)
    rq = np.zeros(B, np.int64)
    evi = np.zeros(B, np.int64)
    for gi, idx in enumerate(members):
        for j, i in enumerate(idx):
            m = last_mst[gi][j]
            ev[i], mg[i], wn[i], rq[i], evi[i] = m.merge_evals, m.merges, m.merge_windows, m.merge_requeued, m.evicted
    # one more step of group 0, launched kernel by kernel between the
    # library's HIP events: the batched form's kernels against their byte
    # models (the timed steps replay no graph here either, but run without
    # events)
    lead = regs[members[0][0]]
    kx = W + K + 2
    lead.set_timing(True)
    obs0, map0 = step(kx, members[0])
    torch.cuda.synchronize()
    tt0 = lead.timing_totals()
    kt0 = ctx.kernel_times()
    lead.set_timing(False)
    ppt_tab = tuned["group0"]["points_per_thread"]
    bst = {"points": 0, "raw_points": 0, "refits": 0, "merge_evals": 0, "created": 0, "evicted": 0, "touched": 0,
           "ieskf_bytes": 0, "pred_bytes": 0}
    for j, i in enumerate(members[0]):
        n_i = int(sizes[kx, i])
        ppt_i = ppt_tab[0 if n_i < 8192 else 1 if n_i < 32768 else 2 if n_i < 131072 else 3]
        m, o = map0[j], obs0[j]
        bst["points"] += m.points
        bst["raw_points"] += n_i
        bst["touched"] += m.touched
        bst["ieskf_bytes"] += o.iterations * (PART_ROW_BYTES * -(-m.points // (256 * ppt_i)) + IESKF_ITER_BYTES)
        bst["pred_bytes"] += PRED_BYTES + PRED_SAMPLE_BYTES * len(steps[kx][2][i][0])
        for f in ("refits", "merge_evals", "created", "evicted"):
            bst[f] += getattr(m, f)
    bkr = kernel_rooflines(tt0, kt0, bst, None, 1)
    bdom = max(bkr, key=lambda k: bkr[k]["ms_per_scan"])
    bd = bkr[bdom]
    log(f"batched step, group 0 ({len(members[0])} sessions, {bst['points']} points), per launch: " +
        ", ".join(f"{n} {r['avg_launch_ms'] * 1e3:.0f} us" for n, r in sorted(bkr.items(), key=lambda kv: -kv[1]["ms_per_scan"])))
    h = int(ev.argmax())
    log(f"batched last step: merge evals room max {ev[0::2].max()} mean {ev[0::2].mean():.0f}, "
        f"street max {ev[1::2].max()} mean {ev[1::2].mean():.0f}; merges room mean {mg[0::2].mean():.0f} "
        f"street mean {mg[1::2].mean():.1f}; heaviest session: {ev[h]} evals, {mg[h]} merges, {wn[h]} windows, "
        f"{rq[h]} requeued; evictions street mean {evi[1::2].mean():.0f}")
    lat = [x for grp in lat for x in grp]
    for g in regs:
        g.close()
    for c in ctxs[1:]:
        c.close()
    lat_ms = np.array(lat) * 1e3
    return {
        "value": pts / elapsed,
        "unit": "points/s",
        "sessions": B,
        "groups": groups,
        "steps": K,
        "warmup": W,
        "points_per_step_mean": pts / K,
        "ms_per_step": elapsed / K * 1e3,
        "step_host_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99))},
        "submission": "asynchronous (vmp_register_batch_submit): step k+1 is prepared while step k runs",
        "map_capacity_per_session": capacity,
        "evicted_per_street_session_last_step": float(evi[1::2].mean()),
        "roofline": {
            "kernel": bdom,
            "bound": "hbm",
            "achieved": bd["achieved_gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": bd["achieved_gbs"] / HBM_PEAK_GBS,
            "traffic": None,
            "bytes_per_launch": bd["bytes_per_launch"],
            "bytes_model": bd["bytes_model"],
            "avg_launch_ms": bd["avg_launch_ms"],
            "timed_over": f"one more step of group 0 ({len(members[0])} sessions, {bst['points']} points) after "
            "the timed region, HIP events around each launch on the group's stream",
        },
        "kernels": bkr,
        "observe_autotune": tuned,
        "workload": f"{B} sequences room/street alternating, raw scans of U[2000,{nmax}] points with their 200 Hz "
        "IMU packages (undistortion on the device)",
    }


def run_serving(sessions, rate, duration, nmax=120000, nmin=2000, capacity=200000, budget_gb=200, executors=2,
                max_batch_sessions=128, seed=0x5E27E, uniform_scans=4):
    """Continuous batching: `sessions` sequences (room / street alternating),
    Poisson scan arrivals at `rate` scans/s in aggregate for `duration` s,
    scan sizes log-uniform in [nmin, nmax], maps and states resident in HBM
    under `budget_gb`.  Every session's MAP_INIT scan is served first (all
    arriving at once, not timed); the timed trace holds the registered scans.
    A second replay of the same scans with every request arriving at t = 0
    measures the serving loop's saturation throughput.  A third replay at
    saturation feeds every session `uniform_scans` scans with sizes uniform in
    [nmin, nmax], the batched form's distribution, for a like-for-like rate."""
    import torch

    from voxelmapplus_fastlio2_amd import LIOConfig, initial_covariance, state_vector
    from voxelmapplus_fastlio2_amd.serving import Server, make_requests
    from voxelmapplus_fastlio2_amd.synth import QDIAG, ROOM, SEED0, STREET, Sequence

    rng = np.random.default_rng(seed)
    t_arr = np.cumsum(rng.exponential(1.0 / rate, size=int(rate * duration * 1.2) + 16))
    t_arr = t_arr[t_arr < duration]
    who = rng.integers(0, sessions, size=len(t_arr))
    counts = np.bincount(who, minlength=sessions)
    lo, hi = np.log(nmin), np.log(nmax)
    seqs = [Sequence(ROOM if i % 2 == 0 else STREET, SEED0 + 5000 + i, nmax) for i in range(sessions)]
    # per session: scan 0 (MAP_INIT) then one scan per arrival
    sizes = [np.exp(rng.uniform(lo, hi, size=counts[i] + 1)).astype(np.int64) for i in range(sessions)]
    total = int(sum(int(z.sum()) for z in sizes))
    t0 = time.perf_counter()
    host = np.empty((3, total), np.float32)
    where = {}
    off = 0
    for i in range(sessions):
        for k, n in enumerate(sizes[i]):
            x, y, z = seqs[i].scan(k, int(n))
            host[0, off : off + n], host[1, off : off + n], host[2, off : off + n] = x, y, z
            where[(i, k)] = (off, int(n))
            off += int(n)
    dev = torch.from_numpy(host).to("cuda:0")
    del host
    base, stride = dev.data_ptr(), dev.stride(0) * 4
    log(f"serving: generated {total} points ({len(t_arr)} timed scans, {sessions} sessions) in {time.perf_counter() - t0:.1f} s")

    def item(i, k, arrival):
        o, n = where[(i, k)]
        it = {"session": i, "n": n, "x": base + 4 * o, "y": base + stride + 4 * o, "z": base + 2 * stride + 4 * o,
              "arrival": float(arrival)}
        if k > 0:
            it["dR"], it["dp"] = seqs[i].motion(k)
        return it

    k_next = np.ones(sessions, np.int64)
    timed = []
    for t, i in zip(t_arr, who):
        timed.append(item(int(i), int(k_next[i]), t))
        k_next[i] += 1
    cfg = LIOConfig(max_scan_points=nmax, map_capacity=capacity)
    out = {}
    with Server(cfg, sessions, executors=executors, max_batch_sessions=max_batch_sessions,
                hbm_budget=int(budget_gb * (1 << 30))) as sv:
        for mode in ("poisson", "saturation"):
            for i in range(sessions):
                R, t = seqs[i].pose(0)
                sv.session_init(i, state_vector(R, t), initial_covariance())
            init = make_requests([item(i, 0, 0.0) for i in range(sessions)])
            sv.run(init, QDIAG)
            reqs = make_requests(timed if mode == "poisson" else [dict(it, arrival=0.0) for it in timed])
            torch.cuda.synchronize()
            sv.run(reqs, QDIAG)
            st = sv.stats()
            arr = np.array([r.arrival for r in reqs])
            done = np.array([r.done for r in reqs])
            npts = np.array([r.n for r in reqs])
            lat = (done - arr) * 1e3
            span = done.max() - arr.min()
            out[mode] = {
                "value": float(npts.sum() / span),
                "scans": len(reqs),
                "scans_per_s": float(len(reqs) / span),
                "latency_ms": {q: float(np.percentile(lat, p)) for q, p in (("p50", 50), ("p99", 99), ("p999", 99.9))},
                "mean_batch_sessions": float(st["requests"] / max(st["batches"], 1)),
                "batches": st["batches"],
                "evictions": st["evictions"],
                "resident_gb": st["resident_bytes"] / (1 << 30),
            }
            log(f"serving {mode}: {out[mode]}")
        if uniform_scans > 0:
            del dev
            # per session: MAP_INIT then uniform_scans scans, sizes U[nmin, nmax]
            usz = rng.integers(nmin, nmax + 1, size=(sessions, uniform_scans + 1))
            utot = int(usz.sum())
            host = np.empty((3, utot), np.float32)
            uwhere = {}
            off = 0
            for i in range(sessions):
                for k in range(uniform_scans + 1):
                    n = int(usz[i, k])
                    x, y, z = seqs[i].scan(k, n)
                    host[0, off : off + n], host[1, off : off + n], host[2, off : off + n] = x, y, z
                    uwhere[(i, k)] = (off, n)
                    off += n
            dev = torch.from_numpy(host).to("cuda:0")
            del host
            ubase, ustride = dev.data_ptr(), dev.stride(0) * 4

            def uitem(i, k):
                o, n = uwhere[(i, k)]
                it = {"session": i, "n": n, "x": ubase + 4 * o, "y": ubase + ustride + 4 * o,
                      "z": ubase + 2 * ustride + 4 * o, "arrival": 0.0}
                if k > 0:
                    it["dR"], it["dp"] = seqs[i].motion(k)
                return it

            for i in range(sessions):
                R, t = seqs[i].pose(0)
                sv.session_init(i, state_vector(R, t), initial_covariance())
            sv.run(make_requests([uitem(i, 0) for i in range(sessions)]), QDIAG)
            reqs = make_requests([uitem(i, k) for k in range(1, uniform_scans + 1) for i in range(sessions)])
            torch.cuda.synchronize()
            st0 = sv.stats()
            sv.run(reqs, QDIAG)
            st = sv.stats()
            done = np.array([r.done for r in reqs])
            npts = np.array([r.n for r in reqs])
            span = done.max()
            nb = st["batches"] - st0["batches"]
            out["saturation_uniform"] = {
                "value": float(npts.sum() / span),
                "scans": len(reqs),
                "scans_per_s": float(len(reqs) / span),
                "mean_batch_sessions": float((st["requests"] - st0["requests"]) / max(nb, 1)),
                "batches": nb,
                "sizes": f"U[{nmin},{nmax}], {uniform_scans} scans per session after MAP_INIT",
            }
            log(f"serving saturation_uniform: {out['saturation_uniform']}")
    del dev
    offered = {
        "scans_per_s": rate,
        "points_per_s": float(sum(r["n"] for r in timed) / duration),
    }
    return {
        "unit": "points/s",
        "workload": f"{sessions} sessions room/street alternating, Poisson {rate:g} scans/s for {duration:g} s, "
        f"scan points log-U[{nmin},{nmax}], {budget_gb} GB HBM budget, {executors} executor streams",
        "offered": offered,
        "poisson": out["poisson"],
        "saturation": out["saturation"],
        "saturation_uniform": out.get("saturation_uniform"),
    }


def load_traffic():
    """Measured HBM bytes per dispatch of each kernel, from the committed
    rocprofv3 PMC summaries (lines "traffic <kernel> <bytes>" in
    profiles/*.md); empty when none is there."""
    import glob
    import re

    out = {}
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        with open(path) as f:
            for line in f:
                m = re.match(r"^traffic (\S+) (\d+(?:\.\d+)?)\s*$", line.strip())
                if m:
                    out[m.group(1)] = (float(m.group(2)), os.path.relpath(path, ROOT))
    return out


def kernel_rooflines(tt, kt, stats, N, scans):
    """Per-kernel figures over the roofline scans: time per scan, launches,
    algorithmic bytes per launch (the model above) and achieved GB/s."""
    rows = {}
    if tt["obs_launches"] > 0:
        b = OBS_BYTES_POINT * tt["obs_points"] / tt["obs_launches"]
        rows["observe_kernel<6>"] = (tt["obs_ms"], tt["obs_launches"], b, f"{OBS_BYTES_POINT} B x points")
    units = {
        "k_ieskf_step": (stats["ieskf_bytes"], f"{PART_ROW_BYTES} B x observe blocks + {IESKF_ITER_BYTES} B, "
                         "per iteration"),
        "k_imu_predict": (stats["pred_bytes"], f"{PRED_BYTES} B + {PRED_SAMPLE_BYTES} B x IMU samples"),
        "k_candidates": (CAND_VOXEL_BYTES * stats["touched"], f"{CAND_VOXEL_BYTES} B x touched voxels"),
        "k_ins_keys": (INS_BYTES_POINT * stats["points"], f"{INS_BYTES_POINT} B x points"),
        "k_replay_heavy": (REFIT_BYTES * stats["refits"], f"{REFIT_BYTES} B x refits (k_replay_heavy_wave and "
                           "k_replay_heavy together)"),
        "k_merge_replay": (MERGE_BYTES * stats["merge_evals"], f"{MERGE_BYTES} B x merge() evaluations"),
        "k_create_evict": (LRU_BYTES * (stats["created"] + stats["evicted"]), f"{LRU_BYTES} B x (creations + evictions)"),
        "k_undistort": (28 * stats["raw_points"], "28 B x raw points (xyz + time in, xyz out)"),
    }
    for name, (ms, n) in kt.items():
        if n == 0:
            continue
        tot, desc = units.get(name, (0.0, "latency-bound, no byte model"))
        rows[name] = (ms, n, tot / n, desc)
    out = {}
    for name, (ms, n, bpl, desc) in rows.items():
        avg = ms / n
        out[name] = {
            "ms_per_scan": ms / scans,
            "launches": n,
            "avg_launch_ms": avg,
            "bytes_per_launch": bpl,
            "bytes_model": desc,
            "achieved_gbs": bpl / (avg * 1e-3) / 1e9 if avg > 0 else 0.0,
        }
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--points", type=int, default=120000)
    ap.add_argument("--capacity", type=int, default=2000000)
    ap.add_argument("--evict-capacity", type=int, default=500000,
                    help="LRU capacity of the second plain run, full after the fill (0: skip)")
    ap.add_argument("--fill", type=int, default=300,
                    help="scans registered before the warmup so the map reaches its LRU capacity")
    ap.add_argument("--roofline-scans", type=int, default=10)
    ap.add_argument("--cpu-warm", type=int, default=3)
    ap.add_argument("--cpu-timed", type=int, default=8)
    ap.add_argument("--cpu-threads", type=int, default=0, help="threads of the all-cores CPU leg (0 = all)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-sessions", type=int, default=128)
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--batch-warmup", type=int, default=5)
    ap.add_argument("--batch-capacity", type=int, default=200000)
    ap.add_argument("--batch-groups", type=int, default=4,
                    help="session groups, each on its own context (HIP stream) and host thread")
    ap.add_argument("--obs-ppt", type=int, default=0, help="observe points per thread (0 = by scan size)")
    ap.add_argument("--serve-sessions", type=int, default=512)
    ap.add_argument("--serve-rate", type=float, default=2000.0)
    ap.add_argument("--serve-seconds", type=float, default=4.0)
    ap.add_argument("--serve-executors", type=int, default=4)
    ap.add_argument("--serve-max-batch", type=int, default=128)
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            "bench.py measures one GPU: every box in this build has a single MI355X and the path "
            "shards by session with no exchange; run with --gpus 1.",
            file=sys.stderr,
        )
        sys.exit(2)

    import torch

    from voxelmapplus_fastlio2_amd import Context, LIOConfig, Registrar
    from voxelmapplus_fastlio2_amd.synth import SEED0, STREET, ImuSequence

    N, W, K, F, RF = args.points, args.warmup, args.steps, args.fill, args.roofline_scans
    RG = 10  # scans after the timed region whose graph-replayed stages are timed by their HIP events
    total = 2 + F + W + K + RG + RF
    seq = ImuSequence(STREET, SEED0, N)
    log(f"generating {total} raw packages x {N} points")
    t0 = time.perf_counter()
    pkgs = []
    for k in range(total):
        pkgs.append(seq.package(k))
        if (k + 1) % 100 == 0:
            log(f"  {k + 1} packages ({time.perf_counter() - t0:.1f} s)")
    log(f"generated in {time.perf_counter() - t0:.1f} s")
    dev = upload_packages(pkgs, 1)
    # one elementwise pass of known size (256 MiB read, 256 MiB written): the
    # reference launch that scales the FETCH_SIZE / WRITE_SIZE counters
    cal = torch.ones(64 << 20, dtype=torch.float32, device="cuda:0")
    torch.mul(cal, 2.0, out=torch.empty_like(cal))
    del cal
    torch.cuda.synchronize()

    cfg = LIOConfig(max_scan_points=N, map_capacity=args.capacity)
    ctx = Context(0)
    if args.obs_ppt:
        ctx.set_param("obs_ppt", args.obs_ppt)
    g = Registrar(ctx, cfg)
    stride = dev.stride(1) * 4

    def bind(k):
        base = dev[k - 1].data_ptr()
        _, _, _, _, imu, cs, ce = pkgs[k]
        g.set_raw_device(base, base + stride, base + 2 * stride, base + 3 * stride, N, imu, cs, ce)

    assert g.imu_init(pkgs[0][4], pkgs[0][6]), "IMU_INIT within the first package"
    bind(1)
    g.build()  # MAP_INIT
    tune = None
    if not args.obs_ppt:
        # observe launch shape for this scan size, picked by timing on the
        # session after MAP_INIT (its residual slots are restored afterwards)
        best, ms = ctx.autotune_observe(g, reps=10)
        tune = {"points_per_thread": best, "candidates_ms": dict(zip(["1", "2", "4", "8"], ms))}
        log(f"autotune observe ({N} points): {best} points per thread, medians {ms}")
    t0 = time.perf_counter()
    created = evicted = 0
    for k in range(2, 2 + F):
        bind(k)
        _, m = g.register()
        created += m.created
        evicted += m.evicted
        if (k - 1) % 50 == 0:
            log(f"  fill: scan {k}, map {m.map_size} voxels, created {created}, evicted {evicted} "
                f"({time.perf_counter() - t0:.1f} s)")
    for k in range(2 + F, 2 + F + W):
        bind(k)
        g.register()
    torch.cuda.synchronize()
    log(f"fill + warmup done: map {g.map_size()} voxels, {evicted} evictions so far")

    lat = []
    maint = []
    iters = 0
    mstat = {"created": 0, "evicted": 0, "merges": 0, "merge_evals": 0, "refits": 0, "touched": 0}
    torch.cuda.synchronize()
    gs0 = ctx.graph_stats()
    t_start = time.perf_counter()
    for k in range(2 + F + W, 2 + F + W + K):
        ts = time.perf_counter()
        bind(k)
        o, m = g.register()
        lat.append(time.perf_counter() - ts)
        maint.append(m.maintenance)
        iters += o.iterations
        for f in mstat:
            mstat[f] += getattr(m, f)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    gs1 = ctx.graph_stats()
    graphs = {k: gs1[k] - gs0[k] for k in gs1}
    # the filter stage (IESKF begin, iterations, end) and the map stage as the
    # timed region runs them (hipGraph replays), each between the library's
    # HIP event pair, over RG more scans
    stage_graph = {"filter": [], "map": []}
    for k in range(2 + F + W + K, 2 + F + W + K + RG):
        bind(k)
        g.register()
        f_ms, m_ms, _ = g.timing()
        stage_graph["filter"].append(f_ms)
        stage_graph["map"].append(m_ms)
    # roofline pass: the sequence continues for RF scans with per-launch HIP
    # events on the library's stream (the filter stage then launches kernel
    # by kernel instead of replaying its graph; the kernels are the same)
    g.set_timing(True)
    ctx.set_param("obs_time_reps", OBS_TIME_REPS)
    rstat = {"points": 0, "raw_points": 0, "refits": 0, "merge_evals": 0, "created": 0, "evicted": 0,
             "touched": 0, "ieskf_bytes": 0, "pred_bytes": 0}
    ppt = tune["points_per_thread"] if tune else args.obs_ppt
    for k in range(2 + F + W + K + RG, total):
        bind(k)
        o, m = g.register()
        rstat["points"] += m.points
        rstat["raw_points"] += N
        rstat["touched"] += m.touched
        blocks = -(-m.points // (256 * ppt))
        rstat["ieskf_bytes"] += o.iterations * (PART_ROW_BYTES * blocks + IESKF_ITER_BYTES)
        rstat["pred_bytes"] += PRED_BYTES + PRED_SAMPLE_BYTES * len(pkgs[k][4][0])
        for f in ("refits", "merge_evals", "created", "evicted"):
            rstat[f] += getattr(m, f)
    torch.cuda.synchronize()
    tt = g.timing_totals()
    kt = ctx.kernel_times()
    g.set_timing(False)
    ctx.set_param("obs_time_reps", 1)
    map_size = g.map_size()
    lat_ms = np.array(lat) * 1e3
    for j in np.argsort(lat_ms)[::-1][:3]:
        log(f"slow step {j}: {lat_ms[j]:.2f} ms, maintenance flags {maint[j]}")

    kr = kernel_rooflines(tt, kt, rstat, N, RF)
    for name, r in sorted(kr.items(), key=lambda kv: -kv[1]["ms_per_scan"]):
        log(f"kernel {name}: {r['ms_per_scan'] * 1e3:.1f} us/scan over {r['launches']} launches, "
            f"{r['achieved_gbs']:.1f} GB/s ({r['bytes_model']})")
    dom = max(kr, key=lambda k: kr[k]["ms_per_scan"])
    traffic = load_traffic()
    value = N * K / elapsed
    tr = traffic.get(dom)
    d = kr[dom]

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": elapsed / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 per point, f64 reductions, filter and IMU propagation",
        "data": "synthetic",
        "config": {
            "workload": f"street-{N // 1000}k-plain-steady",
            "form": "plain (one session, one raw scan + IMU package per step, host sync per scan)",
            "points_per_scan": N,
            "voxel_size": cfg.voxel_size,
            "map_capacity": cfg.map_capacity,
            "fill_scans": F,
            "map_voxels_end": map_size,
            "ieskf_iterations_mean": iters / K,
            "per_scan_mean": {f: v / K for f, v in mstat.items()},
            "maintenance_events": {"hash_rebuilds": sum(1 for f in maint if f & 1),
                                   "ring_compactions": sum(1 for f in maint if f & 2)},
            "observe_autotune": tune,
            "filter_stage": "hipGraph replay (IESKF begin, 5 x observe/step, end)",
            "map_stage": "hipGraph replay once the scan shape recurs (one capture per shape)",
            "graphs_in_timed_region": graphs,
        },
        "latency_ms": {
            "p50": float(np.percentile(lat_ms, 50)),
            "p99": float(np.percentile(lat_ms, 99)),
            "max": float(lat_ms.max()),
        },
        "stage_ms_per_scan": {
            "filter_stage_graph": float(np.mean(stage_graph["filter"])),
            "map_stage_graph": float(np.mean(stage_graph["map"])),
            "graph_measured_over": f"{RG} scans after the timed region, hipGraph replays between the library's "
            "HIP events (filter: IESKF begin, iterations and end; map: the whole map update)",
            "observe_kernels": tt["obs_ms"] / max(RF, 1),
            "map_update_eager": tt["map_ms"] / max(tt["map_calls"], 1),
            "measured_over": f"{RF} scans after those, launched kernel by kernel with per-launch HIP events",
        },
        "roofline": {
            "kernel": dom,
            "bound": "hbm",
            "achieved": d["achieved_gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": d["achieved_gbs"] / HBM_PEAK_GBS,
            "traffic": tr[0] if tr else None,
            "traffic_source": tr[1] if tr else None,
            "bytes_per_launch": d["bytes_per_launch"],
            "bytes_model": d["bytes_model"],
            "avg_launch_ms": d["avg_launch_ms"],
            "launches": d["launches"],
            "timed_over": f"{RF} scans following the timed region, HIP events around each launch on the "
            "library's stream (observe: around back-to-back repeats of each launch, mean per launch)",
        },
        "kernels": kr,
    }
    g.close()
    if args.evict_capacity > 0:
        line["plain_evicting"] = run_plain_evicting(ctx, dev, pkgs, N, args.evict_capacity, F, W, K)
    del dev
    if args.batch_sessions > 0:
        bt = run_batched(ctx, args.batch_sessions, args.batch_warmup, args.batch_steps, args.batch_capacity,
                         groups=args.batch_groups)
        bt["speedup_vs_plain"] = bt["value"] / value
        line["batched"] = bt
    ctx.close()
    if args.serve_sessions > 0:
        sv = run_serving(args.serve_sessions, args.serve_rate, args.serve_seconds, executors=args.serve_executors,
                         max_batch_sessions=args.serve_max_batch)
        sv["saturation_vs_plain"] = sv["saturation"]["value"] / value
        line["serving"] = sv
    if not args.no_cpu:
        info = host_info()
        log(f"timing the FP64 oracle on the host ({info})")
        line["cpu_baseline"] = cpu_baseline(pkgs, cfg, args.cpu_warm, args.cpu_timed)
        line["cpu_baseline"]["host"] = info
        line["vs_cpu"] = value / line["cpu_baseline"]["value"]
        threads = args.cpu_threads or info["affinity_cores"]
        log(f"timing the FP64 oracle with {threads} concurrent sessions")
        mt = cpu_baseline_sessions(pkgs, cfg, 2, 3, threads)
        mt["host"] = info
        line["cpu_baseline_all_cores"] = mt
        if "batched" in line:
            line["batched"]["vs_cpu_all_cores"] = line["batched"]["value"] / mt["value"]
            line["batched"]["vs_cpu_1_thread"] = line["batched"]["value"] / line["cpu_baseline"]["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
