"""The measured configurations, on the GPU, at their full sizes, against the
oracle (and, for the multi-session forms, against the same sessions run
alone):

  C1  room, 20k points per raw scan with its 200 Hz IMU package, the full
      300-scan sequence through LIOBuilder::process (IMU_INIT, MAP_INIT,
      undistortion, IESKF, map update), the oracle taking the device's state
      before every package (teacher forcing)
  C2  street, 120k points per scan, 25 teacher-forced scans with a 100k-voxel
      LRU, so that scans evict from the eighth on: keys and LRU order
      bit-exact, poses, P, plane records and Plane::cov traces
  C4  128 sessions (room / street alternating), raw scans of U[2k, 120k]
      points with their IMU packages, 4 batch steps after MAP_INIT: 4 sampled
      sessions bitwise equal to the same sessions run alone, 2 of them also
      against the oracle
  C5  512 serving sessions, log-U[2k, 120k] points, all arriving at once
      (continuous batching), 2 scans each after MAP_INIT: 4 sampled sessions
      bitwise equal alone, 2 of them against the oracle
Each test stops at the first mismatch.  Tolerances as in test_gpu_parity.py
and test_gpu_imu.py.
"""
import sys

import numpy as np
import pytest

from mapcmp import compare_maps, lio_both, register_both, state_err, ulp_err
from oracle import Oracle
from voxelmapplus_fastlio2_amd import (
    Context,
    ImuConfig,
    LIOBuilder,
    LIOConfig,
    Registrar,
    initial_covariance,
    register_batch,
    set_raw_device_batch,
    state_vector,
)
from voxelmapplus_fastlio2_amd.synth import QDIAG, ROOM, SEED0, STREET, ImuSequence, Sequence

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def imu_oracle(cfg):
    o = Oracle(cfg)
    o.imu_config()
    return o


def test_c1_room_300_scans():
    """Every package from the device's state on both sides (lio_both): the
    oracle follows the device through all 300 packages, so each scan checks
    one device step from the same state and map."""
    cfg = LIOConfig(max_scan_points=20000, map_capacity=100000)
    seq = ImuSequence(ROOM, SEED0, 20000)
    orc = imu_oracle(cfg)
    ctx = Context(0)
    worst = {"pos": 0.0, "rot": 0.0, "P": 0.0, "ulp": 0.0}
    with LIOBuilder(ctx, cfg) as lb:
        for k in range(300):
            r = lio_both(orc, lb, seq.package(k))
            assert r["status"][0] == r["status"][1], (k, r["status"])
            ep, er = state_err(r["so"], r["sg"])
            worst["pos"] = max(worst["pos"], ep)
            worst["rot"] = max(worst["rot"], er)
            if r["points"] is not None:
                po, pg = r["points"]
                assert po.shape == pg.shape, (k, po.shape, pg.shape)
                worst["ulp"] = max(worst["ulp"], ulp_err(po, pg))
                if r["it"][0] == r["it"][1]:
                    eP = np.abs(r["Po"] - r["Pg"]).max() / np.abs(r["Po"]).max()
                    worst["P"] = max(worst["P"], eP)
            assert ep <= 1e-4 and er <= 1e-5, (k, ep, er)
            if k % 25 == 0 or k == 299:
                m = compare_maps(orc.export_map(), lb.reg.export_map())
                log(f"C1 scan {k}: status {lb.status} map {m['n']} worst {worst}", m)
                margin = m["plane_flag_diff"] + m["merged_flag_diff"]
                assert m["plane_flag_diff"] <= max(2, m["n"] // 1000)
                assert m["mean_bad"] <= margin * 8 and m["norm_bad"] <= margin * 8 and m["trace_bad"] <= margin * 8
                assert m["ill_conditioned_bad"] == 0
        assert worst["ulp"] <= 2.0 and worst["P"] <= 1e-6, worst
    ctx.close()


def test_c2_street_120k_with_eviction():
    npts, cap, nscans = 120000, 100000, 25
    seq = Sequence(STREET, SEED0 + 2, npts)
    cfg = LIOConfig(max_scan_points=npts, map_capacity=cap)
    R, t = seq.pose(0)
    s0, P0 = state_vector(R, t), initial_covariance()
    x0, y0, z0 = seq.scan(0)
    orc = Oracle(cfg)
    orc.set_state(s0, P0)
    orc.build(x0, y0, z0)
    ctx = Context(0)
    evicted = 0
    with Registrar(ctx, cfg) as g:
        g.set_state(s0, P0)
        g.set_points(x0, y0, z0)
        g.build()
        for k in range(1, nscans):
            orc.apply_motion(*seq.motion(k), QDIAG)
            x, y, z = seq.scan(k)
            it, so, Po, ost, mst, sg, Pg = register_both(orc, g, x, y, z)
            evicted += mst.evicted
            ep, er = state_err(so, sg)
            eP = np.abs(Po - Pg).max() / np.abs(Po).max()
            a, b = orc.export_map(), g.export_map()
            r = compare_maps(a, b)
            log(f"C2 scan {k}: created {mst.created} evicted {mst.evicted} map {mst.map_size} iters o/g "
                f"{it}/{ost.iterations} pos {ep:.2e} rot {er:.2e} P {eP:.2e}", r)
            assert ep <= 1e-4 and er <= 1e-5
            if it == ost.iterations:
                assert eP <= 1e-9
            assert np.array_equal(a["keys"], b["keys"]), f"LRU order differs after scan {k}"
            margin = r["plane_flag_diff"] + r["merged_flag_diff"]
            assert r["plane_flag_diff"] <= max(2, r["n"] // 1000)
            assert r["mean_bad"] <= margin * 8 and r["norm_bad"] <= margin * 8 and r["trace_bad"] <= margin * 8
            assert r["ill_conditioned_bad"] == 0
    assert evicted > 50000 and orc.stats()["evicted"] == evicted
    ctx.close()


def raw_batch(B, nsteps, nmax, seed):
    """Raw packages of B sequences: per step a list of (x, y, z, t, imu, cs, ce).
    The first (IMU_INIT) package is a full sweep, so it carries the 20
    samples that initialise the IMU; the later ones hold U[2k, nmax] points
    (a shorter scan ends earlier in its sweep)."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(2000, nmax + 1, size=(nsteps, B))
    sizes[0, :] = nmax
    seqs = [ImuSequence(ROOM if i % 2 == 0 else STREET, SEED0 + 7000 + i, nmax) for i in range(B)]
    return [[seqs[i].package(k, int(sizes[k, i])) for i in range(B)] for k in range(nsteps)]


def test_c4_batch_128_sessions():
    import torch

    B, nsteps, nmax = 128, 6, 120000
    pk = raw_batch(B, nsteps, nmax, 0xC4)
    cfg = LIOConfig(max_scan_points=nmax, map_capacity=200000)
    ctx = Context(0)
    regs = [Registrar(ctx, cfg) for _ in range(B)]
    for i, g in enumerate(regs):
        assert g.imu_init(pk[0][i][4], pk[0][i][6])
    for k in range(1, nsteps):
        sizes = [len(p[0]) for p in pk[k]]
        offs = np.concatenate([[0], np.cumsum(sizes)])
        dev = torch.from_numpy(np.concatenate([np.stack(p[:4]) for p in pk[k]], axis=1)).to("cuda:0")
        base, st = dev.data_ptr(), dev.stride(0) * 4
        xo = base + offs[:-1].astype(np.uint64) * 4
        set_raw_device_batch(regs, xo, xo + st, xo + 2 * st, xo + 3 * st, sizes, [p[4] for p in pk[k]],
                             [p[5] for p in pk[k]], [p[6] for p in pk[k]])
        if k == 1:
            for g in regs:
                g.build()
        else:
            _, m = register_batch(regs)
            log(f"C4 step {k}: {sum(sizes)} points, evicted {sum(x.evicted for x in m)}, merges "
                f"{sum(x.merges for x in m)}")
        torch.cuda.synchronize()
        del dev
    sample = [0, 1, 64, 127]
    for i in sample:  # the same session alone through LIOBuilder
        with LIOBuilder(ctx, cfg) as lb:
            for k in range(nsteps):
                lb.process(*pk[k][i])
            sa, Pa = lb.reg.get_state()
            sb, Pb = regs[i].get_state()
            assert np.array_equal(sa, sb) and np.array_equal(Pa, Pb), f"C4 session {i}: batch differs from alone"
            ea, eb = lb.reg.export_map(), regs[i].export_map()
            for key in ("keys", "ints", "group", "dbl"):
                assert np.array_equal(ea[key], eb[key]), f"C4 session {i}: map '{key}' differs"
        if i in (0, 1):  # and against the oracle
            orc = imu_oracle(cfg)
            for k in range(nsteps):
                orc.process(*pk[k][i])
            so, _ = orc.get_state()
            ep, er = state_err(so, sb)
            r = compare_maps(orc.export_map(), regs[i].export_map())
            log(f"C4 session {i}: alone == batch; oracle pos {ep:.2e} rot {er:.2e}", r)
            assert ep <= 1e-4 and er <= 1e-5
            assert r["plane_flag_diff"] <= max(2, r["n"] // 1000)
    for g in regs:
        g.close()
    ctx.close()


def test_c5_serving_512_sessions():
    import torch

    from voxelmapplus_fastlio2_amd.serving import Server, make_requests

    S, nscans, nmin, nmax = 512, 3, 2000, 120000
    rng = np.random.default_rng(0xC5)
    sizes = np.exp(rng.uniform(np.log(nmin), np.log(nmax), size=(S, nscans))).astype(np.int64)
    seqs = [Sequence(ROOM if i % 2 == 0 else STREET, SEED0 + 9000 + i, nmax) for i in range(S)]
    cfg = LIOConfig(max_scan_points=nmax, map_capacity=200000)
    bufs, items = [], []
    for k in range(nscans):
        for i in range(S):
            x, y, z = seqs[i].scan(k, int(sizes[i, k]))
            d = torch.from_numpy(np.stack([x, y, z])).to("cuda:0")
            bufs.append(d)
            base, st = d.data_ptr(), d.stride(0) * 4
            it = {"session": i, "n": int(sizes[i, k]), "x": base, "y": base + st, "z": base + 2 * st,
                  "arrival": 0.0}
            if k > 0:
                it["dR"], it["dp"] = seqs[i].motion(k)
            items.append(it)
    reqs = make_requests(items)
    starts = [(state_vector(*seqs[i].pose(0)), initial_covariance()) for i in range(S)]
    with Server(cfg, S, executors=3, max_batch_sessions=128, hbm_budget=200 << 30) as sv:
        for i in range(S):
            sv.session_init(i, *starts[i])
        sv.run(reqs, QDIAG)
        st = sv.stats()
        log("C5 server stats", st)
        assert all(r.status == 0 for r in reqs)
        assert st["evictions"] == 0 and st["max_batch"] > 8
        ctx = Context(0)
        for i in (0, 1, 255, 511):
            g = Registrar(ctx, cfg)
            g.set_state(*starts[i])
            orc = Oracle(cfg) if i in (0, 1) else None
            if orc is not None:
                orc.set_state(*starts[i])
            for k in range(nscans):
                x, y, z = seqs[i].scan(k, int(sizes[i, k]))
                if k > 0:
                    g.apply_motion(*seqs[i].motion(k), QDIAG)
                g.set_points(x, y, z)
                if k == 0:
                    g.build()
                    if orc is not None:
                        orc.build(x, y, z)
                else:
                    g.register()
                    if orc is not None:
                        orc.apply_motion(*seqs[i].motion(k), QDIAG)
                        orc.register(x, y, z)
            mine = [r for r in reqs if r.session == i]
            s, _ = g.get_state()
            assert np.array_equal(np.array(mine[-1].pos), s[0:3]) and np.array_equal(np.array(mine[-1].rot), s[3:12])
            view = sv.session(i)
            a, b = g.export_map(), view.export_map()
            for key in ("keys", "ints", "group", "dbl"):
                assert np.array_equal(a[key], b[key]), f"C5 session {i}: map '{key}' differs"
            if orc is not None:
                ep, er = state_err(orc.get_state()[0], s)
                r = compare_maps(orc.export_map(), b)
                log(f"C5 session {i}: alone == served; oracle pos {ep:.2e} rot {er:.2e}", r)
                assert ep <= 1e-4 and er <= 1e-5
                assert r["plane_flag_diff"] <= max(2, r["n"] // 1000)
            g.close()
        ctx.close()
