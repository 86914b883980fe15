"""Continuous-batching server (vmp_server_run): scans of several sessions
arrive over time, are batched as sessions become idle, and run on two
overlapping executor streams.  Every result must be bitwise identical to the
same session's scans registered alone through the plain API (MAP_INIT where
the server did a MAP_INIT, registration otherwise), also when a small memory
budget forces sessions to be evicted and re-admitted."""
import numpy as np
import pytest

from voxelmapplus_fastlio2_amd import Context, LIOConfig, Registrar, initial_covariance, state_vector
from voxelmapplus_fastlio2_amd.serving import Server, make_requests
from voxelmapplus_fastlio2_amd.synth import QDIAG, ROOM, SEED0, STREET, Sequence

pytestmark = pytest.mark.gpu

NMAX = 30000
SPECS = [(ROOM, 20000), (STREET, 30000), (ROOM, 6000), (STREET, 12000), (STREET, 3000), (ROOM, 15000)]


def upload(seqs, nscans):
    import torch

    bufs = {}
    for i, (seq, n) in enumerate(seqs):
        for k in range(nscans):
            x, y, z = seq.scan(k)
            t = torch.from_numpy(np.stack([x, y, z])).to("cuda:0")
            bufs[(i, k)] = (t, n)
    return bufs


def build_requests(seqs, bufs, nscans, rate, seed):
    rng = np.random.default_rng(seed)
    items = []
    t = 0.0
    for k in range(nscans):
        for i in rng.permutation(len(seqs)):
            if k > 0:  # every session's MAP_INIT scan arrives at once, then Poisson
                t += rng.exponential(1.0 / rate)
            dev, n = bufs[(int(i), k)]
            base, st = dev.data_ptr(), dev.stride(0) * 4
            it = {"session": int(i), "n": n, "x": base, "y": base + st, "z": base + 2 * st, "arrival": t}
            if k > 0:
                it["dR"], it["dp"] = seqs[int(i)][0].motion(k)
            items.append(it)
    return make_requests(items)


def replay_alone(ctx, cfg, seq, s0, P0, kinds, scans):
    """Plain-API replay following the server's MAP_INIT / register decisions."""
    poses = []
    g = None
    last = (s0, P0)
    for k, kind in enumerate(kinds):
        x, y, z = scans[k]
        if kind in (1, 3):
            if g is not None:
                last = g.get_state()
                g.close()
            g = Registrar(ctx, cfg)
            g.set_state(*last)
            if kind == 3:
                g.apply_motion(*seq.motion(k), QDIAG)
            g.set_points(x, y, z)
            g.build()
        else:
            g.apply_motion(*seq.motion(k), QDIAG)
            g.set_points(x, y, z)
            g.register()
        s, _ = g.get_state()
        poses.append(s)
    return g, poses


def check(budget_sessions, rate, nscans=5, seed=1):
    cfg = LIOConfig(max_scan_points=NMAX, map_capacity=60000)
    seqs = [(Sequence(sc, SEED0 + 300 + i, n), n) for i, (sc, n) in enumerate(SPECS)]
    bufs = upload(seqs, nscans)
    reqs = build_requests(seqs, bufs, nscans, rate, seed)
    from voxelmapplus_fastlio2_amd import _lib
    import ctypes as C

    per = _lib.vmp().vmp_session_bytes(C.byref(cfg.to_c()))
    budget = per * budget_sessions + per // 2
    starts = []
    with Server(cfg, len(seqs), executors=2, max_batch_sessions=4, hbm_budget=budget) as sv:
        for i, (seq, _) in enumerate(seqs):
            R, t = seq.pose(0)
            s0, P0 = state_vector(R, t), initial_covariance()
            starts.append((s0, P0))
            sv.session_init(i, s0, P0)
        sv.run(reqs, QDIAG)
        st = sv.stats()
        print("server stats", st, flush=True)
        assert all(r.status == 0 for r in reqs)
        ctx = Context(0)
        for i, (seq, n) in enumerate(seqs):
            mine = [r for r in reqs if r.session == i]
            kinds = [r.kind for r in mine]
            scans = [seq.scan(k) for k in range(nscans)]
            g, poses = replay_alone(ctx, cfg, seq, *starts[i], kinds, scans)
            for k, (r, s) in enumerate(zip(mine, poses)):
                assert np.array_equal(np.array(r.pos), s[0:3]), f"session {i} scan {k}: position differs"
                assert np.array_equal(np.array(r.rot), s[3:12]), f"session {i} scan {k}: rotation differs"
            view = sv.session(i)
            if view is not None:
                a, b = g.export_map(), view.export_map()
                for key in ("keys", "ints", "group", "dbl"):
                    assert np.array_equal(a[key], b[key]), f"session {i}: map '{key}' differs"
            print(f"session {i}: kinds {kinds}, batches {[r.batch_sessions for r in mine]} identical to alone",
                  flush=True)
            g.close()
        ctx.close()
        lat = np.array([r.done - r.arrival for r in reqs])
        print(f"latency p50 {np.percentile(lat, 50) * 1e3:.2f} ms max {lat.max() * 1e3:.2f} ms", flush=True)
        return st


def test_serving_matches_alone():
    st = check(budget_sessions=len(SPECS), rate=300.0)
    assert st["evictions"] == 0 and st["requests"] == len(SPECS) * 5
    assert st["batches"] < st["requests"], "requests were batched"


def test_serving_with_eviction_matches_alone():
    st = check(budget_sessions=2, rate=300.0, seed=2)
    assert st["evictions"] > 0
