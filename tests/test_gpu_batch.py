"""Batched registration (vmp_state_apply_motion_batch + vmp_register_batch):
sessions of different scenes and scan sizes advance together, one launch per
stage.  Every session's pose, P
and map must be bitwise identical to the same session registered alone
(each session's reductions use a block decomposition that depends only on
its own point count), and the alone run is itself checked against the oracle
in test_gpu_parity.py."""
import numpy as np
import pytest

from voxelmapplus_fastlio2_amd import (
    Context,
    LIOConfig,
    Registrar,
    apply_motion_batch,
    initial_covariance,
    register_batch,
    state_vector,
)
from voxelmapplus_fastlio2_amd.synth import QDIAG, ROOM, SEED0, STREET, Sequence

pytestmark = pytest.mark.gpu

SPECS = [(ROOM, 20000, 1), (STREET, 30000, 2), (ROOM, 7000, 3), (STREET, 12000, 4), (STREET, 2000, 5)]


def make(ctx, scene, n, seed):
    seq = Sequence(scene, SEED0 + seed, n)
    cfg = LIOConfig(max_scan_points=n, map_capacity=60000 if scene == STREET else 100000)
    g = Registrar(ctx, cfg)
    R, t = seq.pose(0)
    g.set_state(state_vector(R, t), initial_covariance())
    g.set_points(*seq.scan(0))
    g.build()
    return seq, g


def test_batch_matches_alone():
    nscans = 6
    ctx_a = Context(0)
    ctx_b = Context(0)
    alone = [make(ctx_a, *sp) for sp in SPECS]
    mixed = [make(ctx_b, *sp) for sp in SPECS]
    for k in range(1, nscans):
        for (seq, g) in alone:
            g.apply_motion(*seq.motion(k), QDIAG)
            g.set_points(*seq.scan(k))
            g.register()
        mot = [seq.motion(k) for seq, _ in mixed]
        apply_motion_batch([g for _, g in mixed], [a for a, _ in mot], [b for _, b in mot], QDIAG)
        for (seq, g) in mixed:
            g.set_points(*seq.scan(k))
        o, m = register_batch([g for _, g in mixed])
        for i, ((_, ga), (_, gb)) in enumerate(zip(alone, mixed)):
            sa, Pa = ga.get_state()
            sb, Pb = gb.get_state()
            assert np.array_equal(sa, sb), f"session {i} scan {k}: state differs alone vs batch"
            assert np.array_equal(Pa, Pb), f"session {i} scan {k}: P differs alone vs batch"
            ea, eb = ga.export_map(), gb.export_map()
            for key in ("keys", "ints", "group", "dbl"):
                assert np.array_equal(ea[key], eb[key]), f"session {i} scan {k}: map '{key}' differs"
        print(f"scan {k}: batch of {len(mixed)} identical to alone; merges", [x.merges for x in m], flush=True)
    for _, g in alone + mixed:
        g.close()
    ctx_a.close()
    ctx_b.close()


def test_filter_graph_matches_eager():
    """The filter stage replayed from its captured hipGraph (the default)
    gives bitwise the results of launching its kernels one by one, over
    several scans with the eager map stage in between every replay, alone
    and in a batch."""
    nscans = 5
    ctx_g = Context(0)
    ctx_e = Context(0)
    ctx_e.set_param("filter_graph", 0)
    graph = [make(ctx_g, *sp) for sp in SPECS]
    eager = [make(ctx_e, *sp) for sp in SPECS]
    for k in range(1, nscans):
        for group in (graph, eager):
            mot = [seq.motion(k) for seq, _ in group]
            apply_motion_batch([g for _, g in group], [a for a, _ in mot], [b for _, b in mot], QDIAG)
            for seq, g in group:
                g.set_points(*seq.scan(k))
            if k % 2:
                register_batch([g for _, g in group])
            else:
                for _, g in group:
                    g.register()
        for i, ((_, ga), (_, gb)) in enumerate(zip(graph, eager)):
            sa, Pa = ga.get_state()
            sb, Pb = gb.get_state()
            assert np.array_equal(sa, sb), f"session {i} scan {k}: state differs graph vs eager"
            assert np.array_equal(Pa, Pb), f"session {i} scan {k}: P differs graph vs eager"
        print(f"scan {k}: filter graph replay identical to eager launches", flush=True)
    for _, g in graph + eager:
        g.close()
    ctx_g.close()
    ctx_e.close()


def test_map_graph_matches_eager():
    """The map stage replayed from its captured hipGraph (captured once a
    step's shape recurs: scans of a fixed size) gives bitwise the results of
    launching its kernels one by one, scan after scan, with the program's own
    eager work (motion prior, point upload, state and map reads) between the
    replays.  The 14k-voxel LRU fills at the second scan, so the replays
    cover eviction too."""
    nscans = 10
    n = 20000
    ctx_g = Context(0)
    ctx_e = Context(0)
    ctx_e.set_param("map_graph", 0)
    seq = Sequence(STREET, SEED0 + 71, n)
    cfg = LIOConfig(max_scan_points=n, map_capacity=14000)
    runs = []
    for ctx in (ctx_g, ctx_e):
        g = Registrar(ctx, cfg)
        R, t = seq.pose(0)
        g.set_state(state_vector(R, t), initial_covariance())
        g.set_points(*seq.scan(0))
        g.build()
        runs.append(g)
    ga, ge = runs
    evicted = 0
    for k in range(1, nscans):
        x, y, z = seq.scan(k)
        res = []
        for g in runs:
            g.apply_motion(*seq.motion(k), QDIAG)
            g.set_points(x, y, z)
            res.append(g.register())
        (oa, ma), (ob, mb) = res
        for f, _ in ma._fields_:
            assert getattr(ma, f) == getattr(mb, f), f"scan {k}: map stat {f} differs graph vs eager"
        evicted += ma.evicted
        sa, Pa = ga.get_state()
        sb, Pb = ge.get_state()
        assert np.array_equal(sa, sb) and np.array_equal(Pa, Pb), f"scan {k}: posterior differs graph vs eager"
        ea, eb = ga.export_map(), ge.export_map()
        for key in ("keys", "ints", "group", "dbl"):
            assert np.array_equal(ea[key], eb[key]), f"scan {k}: map '{key}' differs graph vs eager"
        print(f"scan {k}: map stage {ctx_g.graph_stats()}, evicted {ma.evicted}: identical", flush=True)
    st = ctx_g.graph_stats()
    assert st["map_replays"] >= nscans - 3 and st["map_builds"] >= 1, st
    assert ctx_e.graph_stats()["map_replays"] == 0
    assert evicted > 0
    for g in runs:
        g.close()
    ctx_g.close()
    ctx_e.close()


def test_replay_forms_match():
    """The refitting voxels' replay in its two forms, one wave per voxel
    (moments per lane, refits side by side from snapshots) and one lane per
    voxel (a sequential walk), gives bitwise the same map: a batch of mixed
    sessions registered with every voxel on its own wave (the default at
    these sizes), with every voxel on a lane, and with a split at 8 points.
    The MAP_INIT builds take the same forms."""
    nscans = 5
    forms = {"wave": {}, "lane": {"replay_wave_max": 0, "replay_long_seg": 0},
             "split": {"replay_wave_max": 0, "replay_long_seg": 8}}
    ctxs, sess = {}, {}
    for name, params in forms.items():
        ctxs[name] = Context(0)
        for k, v in params.items():
            ctxs[name].set_param(k, v)
        sess[name] = [make(ctxs[name], *sp) for sp in SPECS]
    refits = 0
    for k in range(1, nscans):
        stats = {}
        for name, group in sess.items():
            mot = [seq.motion(k) for seq, _ in group]
            apply_motion_batch([g for _, g in group], [a for a, _ in mot], [b for _, b in mot], QDIAG)
            for (seq, g) in group:
                g.set_points(*seq.scan(k))
            stats[name] = register_batch([g for _, g in group])[1]
        for i, (_, ref) in enumerate(sess["wave"]):
            sa, Pa = ref.get_state()
            ea = ref.export_map()
            for name in ("lane", "split"):
                g = sess[name][i][1]
                sb, Pb = g.get_state()
                assert np.array_equal(sa, sb) and np.array_equal(Pa, Pb), f"session {i} scan {k}: {name} differs"
                eb = g.export_map()
                for key in ("keys", "ints", "group", "dbl"):
                    assert np.array_equal(ea[key], eb[key]), f"session {i} scan {k}: {name} map '{key}' differs"
                for f, _ in stats["wave"][i]._fields_:
                    assert getattr(stats["wave"][i], f) == getattr(stats[name][i], f), f"{name} stat {f}"
            refits += stats["wave"][i].refits
        print(f"scan {k}: replay forms identical; refits", [m.refits for m in stats["wave"]], flush=True)
    assert refits > 0
    for name in sess:
        for _, g in sess[name]:
            g.close()
        ctxs[name].close()


def test_repeated_observe_timing_leaves_results_unchanged():
    """Per-launch timing with obs_time_reps > 1 runs every observe launch
    several times back to back (the kernel is idempotent) and once more for
    the pass's counters: poses, P, maps and observe statistics stay bitwise
    those of an untimed context."""
    nscans = 4
    ctx_t = Context(0)
    ctx_r = Context(0)
    ctx_t.set_param("obs_time_reps", 5)
    timed = [make(ctx_t, *sp) for sp in SPECS[:3]]
    ref = [make(ctx_r, *sp) for sp in SPECS[:3]]
    for _, g in timed:
        g.set_timing(True)
    for k in range(1, nscans):
        stats = []
        for group in (timed, ref):
            mot = [seq.motion(k) for seq, _ in group]
            apply_motion_batch([g for _, g in group], [a for a, _ in mot], [b for _, b in mot], QDIAG)
            for seq, g in group:
                g.set_points(*seq.scan(k))
            stats.append(register_batch([g for _, g in group]))
        (ot, mt), (orf, mrf) = stats
        for i, ((_, ga), (_, gb)) in enumerate(zip(timed, ref)):
            sa, Pa = ga.get_state()
            sb, Pb = gb.get_state()
            assert np.array_equal(sa, sb), f"session {i} scan {k}: state differs with repeated timing"
            assert np.array_equal(Pa, Pb), f"session {i} scan {k}: P differs with repeated timing"
            for f in ("points", "hits", "plane_hits", "effective", "iterations"):
                assert getattr(ot[i], f) == getattr(orf[i], f), f"session {i} scan {k}: observe stat {f} differs"
            ea, eb = ga.export_map(), gb.export_map()
            for key in ("keys", "ints", "group", "dbl"):
                assert np.array_equal(ea[key], eb[key]), f"session {i} scan {k}: map '{key}' differs"
    tt = timed[0][1].timing_totals()
    assert tt["obs_launches"] > 0 and tt["obs_ms"] > 0.0
    for _, g in timed + ref:
        g.close()
    ctx_t.close()
    ctx_r.close()


def test_autotune_observe_leaves_results_unchanged():
    """vmp_autotune_observe picks points per thread for the scan's size
    bucket and restores the residual slots its timing passes wrote: the next
    registration is bitwise the one of a context given the same table by
    hand, with no tuning passes."""
    scene, n, seed = STREET, 30000, 9
    ctx_t = Context(0)
    ctx_r = Context(0)
    seq_t, g_t = make(ctx_t, scene, n, seed)
    seq_r, g_r = make(ctx_r, scene, n, seed)
    for k in range(1, 4):
        for seq, g in ((seq_t, g_t), (seq_r, g_r)):
            g.apply_motion(*seq.motion(k), QDIAG)
            g.set_points(*seq.scan(k))
        if k == 1:
            best, ms = ctx_t.autotune_observe(g_t, reps=5)
            print("autotune", best, ms, flush=True)
            assert best in (1, 2, 4, 8) and all(t > 0 for t in ms)
            ctx_r.set_param("obs_ppt_bucket1", best)  # 30k points: the < 32k bucket
        g_t.register()
        g_r.register()
        st, Pt = g_t.get_state()
        sr, Pr = g_r.get_state()
        assert np.array_equal(st, sr) and np.array_equal(Pt, Pr), f"scan {k}: tuned context differs"
        vt, rt, _ = g_t.export_slots(n)
        vr, rr, _ = g_r.export_slots(n)
        assert np.array_equal(vt, vr) and np.array_equal(rt[vt], rr[vr])
    for g in (g_t, g_r):
        g.close()
    ctx_t.close()
    ctx_r.close()


def test_autotune_observe_batch_leaves_results_unchanged():
    """The batch form tunes every occupied size bucket on one observe launch
    over all sessions and restores their residual slots: the batch's next
    registrations equal those of a context given the same table by hand."""
    from voxelmapplus_fastlio2_amd import autotune_observe_batch

    ctx_t = Context(0)
    ctx_r = Context(0)
    tuned = [make(ctx_t, *sp) for sp in SPECS]
    ref = [make(ctx_r, *sp) for sp in SPECS]
    table, ms = autotune_observe_batch([g for _, g in tuned], reps=3)
    print("batch autotune", table, ms, flush=True)
    assert all(v in (1, 2, 4, 8) for v in table)
    for k, v in enumerate(table):
        ctx_r.set_param(f"obs_ppt_bucket{k}", v)
    for k in range(1, 4):
        for group in (tuned, ref):
            mot = [seq.motion(k) for seq, _ in group]
            apply_motion_batch([g for _, g in group], [a for a, _ in mot], [b for _, b in mot], QDIAG)
            for seq, g in group:
                g.set_points(*seq.scan(k))
            register_batch([g for _, g in group])
        for i, ((_, ga), (_, gb)) in enumerate(zip(tuned, ref)):
            sa, Pa = ga.get_state()
            sb, Pb = gb.get_state()
            assert np.array_equal(sa, sb) and np.array_equal(Pa, Pb), f"session {i} scan {k}: tuned batch differs"
    for _, g in tuned + ref:
        g.close()
    ctx_t.close()
    ctx_r.close()
