"""GPU parity of the registration path on edge cases, against the oracle.

  empty scan          no points: the prior-only update (x = prior, P kept)
  all-miss scan       a fresh session's scan that lands outside the map: no
                      valid slot, "NO EFFECTIVE POINT" (lio_builder.cpp:308-309),
                      prior-only update, then the points are inserted
  z == 0 points       calcBodyCov's z = 0 -> 0.001 mutation (commons.cpp:20-23)
                      applies to observation and insertion covariance, not to
                      the insertion point
  non-finite input    NaN / Inf coordinates are rejected (VMP_EINVAL): host
                      buffers when bound, device buffers by the registration
                      call, before it launches anything; the session is unchanged
  dense merges        a flat 40 x 40-voxel floor whose voxels all freeze in one
                      scan, then one merge() call per voxel in a strided order:
                      merge-replay windows with more than 32 acting entries
                      (more than one thread per record pair of the block)
All calls go through the C ABI.  Tolerances as in test_gpu_parity.py.
"""
import sys

import numpy as np
import pytest

from mapcmp import compare_maps, register_both, state_err
from oracle import Oracle
from voxelmapplus_fastlio2_amd import Context, LIOConfig, Registrar, VmpError, initial_covariance, state_vector
from voxelmapplus_fastlio2_amd.synth import QDIAG, ROOM, SEED0, Sequence

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def started(ctx, seq, cfg):
    R, t = seq.pose(0)
    s0, P0 = state_vector(R, t), initial_covariance()
    x, y, z = seq.scan(0)
    orc = Oracle(cfg)
    orc.set_state(s0, P0)
    orc.build(x, y, z)
    g = Registrar(ctx, cfg)
    g.set_state(s0, P0)
    g.set_points(x, y, z)
    g.build()
    return orc, g


def step_both(orc, g, seq, k, pts):
    orc.apply_motion(*seq.motion(k), QDIAG)
    sp, Pp = orc.get_state()
    it, so, Po, ost, mst, sg, Pg = register_both(orc, g, *pts)
    return sp, Pp, so, Po, sg, Pg, it, ost, mst


def test_empty_scan_is_a_prior_only_update(ctx):
    seq = Sequence(ROOM, SEED0 + 41, 8000)
    cfg = LIOConfig(max_scan_points=8000)
    orc, g = started(ctx, seq, cfg)
    with g:
        empty = (np.zeros(0, np.float32),) * 3
        sp, Pp, so, Po, sg, Pg, it, ost, mst = step_both(orc, g, seq, 1, empty)
        log("empty scan: iterations o/g", it, ost.iterations, "effective", ost.effective)
        assert it == 1 and ost.iterations == 1 and ost.effective == 0
        np.testing.assert_array_equal(so, sp)
        assert np.abs(sg - sp).max() <= 1e-12
        np.testing.assert_allclose(Pg, Pp, rtol=1e-9, atol=1e-15)
        np.testing.assert_allclose(Po, Pp, rtol=1e-9, atol=1e-15)
        assert mst.points == 0 and mst.created == 0
        assert np.array_equal(orc.export_map()["keys"], g.export_map()["keys"])
        # the next scan registers normally
        sp, Pp, so, Po, sg, Pg, it, ost, mst = step_both(orc, g, seq, 2, seq.scan(2))
        ep, er = state_err(so, sg)
        assert ep <= 1e-4 and er <= 1e-5


def test_all_miss_scan_is_a_prior_only_update(ctx):
    seq = Sequence(ROOM, SEED0 + 42, 8000)
    cfg = LIOConfig(max_scan_points=8000)
    orc, g = started(ctx, seq, cfg)
    with g:
        x, y, z = seq.scan(1)
        far = (x + np.float32(500.0), y, z)  # 500 m beyond the room: every lookup misses
        sp, Pp, so, Po, sg, Pg, it, ost, mst = step_both(orc, g, seq, 1, far)
        log("all-miss scan: hits", ost.hits, "effective", ost.effective, "iterations o/g", it, ost.iterations,
            "created", mst.created)
        assert ost.hits == 0 and ost.effective == 0
        assert it == 1 and ost.iterations == 1
        np.testing.assert_array_equal(so, sp)
        assert np.abs(sg - sp).max() <= 1e-12
        np.testing.assert_allclose(Pg, Po, rtol=1e-9, atol=1e-15)
        assert mst.created > 0
        r = compare_maps(orc.export_map(), g.export_map(), report=lambda o: log("  map", o))
        assert r["plane_flag_diff"] <= max(2, r["n"] // 1000)


def test_zero_z_points(ctx):
    seq = Sequence(ROOM, SEED0 + 43, 12000)
    cfg = LIOConfig(max_scan_points=12000)
    orc, g = started(ctx, seq, cfg)
    with g:
        for k in range(1, 5):
            x, y, z = seq.scan(k)
            z = z.copy()
            z[::9] = 0.0  # one point in nine on the sensor's horizontal plane
            sp, Pp, so, Po, sg, Pg, it, ost, mst = step_both(orc, g, seq, k, (x, y, z))
            ep, er = state_err(so, sg)
            vo, ro, no = orc.export_slots(len(x))
            vg, rg, ng = g.export_slots(len(x))
            flips = int((vo != vg).sum())
            both = vo & vg
            log(f"z=0 scan {k}: pos {ep:.2e} rot {er:.2e} flips {flips} iters o/g {it}/{ost.iterations}")
            assert ep <= 1e-4 and er <= 1e-5
            assert flips <= max(3, len(x) // 2000)
            assert np.abs(ro[both] - rg[both]).max() < 1e-5
            zero = np.zeros(len(x), bool)
            zero[::9] = True
            assert (both & zero).sum() > 10, "z == 0 points take part in the update"
            r = compare_maps(orc.export_map(), g.export_map())
            assert r["plane_flag_diff"] <= max(2, r["n"] // 1000)
            assert r["mean_bad"] <= r["plane_flag_diff"] * 8


def test_non_finite_points_are_rejected(ctx):
    import torch

    seq = Sequence(ROOM, SEED0 + 44, 6000)
    cfg = LIOConfig(max_scan_points=6000)
    orc, g = started(ctx, seq, cfg)
    with g:
        x, y, z = seq.scan(1)
        before, Pb = g.get_state()
        for bad in (np.nan, np.inf, -np.inf):
            xb = x.copy()
            xb[1234] = bad
            with pytest.raises(VmpError, match="-1"):
                g.set_points(xb, y, z)
            # device buffers: checked when the registration call starts,
            # which then fails before launching anything
            dev = torch.from_numpy(np.stack([xb, y, z])).to("cuda:0")
            torch.cuda.synchronize()
            g.set_points_device(dev[0].data_ptr(), dev[1].data_ptr(), dev[2].data_ptr(), len(x))
            with pytest.raises(VmpError, match="-1"):
                g.register()
            with pytest.raises(VmpError, match="-6"):  # the binding was dropped
                g.register()
        after, Pa = g.get_state()
        np.testing.assert_array_equal(before, after)
        np.testing.assert_array_equal(Pb, Pa)
        # finite points bound through the device path register as usual
        dev = torch.from_numpy(np.stack([x, y, z])).to("cuda:0")
        torch.cuda.synchronize()
        orc.apply_motion(*seq.motion(1), QDIAG)
        sp, Pp = orc.get_state()
        orc.register(x, y, z)
        g.set_state(sp, Pp)
        g.set_points_device(dev[0].data_ptr(), dev[1].data_ptr(), dev[2].data_ptr(), len(x))
        g.register()
        torch.cuda.synchronize()
        ep, er = state_err(orc.get_state()[0], g.get_state()[0])
        assert ep <= 1e-4 and er <= 1e-5


def floor_scene(nx=40, ny=40, per_voxel=110, seed=9):
    """World points on a flat floor z = 0.02 over nx x ny voxels of 0.5 m,
    kept 0.1 m inside each voxel's faces."""
    rng = np.random.default_rng(seed)
    ix, iy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    ix, iy = ix.reshape(-1), iy.reshape(-1)

    def pts(order, k):
        u = rng.uniform(0.1, 0.4, size=(len(order) * k, 2))
        vx = np.repeat(ix[order], k) * 0.5 + u[:, 0] - 10.0
        vy = np.repeat(iy[order], k) * 0.5 + u[:, 1] - 10.0
        vz = 0.02 + 0.001 * rng.normal(size=len(vx))
        return np.stack([vx, vy, vz], 1)

    nv = nx * ny
    build = pts(np.arange(nv), per_voxel)
    raster = pts(np.arange(nv), 1)
    # strided visit order: consecutive points land in voxels several cells apart
    strided = np.array(sorted(range(nv), key=lambda v: ((v * 37) % nv)))
    second = pts(strided, 1)
    return build, np.concatenate([raster, second])


def test_dense_merge_windows(ctx):
    build, scan = floor_scene()
    t = np.array([0.3, -0.2, 1.8])
    lidar = lambda pw: (pw - t).astype(np.float32)
    b = lidar(build)
    s1 = lidar(scan)
    cfg = LIOConfig(max_scan_points=len(b))
    s0, P0 = state_vector(np.eye(3), t), initial_covariance()
    orc = Oracle(cfg)
    orc.set_state(s0, P0)
    orc.build(b[:, 0], b[:, 1], b[:, 2])
    with Registrar(ctx, cfg) as g:
        g.set_state(s0, P0)
        g.set_points(b[:, 0], b[:, 1], b[:, 2])
        g.build()
        orc.set_state(s0, P0)
        _, so, _, ost, mst, sg, _ = register_both(orc, g, s1[:, 0], s1[:, 1], s1[:, 2])
        st = orc.stats()
        log("dense merges:", {f: getattr(mst, f) for f, _ in mst._fields_}, "oracle", st)
        assert mst.merge_max_acting > 32, "a window with more acting entries than one thread per record pair covers"
        assert mst.merges == st["merges"] and mst.merge_calls == st["merge_calls"]
        ep, er = state_err(so, sg)
        assert ep <= 1e-4 and er <= 1e-5
        a, m = orc.export_map(), g.export_map()
        r = compare_maps(a, m, report=lambda o: log("  map", o))
        assert r["plane_flag_diff"] == 0 and r["group_diff"] == 0 and r["merged_flag_diff"] == 0
        assert r["mean_bad"] == 0 and r["norm_bad"] == 0
