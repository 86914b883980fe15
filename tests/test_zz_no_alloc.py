"""vmp_register_scan_noalloc, the registration entry point that allocates
nothing, against vmp_register_scan: two sessions start from the same state
and map, one registers through each entry point, and every result (posterior
state and P bit for bit, observe and map statistics, the exported map) must
be identical, over two calls in a row, the second with a smaller input.
Also: sizes beyond the scratch are refused before anything runs, raw IMU
packages with the scan filter, and LRU ring compaction inside the call.
All calls go through the C ABI."""
import sys

import numpy as np
import pytest

from voxelmapplus_fastlio2_amd import Context, LIOBuilder, LIOConfig, Registrar, VmpError, initial_covariance
from voxelmapplus_fastlio2_amd import state_vector
from voxelmapplus_fastlio2_amd.synth import QDIAG, ROOM, SEED0, STREET, ImuSequence, Sequence

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same_results(a, b, ra, rb, what):
    (oa, ma), (ob, mb) = ra, rb
    for f, _ in oa._fields_:
        assert getattr(oa, f) == getattr(ob, f), f"{what}: observe stat {f}"
    for f, _ in ma._fields_:
        assert getattr(ma, f) == getattr(mb, f), f"{what}: map stat {f}"
    sa, Pa = a.get_state()
    sb, Pb = b.get_state()
    assert np.array_equal(sa, sb) and np.array_equal(Pa, Pb), f"{what}: posterior differs"
    ea, eb = a.export_map(), b.export_map()
    for key in ("keys", "ints", "group", "dbl"):
        assert np.array_equal(ea[key], eb[key]), f"{what}: map '{key}' differs"


def pair(ctx, cfg, seq):
    R, t = seq.pose(0)
    s0, P0 = state_vector(R, t), initial_covariance()
    x, y, z = seq.scan(0)
    out = []
    for _ in range(2):
        g = Registrar(ctx, cfg)
        g.set_state(s0, P0)
        g.set_points(x, y, z)
        g.build()
        out.append(g)
    return out


def test_noalloc_matches_register_scan_device_points():
    import torch

    cfg = LIOConfig(max_scan_points=131072, map_capacity=600000)
    seq = Sequence(STREET, SEED0 + 61, 120000)
    ctx = Context(0)
    a, b = pair(ctx, cfg, seq)
    for k, n in ((1, 120000), (2, 30000)):  # the second call with a smaller scan
        x, y, z = seq.scan(k, n)
        dev = torch.from_numpy(np.stack([x, y, z])).to("cuda:0")
        torch.cuda.synchronize()
        for g in (a, b):
            g.apply_motion(*seq.motion(k), QDIAG)
            g.set_points_device(dev[0].data_ptr(), dev[1].data_ptr(), dev[2].data_ptr(), n)
        ra = a.register()
        rb = b.register_noalloc()
        log(f"scan {k} ({n} points): iterations {ra[0].iterations}, created {ra[1].created}, map {ra[1].map_size}")
        same_results(a, b, ra, rb, f"scan {k}")
        del dev
    a.close()
    b.close()
    ctx.close()


def test_noalloc_raw_packages_with_filter():
    cfg = LIOConfig(max_scan_points=131072, map_capacity=400000, scan_resolution=0.1)
    seq = ImuSequence(ROOM, SEED0 + 62, 100000)
    ctx = Context(0)
    la, lb = LIOBuilder(ctx, cfg), LIOBuilder(ctx, cfg)
    sizes = [100000, 100000, 100000, 40000]  # IMU_INIT, MAP_INIT, then two registered packages
    for k, n in enumerate(sizes):
        pkg = seq.package(k, n)
        if k < 2:
            la.process(*pkg)
            lb.process(*pkg)
            continue
        la.reg.set_raw(*pkg[:7])
        lb.reg.set_raw(*pkg[:7])
        ra = la.reg.register()
        rb = lb.reg.register_noalloc()
        log(f"package {k} ({n} raw points): filtered {ra[1].points}, iterations {ra[0].iterations}")
        assert ra[1].points < n
        same_results(la.reg, lb.reg, ra, rb, f"package {k}")
        assert np.array_equal(la.reg.export_scan(), lb.reg.export_scan())
    la.close()
    lb.close()
    ctx.close()


def test_noalloc_ring_compaction():
    """A room rescanned from rest: every scan appends its ~5,200 touched
    voxels to the LRU ring again and evicts nothing, so the ring (45,024
    entries at these limits) compacts after about seven scans, inside the
    no-allocation call, on the scratch."""
    cfg = LIOConfig(max_scan_points=8000, map_capacity=6000)
    seq = Sequence(ROOM, SEED0 + 31, 8000)
    ctx = Context(0)
    a, b = pair(ctx, cfg, seq)
    x, y, z = seq.scan(0)
    still = (np.eye(3), np.zeros(3))
    compactions = 0
    for k in range(1, 13):
        for g in (a, b):
            g.apply_motion(*still, QDIAG)
            g.set_points(x, y, z)
        ra = a.register()
        rb = b.register_noalloc()
        compactions += (rb[1].maintenance & 2) != 0
        same_results(a, b, ra, rb, f"rescan {k}")
    log(f"ring compactions: {compactions}")
    assert compactions >= 1
    a.close()
    b.close()
    ctx.close()


def test_noalloc_limits():
    ctx = Context(0)
    big = LIOConfig(max_scan_points=131073, map_capacity=100000)
    seq = Sequence(ROOM, SEED0 + 64, 131073)
    with Registrar(ctx, big) as g:
        R, t = seq.pose(0)
        g.set_state(state_vector(R, t), initial_covariance())
        x, y, z = seq.scan(0)
        g.set_points(x, y, z)
        g.build()
        before = g.get_state()
        g.set_points(*seq.scan(1))
        with pytest.raises(VmpError, match="131072"):
            g.register_noalloc()
        after = g.get_state()
        assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])
    wide = LIOConfig(max_scan_points=120000, map_capacity=2000000)  # ring of 4.5M entries
    seq2 = Sequence(ROOM, SEED0 + 65, 20000)
    with Registrar(ctx, wide) as g:
        R, t = seq2.pose(0)
        g.set_state(state_vector(R, t), initial_covariance())
        g.set_points(*seq2.scan(0))
        g.build()
        g.set_points(*seq2.scan(1))
        with pytest.raises(VmpError, match="LRU ring"):
            g.register_noalloc()
        g.register()  # the regular entry point still serves it
    ctx.close()
