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
