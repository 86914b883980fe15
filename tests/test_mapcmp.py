"""The map comparison used by the GPU parity tests, checked on the CPU with
mistakes planted in a copy of an oracle map: a tolerance widened by the
fit's conditioning must still catch a wrong normal on a well-conditioned
plane, and a normal outside the near-degenerate eigenspace of an
ill-conditioned one."""
import numpy as np

from mapcmp import compare_maps
from oracle import Oracle
from voxelmapplus_fastlio2_amd import LIOConfig, initial_covariance, state_vector
from voxelmapplus_fastlio2_amd.synth import QDIAG, ROOM, SEED0, Sequence


def static_room_map():
    """A room rescanned from rest: voxels fill with copies of a few distinct
    points, so some fits are near-degenerate."""
    room = Sequence(ROOM, SEED0 + 31, 8000)
    o = Oracle(LIOConfig(max_scan_points=8000, map_capacity=6000))
    R, t = room.pose(0)
    o.set_state(state_vector(R, t), initial_covariance())
    x, y, z = room.scan(0)
    o.build(x, y, z)
    for _ in range(4):
        o.apply_motion(np.eye(3), np.zeros(3), QDIAG)
        o.register(x, y, z)
    return o.export_map()


def test_conditioning_aware_normal_check():
    ref = static_room_map()
    r = compare_maps(ref, ref)
    assert r["norm_bad"] == 0 and r["ill_conditioned_bad"] == 0 and r["trace_bad"] == 0
    assert r["ill_conditioned"] > 0, "the rescans make near-degenerate fits"
    pl = (ref["ints"][:, 1] & 10) == 2  # planes, not merged
    ev = ref["fit"]
    gap = ev[:, 1] - ev[:, 0]
    good = np.flatnonzero(pl & (gap > 1e-4))
    sc = np.maximum(1.0, (ref["dbl"][:, 0:3] ** 2).sum(axis=1))
    cond = 1e-13 * sc / np.maximum(gap, 1e-300)
    ill = np.flatnonzero(pl & (cond > 1e-3) & (ev[:, 2] - ev[:, 1] > 1e-4))
    assert len(good) and len(ill)

    # a well-conditioned normal tilted by 1e-4 rad is caught
    got = {k: v.copy() for k, v in ref.items()}
    i = good[0]
    n = got["dbl"][i, 3:6]
    m = ev[i, 3:6]
    got["dbl"][i, 3:6] = (n + 1e-4 * m) / np.linalg.norm(n + 1e-4 * m)
    assert compare_maps(ref, got)["norm_bad"] == 1

    # an ill-conditioned one may turn within its near-degenerate eigenspace
    # by what its conditioning allows ...
    got = {k: v.copy() for k, v in ref.items()}
    j = ill[0]
    n, m = ref["dbl"][j, 3:6], ev[j, 3:6]
    w = np.cross(n, m)
    w /= np.linalg.norm(w)
    a = min(0.5 * cond[j], 1.5)
    got["dbl"][j, 3:6] = np.cos(a) * n + np.sin(a) * w
    r = compare_maps(ref, got)
    assert r["norm_bad"] == 0 and r["ill_conditioned_bad"] == 0
    got["dbl"][j, 3:6] = np.cos(4 * a) * n + np.sin(4 * a) * w
    assert compare_maps(ref, got)["norm_bad"] == 1
    # ... but not onto the major axis, nor lose unit length
    got["dbl"][j, 3:6] = m
    assert compare_maps(ref, got)["ill_conditioned_bad"] == 1
    got["dbl"][j, 3:6] = 1.01 * n
    assert compare_maps(ref, got)["ill_conditioned_bad"] == 1
