"""Further pins of the oracle on the CPU.

  stale slots (Q2)    a point whose lookup misses keeps the slot of its
                      index from the previous pass or scan, and that slot
                      re-enters H and b with the *current* point_lidar and
                      state (voxel_map.cpp:258-276, lio_builder.cpp:269-273,
                      285-306); a "clear the slot on a miss" variant gives a
                      different H and is rejected by the same check
  insertion transform the fp32 PCL transform (lio_builder.cpp:155-163) stays
                      within the fp32 rounding bound of the fp64 transform
                      R_wl p + p_wl, its rounding is really fp32 (an fp64
                      transform would fail), and insertion keys equal the
                      fp64 keys except within that bound of a voxel face
"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import obsref
import oracle
from voxelmapplus_fastlio2_amd import LIOConfig, initial_covariance, state_vector


def room_points(n=21000, seed=5):
    rng = np.random.default_rng(seed)
    m = n // 6
    u, v = rng.uniform(size=(2, 6, m))
    faces = [
        np.stack([u[0] * 8 - 3, v[0] * 6 - 2, np.zeros(m)], 1),
        np.stack([u[1] * 8 - 3, v[1] * 6 - 2, np.full(m, 3.0)], 1),
        np.stack([np.full(m, -3.0), u[2] * 6 - 2, v[2] * 3], 1),
        np.stack([np.full(m, 5.0), u[3] * 6 - 2, v[3] * 3], 1),
        np.stack([u[4] * 8 - 3, np.full(m, -2.0), v[4] * 3], 1),
        np.stack([u[5] * 8 - 3, np.full(m, 4.0), v[5] * 3], 1),
    ]
    pw = np.concatenate(faces)
    return pw + 0.003 * rng.normal(size=pw.shape)


@pytest.fixture(scope="module")
def room():
    R0 = Rotation.from_rotvec([0.03, -0.02, 0.4]).as_matrix()
    t0 = np.array([0.9, 0.5, 1.2])
    pw = room_points()
    pl = ((pw - t0) @ R0).astype(np.float32)
    return R0, t0, pl


def shifted(s, dx, dr=(0.0, 0.0, 0.0)):
    d = np.zeros(23)
    d[0:3] = dx
    d[3:6] = dr
    return oracle.boxplus(s, d)


def test_stale_slots_reenter_with_current_point_and_state(room):
    R0, t0, pl = room
    s0, P0 = state_vector(R0, t0), initial_covariance()
    o = oracle.Oracle(LIOConfig(max_scan_points=len(pl)))
    o.set_state(s0, P0)
    o.build(pl[:, 0], pl[:, 1], pl[:, 2])
    mp = o.export_map()
    pts = obsref.lidar_points(pl[:, 0], pl[:, 1], pl[:, 2])

    # pass 1 at a slightly perturbed state: most slots become valid
    s1 = shifted(s0, [0.01, -0.01, 0.005])
    o.set_points(pl[:, 0], pl[:, 1], pl[:, 2])
    H1, b1, _ = o.observe(s1)
    slots1 = o.export_slots(len(pl))
    v1, r1, n1, hit1, mg1, H1e, b1e = obsref.observe(s1, P0, pts, mp)
    clear1 = mg1 > 1e-9
    np.testing.assert_array_equal(slots1[0][clear1], v1[clear1])
    assert slots1[0].sum() > 0.6 * len(pl)

    # scan k+1: other points (every point moved 2 cm along its ray) observed
    # at a state 2.6 m away, so a large share of the lookups miss the map
    pl2 = (pl.astype(np.float64) * 1.004).astype(np.float32)
    pts2 = obsref.lidar_points(pl2[:, 0], pl2[:, 1], pl2[:, 2])
    s2 = shifted(s1, [2.6, 1.7, 0.0], [0.0, 0.0, 0.06])
    o.set_points(pl2[:, 0], pl2[:, 1], pl2[:, 2])
    H2, b2, eff2 = o.observe(s2)
    slots2 = o.export_slots(len(pl))
    v2, r2, nn2, hit2, mg2, H2e, b2e = obsref.observe(s2, P0, pts2, mp, prev=slots1)
    miss = ~hit2
    stale_valid = miss & slots1[0]
    assert miss.sum() > 0.2 * len(pl) and stale_valid.sum() > 1000, (miss.sum(), stale_valid.sum())
    assert (hit2 & v2).sum() > 100, "fresh valid slots take part as well"
    clear2 = mg2 > 1e-9
    np.testing.assert_array_equal(slots2[0][clear2], v2[clear2])
    # the missed slots are bit-for-bit what the previous pass left
    np.testing.assert_array_equal(slots2[0][miss], slots1[0][miss])
    np.testing.assert_array_equal(slots2[1][miss], slots1[1][miss])
    np.testing.assert_array_equal(slots2[2][miss], slots1[2][miss])
    assert eff2 == int(slots2[0].sum())
    scale = np.abs(H2e).max()
    np.testing.assert_allclose(H2, H2e, rtol=1e-9, atol=1e-9 * scale)
    np.testing.assert_allclose(b2, b2e, rtol=1e-9, atol=1e-9 * np.abs(b2e).max())

    # the same check rejects the variants a plausible mistake would give
    Hc, bc = obsref.hb(s2, pts2, v2 & hit2, r2, nn2)  # slots cleared on a miss
    assert np.linalg.norm(Hc - H2) > 1e-2 * np.linalg.norm(H2)
    Hp, _ = obsref.hb(s2, pts, v2, r2, nn2)  # stale slots with the previous scan's points
    assert np.linalg.norm(Hp - H2) > 1e-4 * np.linalg.norm(H2)
    Hs, _ = obsref.hb(s1, pts2, v2, r2, nn2)  # stale slots with the previous pass's state
    assert np.linalg.norm(Hs - H2) > 1e-4 * np.linalg.norm(H2)


def test_shorter_scan_keeps_the_slots_beyond_it(room):
    """residual_info is indexed by point: a shorter scan leaves the slots
    past its end untouched, and a later longer scan that misses there sees
    them again (lio_builder.cpp:224-229, 260)."""
    R0, t0, pl = room
    s0, P0 = state_vector(R0, t0), initial_covariance()
    o = oracle.Oracle(LIOConfig(max_scan_points=len(pl)))
    o.set_state(s0, P0)
    o.build(pl[:, 0], pl[:, 1], pl[:, 2])
    o.set_points(pl[:, 0], pl[:, 1], pl[:, 2])
    o.observe(shifted(s0, [0.01, 0.0, 0.0]))
    before = o.export_slots(len(pl))
    half = len(pl) // 2
    o.set_points(pl[:half, 0], pl[:half, 1], pl[:half, 2])
    o.observe(shifted(s0, [-0.01, 0.02, 0.0]))
    after = o.export_slots(len(pl))
    for a, b in zip(before, after):
        np.testing.assert_array_equal(a[half:], b[half:])


@pytest.mark.parametrize("offset", [0.0, 1.0e4])
def test_insertion_transform_fp32_bound(offset):
    rng = np.random.default_rng(11)
    n = 50000
    R = Rotation.from_rotvec(rng.normal(size=3) * 0.7).as_matrix()
    Rx = Rotation.from_rotvec([0.01, -0.02, 0.015]).as_matrix()
    t = np.array([offset + 12.3456789, -0.5 * offset + 3.21, 45.678])
    s = state_vector(R, t)
    s[12:21] = Rx.T.reshape(-1)
    s[21:24] = [0.04, -0.03, 0.11]
    o = oracle.Oracle(LIOConfig(max_scan_points=n))
    o.set_state(s, initial_covariance())
    d = rng.normal(size=(n, 3))
    pl = (d / np.linalg.norm(d, axis=1, keepdims=True) * rng.uniform(0.2, 40.0, size=(n, 1))).astype(np.float32)
    got = o.lidar_to_world_f32(pl[:, 0], pl[:, 1], pl[:, 2]).astype(np.float64)
    rwl = R @ Rx
    pwl = R @ s[21:24] + t
    p64 = pl.astype(np.float64)
    exact = p64 @ rwl.T + pwl
    # fp32 transform: the matrix and translation rounded to fp32, three
    # products and three sums each rounded (no contraction); the error is
    # bounded by a few units of 2^-24 of the terms' magnitudes
    u = 2.0**-24
    mag = np.abs(p64) @ np.abs(rwl).T + np.abs(pwl)
    err = np.abs(got - exact)
    bound = 6.0 * u * mag
    assert np.all(err <= bound), float((err / bound).max())
    # the rounding really is fp32 (an fp64 transform would leave ~1e-12 here)
    assert err.max() > 0.05 * u * mag.max()
    # insertion keys: equal to the fp64 keys except within the bound of a face
    kg = np.floor(got / 0.5)
    ke = np.floor(exact / 0.5)
    frac = exact / 0.5 - np.floor(exact / 0.5)
    near = np.minimum(frac, 1.0 - frac) * 0.5 <= bound
    assert np.all((kg == ke) | near)
    if offset > 0:
        assert (kg != ke).any(), "at 10 km some fp32 keys differ from the fp64 ones"
