"""Pins the oracle's voxel map (voxel_map.cpp:29-256) on the CPU.

  index          brute-force floor(p / voxel_size), exact multiples, negatives
  LRU/eviction   a brute-force OrderedDict replica on tiny capacities: front-to-
                 back key order, the evicted set and the created count are
                 exact; voxel point counts follow min(#points since creation, 100)
  pushPoint      hand cases: init at the 10th point, a refit every 10 points,
                 freeze at 100, frozen planes call merge(), frozen non-planes drop
  updatePlane    exactly coplanar points: lambda0 = 0, normal = -sign(mean.n) e_z,
                 mean = arithmetic mean; the never-reset plane covariance has the
                 closed-form traces  tm = 3c sum_k 1/n_k  and
                 tn = c sum_k (1/(n_k l1_k) + 1/(n_k l2_k))  for point cov c*I
  merge          hand-computed fusion with the reference's operator precedence
                 (voxel_map.cpp:167-169), both gates, group propagation
  build          MAP_INIT adds every point before fitting (no 100-point cap)
"""
from collections import OrderedDict

import numpy as np
import pytest

import oracle
from voxelmapplus_fastlio2_amd import LIOConfig

VS = 0.5


def make(capacity=100000, **kw):
    return oracle.Oracle(LIOConfig(map_capacity=capacity, **kw))


def push(o, pts, c=1e-4, is_build=False):
    pts = np.atleast_2d(np.asarray(pts, float))
    covs = np.repeat((c * np.eye(3))[None], len(pts), axis=0)
    o.push(pts, covs, is_build)


def voxel_rows(o):
    m = o.export_map()
    return {tuple(k): i for i, k in enumerate(m["keys"].tolist())}, m


FLAG_INIT, FLAG_PLANE, FLAG_ENABLE, FLAG_MERGED = 1, 2, 4, 8


# ----------------------------------------------------------------- index
def test_index_is_floor_of_quotient():
    o = make()
    rng = np.random.default_rng(1)
    pts = np.concatenate(
        [
            rng.uniform(-200, 200, size=(300, 3)),
            np.array([[0.0, -0.0, 0.5], [-0.5, -1e-300, 1.0], [0.49999999999, -0.5000000001, 2.5]]),
            np.round(rng.uniform(-50, 50, size=(50, 3)) * 2) / 2,  # exact multiples of the voxel size
        ]
    )
    for p in pts:
        np.testing.assert_array_equal(o.key(p), np.floor(p / VS).astype(np.int64))


# ----------------------------------------------------------------- LRU
class Replica:
    """Brute-force voxel map: OrderedDict in LRU order plus per-voxel counts."""

    def __init__(self, cap):
        self.cap, self.lru, self.created, self.evicted = cap, OrderedDict(), 0, 0

    def touch(self, k):
        if k in self.lru:
            self.lru.move_to_end(k, last=False)
            self.lru[k] += 1
        else:
            self.lru[k] = 1
            self.lru.move_to_end(k, last=False)
            self.created += 1
            if len(self.lru) > self.cap:
                self.lru.popitem(last=True)
                self.evicted += 1


@pytest.mark.parametrize("cap,seed", [(3, 0), (7, 1), (20, 2), (1, 3)])
def test_lru_order_and_eviction_match_bruteforce(cap, seed):
    rng = np.random.default_rng(seed)
    o = make(capacity=cap)
    rep = Replica(cap)
    # voxels in a 4x3x2 block; points well inside their voxel, spread in 3D
    # so that no voxel becomes a plane (counts stay min(n, 100))
    for batch in range(30):
        m = rng.integers(1, 60)
        vox = rng.integers(0, [4, 3, 2], size=(m, 3))
        off = rng.uniform(0.02, 0.48, size=(m, 3))
        pts = vox * VS + off
        push(o, pts)
        for v in map(tuple, vox.tolist()):
            rep.touch(v)
        rows, mp = voxel_rows(o)
        assert list(rows) == list(rep.lru), f"LRU order differs after batch {batch}"
        st = o.stats()
        assert st["new_voxels"] == rep.created and st["evicted"] == rep.evicted
        np.testing.assert_array_equal(mp["ints"][:, 0], np.minimum(list(rep.lru.values()), 100))


def test_lru_is_touch_order_not_insertion_order():
    o = make(capacity=2)
    a, b, c = (0.1, 0.1, 0.1), (0.6, 0.1, 0.1), (1.1, 0.1, 0.1)
    push(o, [a, b, a])  # a most recent
    push(o, [c])  # evicts b (the least recently touched), not a
    rows, _ = voxel_rows(o)
    assert list(rows) == [(2, 0, 0), (0, 0, 0)]


# ----------------------------------------------------------------- pushPoint / updatePlane
def coplanar(n, z=0.1, x0=0.0, rng=None):
    rng = rng or np.random.default_rng(5)
    p = np.empty((n, 3))
    p[:, 0] = x0 + rng.uniform(0.01, 0.49, n)
    p[:, 1] = rng.uniform(0.01, 0.49, n)
    p[:, 2] = z
    return p


def test_plane_lifecycle_and_closed_form_covariance():
    c = 1e-4
    o = make()
    pts = coplanar(101)
    tn_expect, tm_expect = 0.0, 0.0
    for i in range(101):
        push(o, pts[i], c)
        rows, mp = voxel_rows(o)
        n, flags, newly, tmp = mp["ints"][0, :4]
        st = o.stats()
        k = i + 1
        if k < 10:
            assert flags == FLAG_ENABLE and n == k and st["refits"] == 0
            continue
        if k <= 100:
            assert n == k and flags & FLAG_INIT and flags & FLAG_PLANE
            assert st["refits"] == k // 10
            if k % 10 == 0:
                q = pts[:k]
                lam = np.linalg.eigvalsh(np.cov(q.T, bias=True))
                tn_expect += c * (1.0 / (k * lam[1]) + 1.0 / (k * lam[2]))
                tm_expect += 3.0 * c / k
                np.testing.assert_allclose(mp["dbl"][0, 0:3], q.mean(axis=0), rtol=0, atol=1e-15)
                np.testing.assert_allclose(mp["dbl"][0, 3:6], [0, 0, -1], atol=1e-12)  # mean.z > 0 -> n.z < 0
                np.testing.assert_allclose(mp["dbl"][0, 6], tn_expect, rtol=1e-6)
                np.testing.assert_allclose(mp["dbl"][0, 7], tm_expect, rtol=1e-12)
            if k < 100:
                assert flags & FLAG_ENABLE and tmp == k
            else:
                assert not flags & FLAG_ENABLE and tmp == 0  # frozen, temp points released
        else:  # 101st point: frozen plane -> merge() and no insertion
            assert n == 100 and st["merge_calls"] == 1 and st["dropped"] == 0 and st["refits"] == 10


def test_frozen_non_plane_drops_points():
    o = make()
    corners = np.array([[x, y, z] for x in (0.02, 0.48) for y in (0.02, 0.48) for z in (0.02, 0.48)])
    pts = corners[np.arange(105) % 8]
    push(o, pts)
    _, mp = voxel_rows(o)
    n, flags = mp["ints"][0, :2]
    st = o.stats()
    assert n == 100 and flags == FLAG_INIT  # initialised, not a plane, frozen
    assert st["dropped"] == 5 and st["merge_calls"] == 0 and st["refits"] == 10


def test_build_adds_all_points_then_fits_once():
    o = make()
    pts = coplanar(150)
    push(o, pts, is_build=True)
    _, mp = voxel_rows(o)
    n, flags, newly, tmp = mp["ints"][0, :4]
    assert n == 150 and tmp == 150 and flags == FLAG_INIT | FLAG_PLANE | FLAG_ENABLE
    assert o.stats()["refits"] == 1
    push(o, coplanar(1, rng=np.random.default_rng(9)))  # 151 >= 100 -> freeze on the next insert
    _, mp = voxel_rows(o)
    assert mp["ints"][0, 0] == 151 and not mp["ints"][0, 1] & FLAG_ENABLE and mp["ints"][0, 3] == 0


# ----------------------------------------------------------------- merge
def two_frozen_planes(zb, tilt=0.0):
    o = make()
    rng = np.random.default_rng(11)
    b = coplanar(100, z=zb, x0=VS, rng=rng)
    if tilt:
        # plane n.p = -zb with n = (0, sin t, -cos t): the same offset n.m as
        # A's plane (n = (0, 0, -1), n.m = -0.1), so only the angle gate acts
        b[:, 1] = rng.uniform(0.01, 0.3, len(b))
        b[:, 2] = (b[:, 1] * np.sin(tilt) + zb) / np.cos(tilt)
    push(o, b, 2e-4)
    a = coplanar(100, z=0.1, rng=rng)
    push(o, a, 1e-4)
    return o, a


def test_merge_fusion_with_reference_precedence():
    o, a = two_frozen_planes(0.11)
    rows, mp = voxel_rows(o)
    ia, ib = rows[(0, 0, 0)], rows[(1, 0, 0)]
    g = mp["group"]
    assert g[ia] != g[ib]
    mA, nA, tnA, tmA = mp["dbl"][ia, 0:3], mp["dbl"][ia, 3:6], mp["dbl"][ia, 6], mp["dbl"][ia, 7]
    mB, nB, tnB, tmB = mp["dbl"][ib, 0:3], mp["dbl"][ib, 3:6], mp["dbl"][ib, 6], mp["dbl"][ib, 7]
    tcA, tcB = tnA + tmA, tnB + tmB
    push(o, a[:1])  # the 101st point of A: A is frozen -> merge() visits +x = B
    rows, mp = voxel_rows(o)
    ia, ib = rows[(0, 0, 0)], rows[(1, 0, 0)]
    # the reference writes  tm0 * m_near + tm1 * m_this / (tm0 + tm1): only the
    # second term is divided; the same for the normal
    mean = tmA * mB + (tmB * mA) / (tmA + tmB)
    norm = tnA * nB + (tnB * nA) / (tnA + tnB)
    if -(mean @ norm) < 0.0:
        norm = -norm
    for i in (ia, ib):
        np.testing.assert_allclose(mp["dbl"][i, 0:3], mean, rtol=1e-13)
        np.testing.assert_allclose(mp["dbl"][i, 3:6], norm, rtol=1e-13)
        np.testing.assert_allclose(mp["dbl"][i, 6], (tcA**2 * tnB + tcB**2 * tnA) / (tcA + tcB) ** 2, rtol=1e-12)
        np.testing.assert_allclose(mp["dbl"][i, 7], (tcA**2 * tmB + tcB**2 * tmA) / (tcA + tcB) ** 2, rtol=1e-12)
        assert mp["ints"][i, 1] & FLAG_MERGED
    assert abs(np.linalg.norm(norm) - 1.0) > 1e-3  # the fused normal is not renormalised
    assert mp["group"][ia] == mp["group"][ib]
    st = o.stats()
    assert st["merges"] == 1 and st["merge_calls"] == 1
    # same group now: a second call skips B
    push(o, a[1:2])
    assert o.stats()["merges"] == 1 and o.stats()["merge_calls"] == 2


@pytest.mark.parametrize(
    "zb,tilt,merges",
    [
        (0.13, 0.0, 1),  # |d_axis| = 0.03 <= 0.04
        (0.15, 0.0, 0),  # 0.05 > 0.04
        (0.1, np.arccos(1 - 0.19), 1),  # 1 - cos = 0.19 <= 0.2
        (0.1, np.arccos(1 - 0.25), 0),  # 0.25 > 0.2
    ],
)
def test_merge_gates(zb, tilt, merges):
    o, a = two_frozen_planes(zb, tilt)
    rows, mp = voxel_rows(o)
    if tilt:
        ib = rows[(1, 0, 0)]
        assert mp["ints"][ib, 1] & FLAG_PLANE
    push(o, a[:1])
    assert o.stats()["merges"] == merges
