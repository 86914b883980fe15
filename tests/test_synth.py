"""The seeded scene generator (host C++): bit-identical output against the
committed digests, exact point counts, the configured range window and
prefix-consistent partial scans."""
import json
import os

import numpy as np

from voxelmapplus_fastlio2_amd.synth import ROOM, SEED0, STREET, Sequence

HERE = os.path.dirname(os.path.abspath(__file__))


def test_digests_match_the_committed_fixture():
    import importlib.util

    spec = importlib.util.spec_from_file_location("mk", os.path.join(HERE, "golden", "make_synth_digest.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    want = json.load(open(os.path.join(HERE, "golden", "synth_digest.json")))
    for name, sc, seed, n, ks in mk.CASES:
        assert mk.digest(sc, seed, n, ks) == want[name]["digests"], name
    for name, sc, seed, n, npkg in mk.IMU_CASES:
        assert mk.digest_imu(sc, seed, n, npkg) == want[name]["digests"], name


def test_scan_shape_range_and_prefix():
    for scene in (ROOM, STREET):
        seq = Sequence(scene, SEED0 + 5, 8000)
        x, y, z = seq.scan(3)
        assert len(x) == len(y) == len(z) == 8000
        r = np.sqrt(x.astype(np.float64) ** 2 + y**2 + z**2)
        assert r.min() >= 0.1 - 1e-6 and r.max() <= 40.0 + 1e-4
        xs, ys, zs = seq.scan(3, 1000)
        np.testing.assert_array_equal(xs, x[:1000])
        # a different seed gives a different sequence
        x2, _, _ = Sequence(scene, SEED0 + 6, 8000).scan(3)
        assert not np.array_equal(x, x2)


def test_motion_composes_poses():
    seq = Sequence(ROOM, SEED0, 1000)
    R0, t0 = seq.pose(4)
    R1, t1 = seq.pose(5)
    dR, dp = seq.motion(5)
    # odometry increment is noisy (2 cm / 0.2 deg per axis): it lands near the next pose
    Rp, tp = R0 @ dR, t0 + R0 @ dp
    assert np.linalg.norm(tp - t1) < 0.2
    assert np.arccos(np.clip((np.trace(Rp.T @ R1) - 1) / 2, -1, 1)) < 0.03
