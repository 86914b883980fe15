"""Pins of the oracle's IMU side on the CPU (the LIOBuilder preprocessing in
front of the registration path, lio_builder.cpp:28-153, ieskf.cpp:101-123,
and the scan filter of lio_builder.cpp:13-14, 215-223).

  predict        a static IMU reading (gyro = bg, acc = R^T (-g) + ba) leaves
                 the state unchanged; a constant body rate composes to the
                 closed-form rotation R0 exp(w T); a constant world
                 acceleration gives the Euler closed form v0 + N a dt,
                 p0 + N v0 dt + a dt^2 N (N - 1) / 2; every column of F, read
                 back from P' = F e_i e_i^T F^T with Q = 0, equals the
                 central-difference Jacobian of the step in boxplus/boxminus
                 coordinates; G Q G^T of the gyro and accelerometer noise
                 equals the covariance of central differences with respect
                 to those inputs
  undistortion   a resting sensor leaves the points unchanged (to fp32);
                 a constant yaw rate gives p' = exp(-w (T - t)) p and a
                 constant velocity p' = p + v (t - T); a point taken at time
                 0 stays as it was; the earliest point, if later than the
                 first IMU sample, is compensated again for every earlier
                 IMU interval (the loop of lio_builder.cpp:126-152)
  IMU init       below imu_init_num samples nothing is set; then bg = mean
                 gyro, |g| = 9.81 along -z, and rot takes the mean specific
                 force direction onto +z by the smallest rotation
  VoxelGrid      an independent numpy grouping: leaves ordered by the
                 linear index of fp32 floor(p / leaf), each emitted as the
                 fp32 mean of its points in input order; points already on
                 distinct leaves come back unchanged, reordered
"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import oracle
from voxelmapplus_fastlio2_amd import LIOConfig, initial_covariance, state_vector

G = 9.81


def fresh(s, P=None, noise=(0.01, 0.01, 1e-4, 1e-4)):
    o = oracle.Oracle(LIOConfig(max_scan_points=16))
    o.imu_config(*noise)
    o.set_state(s, initial_covariance() if P is None else P)
    return o


def moving_state(seed=0):
    rng = np.random.default_rng(seed)
    R = Rotation.from_rotvec(rng.normal(size=3) * 0.4).as_matrix()
    s = state_vector(R, rng.normal(size=3))
    s[24:27] = rng.normal(size=3)          # vel
    s[27:30] = rng.normal(size=3) * 0.01   # bg
    s[30:33] = rng.normal(size=3) * 0.05   # ba
    g = np.array([0.3, -0.2, -9.8])
    s[33:36] = g / np.linalg.norm(g) * G
    return s


def test_predict_static_reading_keeps_the_state():
    s = moving_state(1)
    s[24:27] = 0.0
    R = s[3:12].reshape(3, 3).T
    acc = R.T @ (-s[33:36]) + s[30:33]
    gyro = s[27:30].copy()
    o = fresh(s)
    for _ in range(20):
        o.predict(acc, gyro, 0.005)
    s1, _ = o.get_state()
    np.testing.assert_allclose(s1, s, atol=1e-13)


def test_predict_constant_rate_and_acceleration_closed_forms():
    s = state_vector(np.eye(3), np.array([1.0, 2.0, 3.0]))
    s[24:27] = [0.5, -0.2, 0.1]
    w = np.array([0.02, -0.5, 0.3])
    o = fresh(s)
    N, dt = 40, 0.005
    # level sensor, constant world acceleration a while the body turns:
    # feed acc = R_k^T (a - g) at every step from the state's own rotation
    a = np.array([0.3, 0.1, -0.2])
    for k in range(N):
        Rk = Rotation.from_rotvec(w * dt * k).as_matrix()
        o.predict(Rk.T @ (a - s[33:36]), w, dt)
    s1, _ = o.get_state()
    R1 = s1[3:12].reshape(3, 3).T
    np.testing.assert_allclose(R1, Rotation.from_rotvec(w * dt * N).as_matrix(), atol=1e-12)
    np.testing.assert_allclose(s1[24:27], s[24:27] + N * a * dt, atol=1e-9)
    np.testing.assert_allclose(s1[0:3], s[0:3] + N * s[24:27] * dt + a * dt * dt * N * (N - 1) / 2, atol=1e-9)


def step(s, acc, gyro, dt):
    o = fresh(s, noise=(0.0, 0.0, 0.0, 0.0))
    o.predict(acc, gyro, dt)
    return o.get_state()[0]


def test_predict_F_matches_finite_differences():
    s = moving_state(2)
    acc = np.array([0.4, -0.3, 9.6])
    gyro = np.array([0.3, -0.2, 0.5])
    dt = 0.01
    f0 = step(s, acc, gyro, dt)
    eps = 1e-6
    Jfd = np.zeros((23, 23))
    for i in range(23):
        e = np.zeros(23)
        e[i] = eps
        fp = step(oracle.boxplus(s, e), acc, gyro, dt)
        fm = step(oracle.boxplus(s, -e), acc, gyro, dt)
        # both sides in the chart at f(x), so the central difference is O(eps^2)
        Jfd[:, i] = (oracle.boxminus(fp, f0) - oracle.boxminus(fm, f0)) / (2 * eps)
    assert np.abs(oracle.boxminus(step(s, acc, gyro, dt), f0)).max() == 0.0
    for i in range(23):
        P = np.zeros((23, 23))
        P[i, i] = 1.0
        o = fresh(s, P, noise=(0.0, 0.0, 0.0, 0.0))
        o.predict(acc, gyro, dt)
        FF = o.get_state()[1]  # f_i f_i^T
        np.testing.assert_allclose(FF, np.outer(Jfd[:, i], Jfd[:, i]), atol=2e-8, err_msg=f"column {i}")


def test_predict_GQG_matches_input_differences():
    s = moving_state(3)
    acc = np.array([0.2, 0.5, 9.7])
    gyro = np.array([-0.4, 0.1, 0.3])
    dt = 0.01
    eps = 1e-6
    f0 = step(s, acc, gyro, dt)
    Jin = np.zeros((23, 6))
    for i in range(6):
        e = np.zeros(6)
        e[i] = eps
        fp = step(s, acc + e[3:6], gyro + e[0:3], dt)
        fm = step(s, acc - e[3:6], gyro - e[0:3], dt)
        Jin[:, i] = (oracle.boxminus(fp, f0) - oracle.boxminus(fm, f0)) / (2 * eps)
    ng, na = 0.013, 0.021
    o = fresh(s, np.zeros((23, 23)), noise=(na, ng, 0.0, 0.0))
    o.predict(acc, gyro, dt)
    GQG = o.get_state()[1]
    Qin = np.diag([ng] * 3 + [na] * 3)
    np.testing.assert_allclose(GQG, Jin @ Qin @ Jin.T, atol=1e-9)
    # the bias random walks enter as I dt on their own blocks
    o = fresh(s, np.zeros((23, 23)), noise=(0.0, 0.0, 0.005, 0.007))
    o.predict(acc, gyro, dt)
    P = o.get_state()[1]
    np.testing.assert_allclose(np.diag(P)[15:18], 0.007 * dt * dt, rtol=1e-12)
    np.testing.assert_allclose(np.diag(P)[18:21], 0.005 * dt * dt, rtol=1e-12)


def imu_stream(t0, t1, acc, gyro, dt=0.005):
    t = np.arange(int(round(t0 / dt)), int(round(t1 / dt))) * dt
    return t, np.tile(acc, (len(t), 1)), np.tile(gyro, (len(t), 1))


def initialised(acc, gyro, t_end=0.1):
    """An oracle through IMU_INIT on a static package ending at t_end."""
    o = oracle.Oracle(LIOConfig(max_scan_points=64))
    o.imu_config()
    st, _ = o.process(np.zeros(1), np.zeros(1), np.ones(1), np.zeros(1), imu_stream(0.0, t_end, acc, gyro), 0.0, t_end)
    assert st == 1
    return o


def test_undistort_static_leaves_points():
    o = initialised([0.0, 0.0, G], [0.0, 0.0, 0.0])
    rng = np.random.default_rng(4)
    x, y, z = (rng.uniform(-20, 20, 500).astype(np.float32) for _ in range(3))
    t = np.sort(rng.uniform(0.001, 99.0, 500)).astype(np.float32)
    ux, uy, uz, _ = o.undistort(x, y, z, t, imu_stream(0.1, 0.2, [0.0, 0.0, G], [0.0, 0.0, 0.0]), 0.1, 0.2)
    np.testing.assert_allclose(np.stack([ux, uy, uz]), np.stack([x, y, z]), atol=2e-5)


def test_undistort_constant_yaw_rate_and_velocity():
    wz = 0.8
    o = initialised([0.0, 0.0, G], [0.0, 0.0, 0.0])
    s, P = o.get_state()
    v = np.array([1.5, -0.5, 0.2])
    s[24:27] = v
    o.set_state(s, P)
    rng = np.random.default_rng(5)
    n = 400
    x, y, z = (rng.uniform(-20, 20, n).astype(np.float32) for _ in range(3))
    t = np.sort(rng.uniform(0.0004, 99.0, n)).astype(np.float32)
    t[0] = 0.0  # a point at the cloud's start time
    end = 0.1 + float(t[-1]) / 1000.0
    imu = imu_stream(0.1, end, [0.0, 0.0, G], [0.0, 0.0, wz])
    ux, uy, uz, ut = o.undistort(x, y, z, t, imu, 0.1, end)
    T = float(t[-1]) / 1000.0
    p = np.stack([x, y, z], 1).astype(np.float64)
    dtp = T - t.astype(np.float64) / 1000.0
    Rz = Rotation.from_rotvec(np.outer(-wz * dtp, [0, 0, 1])).as_matrix()
    # R_cur^T (R(t) p + p(t) - p_cur): rotation exp(-w (T - t)) and the
    # velocity carried back into the end frame
    R_end = Rotation.from_rotvec([0, 0, wz * T]).as_matrix()
    exp = np.einsum("nij,nj->ni", Rz, p) - (R_end.T @ (np.outer(dtp, v)).T).T
    got = np.stack([ux, uy, uz], 1).astype(np.float64)
    np.testing.assert_allclose(got[1:], exp[1:], atol=5e-5)
    np.testing.assert_array_equal(got[0], p[0])  # time 0: not compensated


def test_undistort_recompensates_the_earliest_point():
    """The earliest point, taken after the first IMU sample of the sweep, is
    transformed once for its own interval and once more for every earlier
    interval (the inner loop breaks at the first point without moving)."""
    wz = 0.6
    o = initialised([0.0, 0.0, G], [0.0, 0.0, 0.0])
    t = np.array([12.0, 30.0, 60.0, 95.0], np.float32)  # first point after samples at 0.1, 0.105, 0.11
    x = np.array([5.0, 4.0, -3.0, 2.0], np.float32)
    y = np.array([1.0, -2.0, 6.0, 3.0], np.float32)
    z = np.array([0.5, 0.4, -0.3, 1.0], np.float32)
    end = 0.1 + 0.095
    ux, uy, uz, _ = o.undistort(x, y, z, t, imu_stream(0.1, end, [0.0, 0.0, G], [0.0, 0.0, wz]), 0.1, end)
    poses = o.imu_poses()
    s, _ = o.get_state()
    Rc, pc = s[3:12].reshape(3, 3).T, s[0:3]

    def comp(q, h, tl):  # one compensation with head h and tail tl
        dt = float(tail_t) - h["offset"]
        Rp = h["rot"] @ Rotation.from_rotvec(tl["gyro"] * dt).as_matrix()
        pp = h["pos"] + h["vel"] * dt + 0.5 * tl["acc"] * dt * dt
        return Rc.T @ (Rp @ q + pp - pc)

    tail_t = np.float64(t[0]) / 1000.0
    offs = [q["offset"] for q in poses]
    k0 = max(i for i in range(len(poses) - 1) if offs[i] < tail_t)
    assert k0 >= 2
    q = np.array([x[0], y[0], z[0]], np.float64)
    for k in range(k0, -1, -1):
        q = comp(q, poses[k], poses[k + 1])
    np.testing.assert_allclose([ux[0], uy[0], uz[0]], q, atol=2e-5)
    once = comp(np.array([x[0], y[0], z[0]], np.float64), poses[k0], poses[k0 + 1])
    assert np.linalg.norm(q - once) > 1e-3


def test_imu_init_alignment():
    rng = np.random.default_rng(6)
    Rt = Rotation.from_rotvec([0.05, -0.08, 1.2]).as_matrix()
    f = Rt.T @ np.array([0.0, 0.0, 9.79])
    bg = np.array([0.003, -0.002, 0.001])
    o = oracle.Oracle(LIOConfig(max_scan_points=64))
    o.imu_config(imu_init_num=20)
    m = 12
    t = np.arange(m) * 0.005
    acc = f + 0.01 * rng.normal(size=(m, 3))
    gyro = bg + 0.001 * rng.normal(size=(m, 3))
    assert not o.initialize_imu((t, acc, gyro))
    t2 = (np.arange(m) + m) * 0.005
    acc2 = f + 0.01 * rng.normal(size=(m, 3))
    gyro2 = bg + 0.001 * rng.normal(size=(m, 3))
    assert o.initialize_imu((t2, acc2, gyro2))
    s, P = o.get_state()
    am = np.concatenate([acc, acc2]).mean(0)
    np.testing.assert_allclose(s[27:30], np.concatenate([gyro, gyro2]).mean(0), atol=1e-15)
    np.testing.assert_allclose(s[33:36], [0, 0, -G], atol=1e-15)
    R = s[3:12].reshape(3, 3).T
    u = am / np.linalg.norm(am)
    np.testing.assert_allclose(R @ (-u), [0, 0, -1], atol=1e-14)
    rv = Rotation.from_matrix(R).as_rotvec()
    ax = np.cross(-u, [0, 0, -1])
    np.testing.assert_allclose(rv, ax / np.linalg.norm(ax) * np.arccos(np.dot(-u, [0, 0, -1])), atol=1e-12)
    np.testing.assert_allclose(P, initial_covariance(), atol=0)
    assert abs(o.imu_state()["gravity_norm"] - np.linalg.norm(am)) < 1e-14


def grouped_mean(x, y, z, leaf):
    inv = np.float32(1.0) / np.float32(leaf)
    p = np.stack([x, y, z]).astype(np.float32)
    minb = np.floor(p.min(1) * inv).astype(np.int64)
    maxb = np.floor(p.max(1) * inv).astype(np.int64)
    div = maxb - minb + 1
    ijk = (np.floor(p * inv) - minb.astype(np.float32)[:, None]).astype(np.int64)
    lin = ijk[0] + ijk[1] * div[0] + ijk[2] * div[0] * div[1]
    keys, inverse = np.unique(lin, return_inverse=True)
    out = np.zeros((3, len(keys)), np.float32)
    cnt = np.zeros(len(keys), np.int64)
    for i in range(len(x)):  # fp32 sums in input order
        out[:, inverse[i]] = (out[:, inverse[i]] + p[:, i]).astype(np.float32)
        cnt[inverse[i]] += 1
    return (out / cnt.astype(np.float32)).astype(np.float32)


@pytest.mark.parametrize("leaf", [0.1, 0.25])
def test_voxel_grid_matches_grouped_means(leaf):
    rng = np.random.default_rng(7)
    n = 6000
    p = rng.uniform([-8, -5, -1], [9, 4, 3], size=(n, 3)).astype(np.float32)
    p[n // 2:] = p[: n - n // 2] + rng.normal(scale=0.01, size=(n - n // 2, 3)).astype(np.float32)  # shared leaves
    got = oracle.voxel_grid(p[:, 0], p[:, 1], p[:, 2], leaf)
    exp = grouped_mean(p[:, 0], p[:, 1], p[:, 2], leaf)
    assert got.shape == exp.shape and got.shape[1] < n
    np.testing.assert_array_equal(got, exp)


def test_voxel_grid_on_leaf_centres_is_a_reordering():
    leaf = 0.5
    ii = np.array([[3, -2, 1], [0, 0, 0], [-4, 7, 2], [1, 1, -3], [2, -5, 0]])
    p = ((ii + 0.5) * leaf).astype(np.float32)
    got = oracle.voxel_grid(p[:, 0], p[:, 1], p[:, 2], leaf)
    order = np.lexsort((ii[:, 0], ii[:, 1], ii[:, 2]))  # z, then y, then x
    np.testing.assert_array_equal(got.T, p[order])
