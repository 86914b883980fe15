"""Pins the oracle's observation build (lio_builder.cpp:250-311) and IESKF
update (ieskf.cpp:125-156) on the CPU, on a planar numpy scene.

  observe   residual = n.(p_w - m) of the point's voxel plane; validity
            |r| < 3 sqrt(n^T C_w n); H = sum w J^T J, b = sum w J^T r with
            w = 1/r_cov clamped to 5000 — all recomputed with numpy;
            J checked against central differences of r(x [+] e) (pins the
            right-perturbation convention and the sign); H symmetric PSD
  update    empty map: one iteration, state and P unchanged;
            one iteration with observations equals the Gauss-Newton / Kalman
            step  delta = -(P^-1 + H)^-1 b,  P+ = L (P^-1 + H)^-1 L^T  (numpy);
            iterations never exceed max_iter
"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import oracle
from voxelmapplus_fastlio2_amd import LIOConfig, initial_covariance, state_vector

VS = 0.5


def hat(v):
    return np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]], float)


def body_cov(p, sr=0.04, deg=0.1):
    r = np.linalg.norm(p)
    d = p / r
    a = r**2 * np.sin(np.deg2rad(deg)) ** 2
    return sr**2 * np.outer(d, d) + a * (np.eye(3) - np.outer(d, d))


def scene(n=30000, seed=3):
    """World points on the six faces of an 8 x 6 x 3 m room and one 30-degree ramp."""
    rng = np.random.default_rng(seed)
    faces = []
    m = n // 7
    u, v = rng.uniform(size=(2, 7, m))
    faces.append(np.stack([u[0] * 8 - 3, v[0] * 6 - 2, np.zeros(m)], 1))  # floor
    faces.append(np.stack([u[1] * 8 - 3, v[1] * 6 - 2, np.full(m, 3.0)], 1))  # ceiling
    faces.append(np.stack([np.full(m, -3.0), u[2] * 6 - 2, v[2] * 3], 1))
    faces.append(np.stack([np.full(m, 5.0), u[3] * 6 - 2, v[3] * 3], 1))
    faces.append(np.stack([u[4] * 8 - 3, np.full(m, -2.0), v[4] * 3], 1))
    faces.append(np.stack([u[5] * 8 - 3, np.full(m, 4.0), v[5] * 3], 1))
    y = u[6] * 2.0
    faces.append(np.stack([v[6] * 2 + 1, y + 1, y * np.tan(np.deg2rad(30)) + 0.3], 1))
    pw = np.concatenate(faces)
    pw += 0.003 * rng.normal(size=pw.shape)
    return pw


@pytest.fixture(scope="module")
def setup():
    R0 = Rotation.from_rotvec([0.05, -0.03, 0.7]).as_matrix()
    t0 = np.array([1.1, 0.4, 1.3])
    pw = scene()
    pl = ((pw - t0) @ R0).astype(np.float32)
    s0 = state_vector(R0, t0)
    P0 = initial_covariance()
    return R0, t0, pl, s0, P0


def built_oracle(setup, **kw):
    R0, t0, pl, s0, P0 = setup
    o = oracle.Oracle(LIOConfig(max_scan_points=len(pl), **kw))
    o.set_state(s0, P0)
    o.build(pl[:, 0], pl[:, 1], pl[:, 2])
    o.set_points(pl[:, 0], pl[:, 1], pl[:, 2])
    return o


def perturbed(s0):
    d = np.zeros(23)
    d[0:3] = [0.02, -0.015, 0.01]
    d[3:6] = [0.004, -0.003, 0.006]
    return oracle.boxplus(s0, d)


def expected_observation(o, s, P, pl):
    """numpy restatement of pass A + pass B for rot_ext = I, pos_ext = 0."""
    R = s[3:12].reshape(3, 3).T
    t = s[0:3]
    mp = o.export_map()
    vox = {tuple(k): i for i, k in enumerate(mp["keys"].tolist())}
    plq = pl.astype(np.float64)
    pw = plq @ R.T + t
    keys = np.floor(pw / VS).astype(np.int64)
    Prr, Ppp = P[3:6, 3:6], P[0:3, 0:3]
    valid = np.zeros(len(pl), bool)
    resid = np.zeros(len(pl))
    norms = np.zeros((len(pl), 3))
    margin = np.full(len(pl), np.inf)
    for i, k in enumerate(map(tuple, keys.tolist())):
        j = vox.get(k)
        if j is None or not mp["ints"][j, 1] & 2:
            continue
        m, n = mp["dbl"][j, 0:3], mp["dbl"][j, 3:6]
        r = n @ (pw[i] - m)
        Cl = body_cov(plq[i])
        hp = hat(plq[i])
        Cw = R @ Cl @ R.T + hp @ Prr @ hp.T + Ppp
        sig = 3.0 * np.sqrt(n @ Cw @ n)
        valid[i] = abs(r) < sig
        margin[i] = abs(abs(r) - sig) / sig
        resid[i], norms[i] = r, n
    H = np.zeros((12, 12))
    b = np.zeros(12)
    J = np.zeros((len(pl), 12))
    w = np.zeros(len(pl))
    for i in np.flatnonzero(valid):
        n = norms[i]
        Cl = body_cov(plq[i])
        r_cov = n @ R @ Cl @ R.T @ n
        w[i] = 5000.0 if r_cov < 2e-4 else 1.0 / r_cov
        J[i, 0:3] = n
        J[i, 3:6] = -n @ R @ hat(plq[i])
    H = (J * w[:, None]).T @ J
    b = J.T @ (w * resid)
    return valid, resid, norms, margin, H, b, J


def test_observe_matches_numpy_restatement(setup):
    R0, t0, pl, s0, P0 = setup
    o = built_oracle(setup)
    s1 = perturbed(s0)
    H, b, eff = o.observe(s1)
    valid, resid, norms, margin, He, be, J = expected_observation(o, s1, P0, pl)
    v, r, nrm = o.export_slots(len(pl))
    clear = margin > 1e-9
    np.testing.assert_array_equal(v[clear], valid[clear])
    assert eff == int(v.sum()) and eff > 0.5 * len(pl)
    np.testing.assert_allclose(r[v], resid[v], atol=1e-12)
    np.testing.assert_allclose(nrm[v], norms[v], atol=0)
    assert np.all(v == valid), "no validity decision sits on the 3-sigma boundary here"
    np.testing.assert_allclose(H, He, rtol=1e-9, atol=1e-9 * np.abs(He).max())
    np.testing.assert_allclose(b, be, rtol=1e-9, atol=1e-9 * np.abs(be).max())
    np.testing.assert_allclose(H, H.T, atol=0)
    assert np.linalg.eigvalsh(H[:6, :6]).min() > 0
    assert np.all(H[6:, :] == 0) and np.all(b[6:] == 0), "estimate_ext = false leaves the extrinsic block empty"
    # the clamp branch (w = 5000 for r_cov < 2e-4) is exercised
    Rm = s1[3:12].reshape(3, 3).T
    rc = np.array([n @ Rm @ body_cov(p.astype(float)) @ Rm.T @ n for p, n in zip(pl[v], nrm[v])])
    assert (rc < 2e-4).any() and (rc >= 2e-4).any()


def test_observe_jacobian_matches_finite_differences(setup):
    R0, t0, pl, s0, P0 = setup
    o = built_oracle(setup)
    s1 = perturbed(s0)
    o.observe(s1)
    v, r, nrm = o.export_slots(len(pl))
    _, _, _, _, _, _, J = expected_observation(o, s1, P0, pl)
    mp = o.export_map()
    vox = {tuple(k): i for i, k in enumerate(mp["keys"].tolist())}
    idx = np.flatnonzero(v)[::997]
    plq = pl.astype(np.float64)

    def resid(s, i, m, n):
        R = s[3:12].reshape(3, 3).T
        return n @ (R @ plq[i] + s[0:3] - m)

    eps = 1e-6
    for i in idx:
        pw = s1[3:12].reshape(3, 3).T @ plq[i] + s1[0:3]
        j = vox[tuple(np.floor(pw / VS).astype(np.int64).tolist())]
        m, n = mp["dbl"][j, 0:3], mp["dbl"][j, 3:6]
        for k in range(6):
            e = np.zeros(23)
            e[k] = eps
            fd = (resid(oracle.boxplus(s1, e), i, m, n) - resid(oracle.boxplus(s1, -e), i, m, n)) / (2 * eps)
            assert abs(fd - J[i, k]) < 1e-7 * max(1.0, abs(J[i, k])), (i, k, fd, J[i, k])


def test_update_without_observations_keeps_state_and_covariance(setup):
    R0, t0, pl, s0, P0 = setup
    o = oracle.Oracle(LIOConfig(max_scan_points=len(pl)))
    o.set_state(s0, P0)
    o.set_points(pl[:, 0], pl[:, 1], pl[:, 2])  # empty map: every lookup misses
    assert o.update() == 1
    s, P = o.get_state()
    np.testing.assert_array_equal(s, s0)
    np.testing.assert_allclose(P, P0, rtol=1e-12, atol=1e-16)


def test_one_iteration_is_the_gauss_newton_step(setup):
    R0, t0, pl, s0, P0 = setup
    o = built_oracle(setup, max_iter=1)
    s1 = perturbed(s0)
    o.set_state(s1, P0)
    H, b, _ = o.observe(s1)
    assert o.update() == 1
    s, P = o.get_state()
    Ht = np.linalg.inv(P0)
    Ht[:12, :12] += H
    bt = np.zeros(23)
    bt[:12] = b
    d = -np.linalg.solve(Ht, bt)
    np.testing.assert_allclose(s, oracle.boxplus(s1, d), atol=1e-12)
    L = np.eye(23)
    L[3:6, 3:6] = oracle.left_jacobian(d[3:6]).T
    L[6:9, 6:9] = oracle.left_jacobian(d[6:9]).T
    Pe = L @ np.linalg.inv(Ht) @ L.T
    np.testing.assert_allclose(P, Pe, rtol=1e-8, atol=1e-14)
    # the step moves the perturbed pose back towards the true one
    assert np.linalg.norm(s[0:3] - t0) < 0.5 * np.linalg.norm(s1[0:3] - t0)


def test_full_update_converges_to_true_pose(setup):
    R0, t0, pl, s0, P0 = setup
    o = built_oracle(setup)
    s1 = perturbed(s0)
    o.set_state(s1, P0)
    it = o.update()
    assert 1 <= it <= 5
    s, _ = o.get_state()
    assert np.linalg.norm(s[0:3] - t0) < 5e-3  # 3.5 cm start, 3 mm noise-limited
    dR = s[3:12].reshape(3, 3).T @ R0.T
    assert np.linalg.norm(Rotation.from_matrix(dR).as_rotvec()) < 1e-3  # 8 mrad start


def test_observe_with_extrinsic_matches_finite_differences(setup):
    """estimate_ext = true (lio_builder.cpp:283-307): all twelve Jacobian
    columns (position, rotation, extrinsic rotation, extrinsic translation)
    come from central differences of r(x [+] e) = n.(R (R_ext p + t_ext) + t - m)
    per valid slot, with a non-identity extrinsic; H = sum w J^T J and
    b = sum w J^T r are rebuilt from them."""
    R0, t0, pl, s0, P0 = setup
    Rx = Rotation.from_rotvec([0.02, -0.01, 0.03]).as_matrix()
    tx = np.array([0.05, -0.02, 0.1])
    R = R0 @ Rx.T  # keeps the LiDAR pose R0, t0 the points were generated at
    t = t0 - R @ tx
    s = state_vector(R, t)
    s[12:21] = Rx.T.reshape(-1)
    s[21:24] = tx
    o = oracle.Oracle(LIOConfig(max_scan_points=len(pl), estimate_ext=1))
    o.set_state(s, P0)
    o.build(pl[:, 0], pl[:, 1], pl[:, 2])
    o.set_points(pl[:, 0], pl[:, 1], pl[:, 2])
    d = np.zeros(23)
    d[0:12] = [0.01, -0.008, 0.005, 0.003, -0.002, 0.004, 0.002, 0.001, -0.003, 0.01, 0.005, -0.004]
    s1 = oracle.boxplus(s, d)
    H, b, eff = o.observe(s1)
    v, r, nrm = o.export_slots(len(pl))
    mp = o.export_map()
    vox = {tuple(k): i for i, k in enumerate(mp["keys"].tolist())}
    plq = pl.astype(np.float64)

    def world(st):
        Rs, Re = st[3:12].reshape(3, 3).T, st[12:21].reshape(3, 3).T
        return plq @ (Rs @ Re).T + (Rs @ st[21:24] + st[0:3])

    pw = world(s1)
    idx = np.flatnonzero(v)
    keys = np.floor(pw[idx] / VS).astype(np.int64)
    rows = np.array([vox[tuple(k)] for k in keys.tolist()])
    m, n = mp["dbl"][rows, 0:3], mp["dbl"][rows, 3:6]
    np.testing.assert_allclose(np.sum(n * (pw[idx] - m), 1), r[idx], atol=1e-12)
    eps = 1e-6
    J = np.zeros((len(idx), 12))
    for k in range(12):
        e = np.zeros(23)
        e[k] = eps
        rp = np.sum(n * (world(oracle.boxplus(s1, e))[idx] - m), 1)
        rm = np.sum(n * (world(oracle.boxplus(s1, -e))[idx] - m), 1)
        J[:, k] = (rp - rm) / (2 * eps)
    Rwl = s1[3:12].reshape(3, 3).T @ s1[12:21].reshape(3, 3).T
    rc = np.array([nn @ Rwl @ body_cov(p) @ Rwl.T @ nn for p, nn in zip(plq[idx], n)])
    w = np.where(rc < 2e-4, 5000.0, 1.0 / rc)
    He = (J * w[:, None]).T @ J
    be = J.T @ (w * r[idx])
    assert eff == len(idx) and eff > 0.5 * len(pl)
    assert np.abs(H[6:12, 6:12]).max() > 1e-3 * np.abs(H).max(), "extrinsic block populated"
    np.testing.assert_allclose(H, He, rtol=1e-6, atol=1e-7 * np.abs(He).max())
    np.testing.assert_allclose(b, be, rtol=1e-6, atol=1e-7 * np.abs(be).max())
