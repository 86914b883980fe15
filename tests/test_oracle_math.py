"""Pins the oracle's small math routines against independent libraries and
closed forms (CPU only).

  eigh3         numpy.linalg.eigh (values, vectors up to sign, reconstruction)
  invert        numpy.linalg.inv; singular input is reported
  so3_exp/log   scipy Rotation; exp(log(R)) == R near 0 and near pi
  left_jacobian integral form  J_l(w) = int_0^1 exp(s [w]x) ds  (scipy expm)
  calc_body_cov eigen-structure: sigma_r^2 along the ray, r^2 sin^2(theta)
                twice across it; the z == 0 -> 0.001 mutation (commons.cpp:20-21)
  boxplus/minus x [+] 0 == x; (x [+] d) [-] x == d; SO3 block is R exp(dr)
                (ieskf.cpp:11-55)
"""
import numpy as np
import pytest
from scipy.linalg import expm
from scipy.spatial.transform import Rotation

import oracle

RNG = np.random.default_rng(20261202)


def hat(v):
    return np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]], float)


# ----------------------------------------------------------------- eigh3
def _sym_cases():
    cases = []
    for _ in range(40):
        A = RNG.normal(size=(3, 3))
        cases.append(A + A.T)
    # covariance of exactly / nearly coplanar points (the plane-fit regime)
    for eps in (0.0, 1e-9, 1e-4):
        p = RNG.uniform(-0.25, 0.25, size=(50, 3))
        p[:, 2] = 0.1 + eps * RNG.normal(size=50)
        c = np.cov(p.T, bias=True)
        cases.append(c)
    # repeated eigenvalues, diagonal, zero
    cases += [np.eye(3), np.diag([3.0, 1.0, 2.0]), np.zeros((3, 3)), np.diag([1.0, 1.0, 1e-12])]
    Q = Rotation.from_rotvec([0.3, -0.2, 0.9]).as_matrix()
    cases.append(Q @ np.diag([1e-6, 0.02, 0.02]) @ Q.T)
    return cases


@pytest.mark.parametrize("i", range(len(_sym_cases())))
def test_eigh3_matches_numpy(i):
    A = _sym_cases()[i]
    ev, V = oracle.eigh3(A)
    ev_np, V_np = np.linalg.eigh(A)
    scale = max(1.0, np.abs(ev_np).max())
    np.testing.assert_allclose(ev, ev_np, atol=1e-13 * scale)
    assert np.all(np.diff(ev) >= 0), "eigenvalues must be ascending"
    np.testing.assert_allclose(V.T @ V, np.eye(3), atol=1e-13)
    np.testing.assert_allclose(V @ np.diag(ev) @ V.T, A, atol=1e-13 * scale)
    # eigenvectors of well separated eigenvalues agree with numpy up to sign
    gaps = np.diff(ev_np)
    for c in range(3):
        lo = gaps[c - 1] if c > 0 else np.inf
        hi = gaps[c] if c < 2 else np.inf
        if min(lo, hi) > 1e-6 * scale:
            assert abs(abs(V[:, c] @ V_np[:, c]) - 1.0) < 1e-9


# ----------------------------------------------------------------- invert
def test_invert_matches_numpy():
    for n in (3, 12, 23):
        for _ in range(5):
            A = RNG.normal(size=(n, n)) + n * np.eye(n)
            np.testing.assert_allclose(oracle.invert(A), np.linalg.inv(A), rtol=1e-11, atol=1e-13)
    # a covariance with entries spanning 1e-10 .. 1 (the IESKF P regime)
    d = np.logspace(-10, 0, 23)
    Q = np.linalg.qr(RNG.normal(size=(23, 23)))[0]
    P = Q @ np.diag(d) @ Q.T
    Pi = oracle.invert(P)
    np.testing.assert_allclose(Pi @ P, np.eye(23), atol=1e-5)
    np.testing.assert_allclose(Pi, np.linalg.inv(P), rtol=1e-4, atol=1e-4 * np.abs(Pi).max())


def test_invert_reports_singular():
    A = np.ones((4, 4))
    with pytest.raises(np.linalg.LinAlgError):
        oracle.invert(A)


# ----------------------------------------------------------------- SO3
def _rotvecs():
    out = [np.zeros(3), np.array([1e-12, 0, 0]), np.array([0, 3e-9, -1e-9]), np.array([0, 0, np.pi - 1e-9])]
    for _ in range(30):
        ax = RNG.normal(size=3)
        ax /= np.linalg.norm(ax)
        out.append(ax * RNG.uniform(0, np.pi - 1e-3))
    for ang in (1e-6, 0.1, 1.0, 3.0, np.pi - 1e-6):
        out.append(np.array([0.6, -0.8, 0.0]) * ang)
    return out


def test_so3_exp_matches_scipy():
    for w in _rotvecs():
        np.testing.assert_allclose(oracle.so3_exp(w), Rotation.from_rotvec(w).as_matrix(), atol=1e-14)


def test_so3_log_matches_scipy_and_inverts_exp():
    for w in _rotvecs():
        R = Rotation.from_rotvec(w).as_matrix()
        lw = oracle.so3_log(R)
        if np.linalg.norm(w) < np.pi - 1e-5:
            np.testing.assert_allclose(lw, Rotation.from_matrix(R).as_rotvec(), atol=1e-12)
            np.testing.assert_allclose(lw, w, atol=1e-11)
        # exp(log(R)) == R everywhere, including the pi boundary (sign-ambiguous there)
        np.testing.assert_allclose(oracle.so3_exp(lw), R, atol=1e-12)


def test_left_jacobian_is_integral_of_exp():
    s = np.linspace(0, 1, 401)
    wts = np.ones_like(s)
    wts[1:-1:2], wts[2:-1:2] = 4, 2  # Simpson
    wts *= (s[1] - s[0]) / 3
    for w in _rotvecs()[:20]:
        Jl = sum(wi * expm(si * hat(w)) for si, wi in zip(s, wts))
        np.testing.assert_allclose(oracle.left_jacobian(w), Jl, atol=1e-11)
        # exp(w + d) ~ exp(J_l(w) d) exp(w) to first order
        d = 1e-6 * RNG.normal(size=3)
        lhs = oracle.so3_exp(w + d)
        rhs = oracle.so3_exp(oracle.left_jacobian(w) @ d) @ oracle.so3_exp(w)
        assert np.abs(lhs - rhs).max() < 1e-11


# ----------------------------------------------------------------- calcBodyCov
@pytest.mark.parametrize(
    "p",
    [(1.0, 2.0, 3.0), (10.0, -4.0, 0.5), (-0.3, 0.1, -2.0), (25.0, 30.0, -1.0), (0.2, 0.0, 0.1), (3.0, -1.0, 1e-3)],
)
def test_calc_body_cov_eigenstructure(p):
    sr, deg = 0.04, 0.1
    q, C = oracle.calc_body_cov(p, sr, deg)
    np.testing.assert_array_equal(q, np.array(p, float))
    r = np.linalg.norm(q)
    d = q / r
    np.testing.assert_allclose(C, C.T, atol=1e-18)
    np.testing.assert_allclose(C @ d, sr**2 * d, rtol=1e-10, atol=1e-16)
    ev = np.linalg.eigvalsh(C)
    across = r**2 * np.sin(np.deg2rad(deg)) ** 2
    np.testing.assert_allclose(sorted(ev), sorted([sr**2, across, across]), rtol=1e-9)
    # closed form: sigma_r^2 d d^T + r^2 sin^2(theta) (I - d d^T)
    np.testing.assert_allclose(C, sr**2 * np.outer(d, d) + across * (np.eye(3) - np.outer(d, d)), atol=1e-15)


def test_calc_body_cov_mutates_zero_z():
    q, C = oracle.calc_body_cov((1.0, 2.0, 0.0))
    assert q[2] == 0.001
    d = q / np.linalg.norm(q)
    np.testing.assert_allclose(C @ d, 0.04**2 * d, rtol=1e-10)


# ----------------------------------------------------------------- state [+] / [-]
def _state(rng):
    R = Rotation.from_rotvec(rng.normal(size=3)).as_matrix()
    Re = Rotation.from_rotvec(0.1 * rng.normal(size=3)).as_matrix()
    g = rng.normal(size=3)
    g = -9.81 * np.array([0.05 * g[0], 0.05 * g[1], 1.0]) / np.linalg.norm([0.05 * g[0], 0.05 * g[1], 1.0])
    s = np.concatenate([rng.normal(size=3), R.T.ravel(), Re.T.ravel(), rng.normal(size=3) * 0.1,
                        rng.normal(size=3), 1e-3 * rng.normal(size=3), 1e-2 * rng.normal(size=3), g])
    return s


def test_boxplus_zero_is_identity_and_minus_inverts_plus():
    rng = np.random.default_rng(7)
    for _ in range(10):
        s = _state(rng)
        np.testing.assert_allclose(oracle.boxplus(s, np.zeros(23)), s, atol=1e-15)
        np.testing.assert_allclose(oracle.boxminus(s, s), np.zeros(23), atol=1e-12)
        d = np.concatenate([rng.normal(size=21) * 0.1, rng.normal(size=2) * 0.05])
        s2 = oracle.boxplus(s, d)
        np.testing.assert_allclose(oracle.boxminus(s2, s), d, atol=1e-10)
        # rotation blocks are right-perturbed: R' = R exp(d_r)
        R = s[3:12].reshape(3, 3).T
        np.testing.assert_allclose(s2[3:12].reshape(3, 3).T, R @ Rotation.from_rotvec(d[3:6]).as_matrix(), atol=1e-14)
        np.testing.assert_allclose(s2[0:3], s[0:3] + d[0:3], atol=0)
        # gravity stays on the sphere of radius 9.81
        assert abs(np.linalg.norm(s2[33:36]) - 9.81) < 1e-12
