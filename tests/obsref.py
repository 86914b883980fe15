"""numpy restatement of one observation pass (lio_builder.cpp:250-311) used
by the parity tests as the per-point reference: slot update with the stale
semantics of buildResidual (voxel_map.cpp:258-276, a miss leaves the slot
as it was), the Jacobian row and weight of every valid slot, and H, b.

Everything here is float64 and written from the reference's formulas, not
from the oracle's code; extrinsic rotation/translation are honoured.
"""
import numpy as np

VS = 0.5


def hat(v):
    return np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]], float)


def body_cov(p, sr=0.04, deg=0.1):
    """calcBodyCov (commons.cpp:18-45) in closed form: sigma_r^2 along the
    ray, r^2 sin^2(angle) across it."""
    r = np.linalg.norm(p)
    d = p / r
    a = r**2 * np.sin(np.deg2rad(deg)) ** 2
    return sr**2 * np.outer(d, d) + a * (np.eye(3) - np.outer(d, d))


def lidar_points(x, y, z):
    """point_lidar after calcBodyCov's z == 0 -> 0.001 mutation (commons.cpp:20-23)."""
    p = np.stack([np.asarray(x, np.float64), np.asarray(y, np.float64), np.asarray(z, np.float64)], 1)
    p[p[:, 2] == 0.0, 2] = 0.001
    return p


def state_parts(s):
    rot = np.asarray(s[3:12]).reshape(3, 3).T
    rext = np.asarray(s[12:21]).reshape(3, 3).T
    return rot, rext, np.asarray(s[0:3]), np.asarray(s[21:24])


def jac_weight(s, p, n, estimate_ext=False):
    """Jacobian row (12) and weight of one valid slot at state s
    (lio_builder.cpp:283-306): r_cov = n^T R_wl C_l R_wl^T n, clamped."""
    rot, rext, pos, pext = state_parts(s)
    rwl = rot @ rext
    rc = n @ rwl @ body_cov(p) @ rwl.T @ n
    w = 5000.0 if rc < 2e-4 else 1.0 / rc
    J = np.zeros(12)
    J[0:3] = n
    J[3:6] = -n @ rot @ hat(rext @ p + pext)
    if estimate_ext:
        J[6:9] = -n @ rwl @ hat(p)
        J[9:12] = n @ rot
    return J, w


def observe(s, P, pts, mp, prev=None, estimate_ext=False):
    """One pass at state s with prior P over lidar points `pts` (N x 3,
    already mutated) against the exported map `mp`.  prev = (valid, resid,
    normals) of the slots before the pass (None: all invalid).  Returns the
    new slots, a hit mask (key found), the 3-sigma margin of every fresh
    decision, and H, b."""
    rot, rext, pos, pext = state_parts(s)
    rwl = rot @ rext
    pwl = rot @ pext + pos
    vox = {tuple(k): i for i, k in enumerate(mp["keys"].tolist())}
    N = len(pts)
    if prev is None:
        valid, resid, norms = np.zeros(N, bool), np.zeros(N), np.zeros((N, 3))
    else:
        valid, resid, norms = (np.array(a, copy=True) for a in prev)
        valid = valid.astype(bool)
    pw = pts @ rwl.T + pwl
    keys = np.floor(pw / VS).astype(np.int64)
    Prr, Ppp = P[3:6, 3:6], P[0:3, 0:3]
    hit = np.zeros(N, bool)
    margin = np.full(N, np.inf)
    for i, k in enumerate(map(tuple, keys.tolist())):
        j = vox.get(k)
        if j is None:
            continue  # stale: the slot keeps what it held
        hit[i] = True
        valid[i] = False
        if not mp["ints"][j, 1] & 2:
            continue
        m, n = mp["dbl"][j, 0:3], mp["dbl"][j, 3:6]
        r = n @ (pw[i] - m)
        hp = hat(pts[i])
        Cw = rwl @ body_cov(pts[i]) @ rwl.T + hp @ Prr @ hp.T + Ppp
        sig = 3.0 * np.sqrt(n @ Cw @ n)
        valid[i] = abs(r) < sig
        margin[i] = abs(abs(r) - sig) / sig
        resid[i], norms[i] = r, n
    H, b = hb(s, pts, valid, resid, norms, estimate_ext)
    return valid, resid, norms, hit, margin, H, b


def hb(s, pts, valid, resid, norms, estimate_ext=False):
    """H = sum w J^T J, b = sum w J^T r over the valid slots."""
    H = np.zeros((12, 12))
    b = np.zeros(12)
    for i in np.flatnonzero(valid):
        J, w = jac_weight(s, pts[i], norms[i], estimate_ext)
        H += w * np.outer(J, J)
        b += w * J * resid[i]
    return H, b
