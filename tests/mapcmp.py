"""Map / state comparison helpers shared by the parity tests."""
import numpy as np

import obsref


def rot_angle(Ra, Rb):
    d = Ra.T @ Rb
    return float(np.arccos(np.clip((np.trace(d) - 1.0) / 2.0, -1.0, 1.0)))


def state_err(sa, sb):
    """(position error m, rotation error rad) between two flat state vectors."""
    Ra = np.asarray(sa[3:12]).reshape(3, 3).T
    Rb = np.asarray(sb[3:12]).reshape(3, 3).T
    return float(np.linalg.norm(np.asarray(sa[0:3]) - np.asarray(sb[0:3]))), rot_angle(Ra, Rb)


def ulp_err(a, b):
    """Largest coordinate difference of two (3, n) point arrays in units of
    the fp32 spacing at each point's largest coordinate: the compensation
    rotates and translates the whole point in double and rounds once, so its
    error scales with the point's size, and a state that differs by ~1e-9
    between the two sides moves a rounding by one unit."""
    if a.size == 0:
        return 0.0
    mag = np.maximum(np.abs(a).max(axis=0), np.abs(b).max(axis=0)).astype(np.float32)
    u = np.spacing(mag).astype(np.float64)
    return float((np.abs(a.astype(np.float64) - b) / u).max())


def compare_maps(ref, got, mean_tol=1e-6, norm_tol=1e-5, tr_rtol=1e-4, report=None):
    """Keys in LRU order and point counts must match exactly; plane flags,
    group ids and plane parameters are compared with the margin cases
    (plane threshold / merge gates decided differently in the last bits)
    counted and returned rather than failed."""
    out = {}
    assert ref["keys"].shape == got["keys"].shape, (ref["keys"].shape, got["keys"].shape)
    assert np.array_equal(ref["keys"], got["keys"]), "LRU key order differs"
    assert np.array_equal(ref["ints"][:, 0], got["ints"][:, 0]), "point counts differ"
    fl_r, fl_g = ref["ints"][:, 1], got["ints"][:, 1]
    # init and update_enable are functions of counts only: exact
    assert np.array_equal(fl_r & 5, fl_g & 5), "init / update_enable flags differ"
    plane_diff = (fl_r & 2) != (fl_g & 2)
    out["plane_flag_diff"] = int(plane_diff.sum())
    out["merged_flag_diff"] = int(((fl_r & 8) != (fl_g & 8)).sum())
    out["group_diff"] = int((ref["group"] != got["group"]).sum())
    same = ~plane_diff
    # merged planes: the fused mean scales with the plane's distance from the
    # origin (the reference divides only the second term, Q4), so its
    # tolerance is relative to |mean| (the cov-trace weights carry ~1e-6
    # relative error from fp32 per-point covariances); the fused normal is a
    # weighted sum of unit normals and keeps an absolute tolerance, relative
    # only to its own length when that exceeds 1
    merged = (fl_r & 8) != 0
    scale = np.where(merged, 10.0 * np.maximum(1.0, np.abs(ref["dbl"][:, 0:3]).max(axis=1)), 1.0)
    dm = np.abs(ref["dbl"][:, 0:3] - got["dbl"][:, 0:3]).max(axis=1) / scale
    out["mean_err_max"] = float(dm[same].max()) if same.any() else 0.0
    pl = same & ((fl_r & 2) != 0)
    n_tol = np.full(len(pl), norm_tol)
    t_tol = np.full(len(pl), tr_rtol)
    if "fit" in ref:
        # Conditioning of the fit: the normal is the eigenvector of the
        # smallest eigenvalue of cov = ppt / n - mean mean^T and moves by
        # about |dcov| / (l1 - l0); the Plane::cov traces carry 1 / (l0 - lm)^2
        # factors and move by about 2 |dcov| / (l1 - l0) relative.  Summation
        # order differences give |dcov| ~ 1e-15 max(1, |mean|^2), so each
        # voxel's tolerance grows by 1e-13 max(1, |mean|^2) / (l1 - l0)
        # (100x that).  A voxel holding copies of a few distinct points
        # (rescans from rest) has l1 - l0 ~ 1e-10 and a normal that rounding
        # picks on either side; for those (tolerance above 1e-3) the device's
        # normal must still be a unit vector of the near-degenerate
        # eigenspace: orthogonal to the major axis when l2 is separated.
        ev, major = ref["fit"][:, 0:3], ref["fit"][:, 3:6]
        sc = np.maximum(1.0, (ref["dbl"][:, 0:3] ** 2).sum(axis=1))
        cond = 1e-13 * sc / np.maximum(ev[:, 1] - ev[:, 0], 1e-300)
        live = pl & ~merged
        n_tol = np.where(live, norm_tol + cond, n_tol)
        t_tol = np.where(live, tr_rtol + 2.0 * cond, t_tol)
        ill = live & (cond > 1e-3)
        ng = got["dbl"][:, 3:6]
        unit_bad = np.abs(np.linalg.norm(ng, axis=1) - 1.0) > 1e-9
        perp_bad = (ev[:, 2] - ev[:, 1] > 1e-9 * sc) & (np.abs((ng * major).sum(axis=1)) > 1e-6)
        out["ill_conditioned"] = int(ill.sum())
        out["ill_conditioned_bad"] = int((ill & (unit_bad | perp_bad)).sum())
    nscale = np.maximum(1.0, np.linalg.norm(ref["dbl"][:, 3:6], axis=1))
    dn = np.abs(ref["dbl"][:, 3:6] - got["dbl"][:, 3:6]).max(axis=1) / nscale
    out["norm_err_max"] = float(dn[pl].max()) if pl.any() else 0.0
    out["norm_bad"] = int((dn[pl] > n_tol[pl]).sum())
    out["mean_bad"] = int((dm[same] > mean_tol).sum())
    out["merged"] = int(merged.sum())
    tr_r, tr_g = ref["dbl"][:, 6:8], got["dbl"][:, 6:8]
    rel = np.abs(tr_r - tr_g) / np.maximum(np.abs(tr_r), 1e-300)
    out["trace_rel_max"] = float(rel[pl].max()) if pl.any() else 0.0
    out["trace_bad"] = int((rel[pl] > t_tol[pl, None]).max(axis=1).sum()) if pl.any() else 0
    out["n"] = int(len(ref["keys"]))
    if report is not None:
        report(out)
    return out


def hb_on_common_set(Ho, bo, s, x, y, z, oslots, gslots, estimate_ext=False):
    """The oracle's H, b moved onto the device's valid set: slots valid only
    on the oracle's side are taken out (with the oracle's normal and
    residual), slots valid only on the device's side are added (with the
    device's), each with the Jacobian row and weight of lio_builder.cpp:283-306
    at state s.  Equal to the device's H, b up to rounding whenever the two
    sides agree on every slot they share."""
    pts = obsref.lidar_points(x, y, z)
    vo, ro, no = oslots
    vg, rg, ng = gslots
    H, b = np.array(Ho, copy=True), np.array(bo, copy=True)
    for i in np.flatnonzero(vo & ~vg):
        J, w = obsref.jac_weight(s, pts[i], np.asarray(no[i], np.float64), estimate_ext)
        H -= w * np.outer(J, J)
        b -= w * J * float(ro[i])
    for i in np.flatnonzero(vg & ~vo):
        J, w = obsref.jac_weight(s, pts[i], np.asarray(ng[i], np.float64), estimate_ext)
        H += w * np.outer(J, J)
        b += w * J * float(rg[i])
    return H, b


def register_both(orc, g, x, y, z):
    """One registered scan on both sides from the oracle's current state as
    the prior (teacher forcing).  The oracle then inserts the scan with the
    device's posterior, so both maps receive bitwise the same fp32 insertion
    points: otherwise a posterior difference of ~1e-8 m moves 0.1-0.3 % of
    the fp32 world points across a rounding boundary (Q10), and the voxels
    they land in differ by up to one fp32 unit in their means.  The IESKF
    parity is the oracle's own posterior against the device's.  Returns
    (oracle iterations, oracle posterior s, P, device observe and map stats,
    device posterior s, P)."""
    sp, Pp = orc.get_state()
    orc.set_points(x, y, z)
    it = orc.update()
    so, Po = orc.get_state()
    g.set_state(sp, Pp)
    g.set_points(x, y, z)
    ost, mst = g.register()
    sg, Pg = g.get_state()
    orc.set_state(sg, Pg)
    orc.map_update()
    return it, so, Po, ost, mst, sg, Pg


def lio_both(orc, lb, pkg):
    """One SyncPackage (x, y, z, t_ms, imu, cloud_start, cloud_end) through
    LIOBuilder::process on both sides, the oracle starting from the device's
    state (teacher forcing): at LIO_MAPPING the oracle undistorts, updates
    and then inserts with the device's posterior, as register_both does, so
    every package checks one step of the device against the oracle from
    the same state and map.  Returns a dict: status (device, oracle),
    iterations (oracle, device), prior (the oracle's propagated state),
    oracle posterior so, Po, device posterior sg, Pg, and the undistorted
    points of both sides (None before LIO_MAPPING)."""
    x, y, z, t, imu, cs, ce = pkg
    out = {"points": None, "it": None}
    if lb.status != 0:
        s, P = lb.reg.get_state()
        orc.set_state(s, P)
    if lb.status in (0, 1):  # IMU_INIT, MAP_INIT: the whole package on both sides
        st_o, _ = orc.process(*pkg)
        st_g = lb.process(*pkg)
    else:
        ux, uy, uz, _ = orc.undistort(*pkg)
        out["prior"] = orc.get_state()
        if orc.scan_resolution > 0.0:
            import oracle

            ux, uy, uz = oracle.voxel_grid(ux, uy, uz, float(np.float32(orc.scan_resolution)))
        orc.set_points(ux, uy, uz)
        it_o = orc.update()
        so, Po = orc.get_state()
        st_g = lb.process(*pkg)
        st_o = st_g if lb.status == 2 else -1
        sg, Pg = lb.reg.get_state()
        orc.set_state(sg, Pg)
        orc.map_update()
        out["it"] = (it_o, lb.last_obs.iterations)
        out["points"] = (orc.scan_points(), lb.reg.export_scan())
        out["so"], out["Po"] = so, Po
    out["status"] = (st_g, st_o)
    out["sg"], out["Pg"] = lb.reg.get_state()
    if "so" not in out:
        out["so"], out["Po"] = orc.get_state()
    return out
