"""GPU parity of the IMU side of the registration step, against the oracle:
the device runs LIOBuilder::process (lio_builder.cpp:175-248) package by
package from raw scans and a 200 Hz IMU stream: IMU_INIT, then MAP_INIT and
LIO_MAPPING with undistortCloud (IMU propagation of x and P, backward point
compensation, lio_builder.cpp:65-153) on the device before every update.

Tolerances:
  state after IMU_INIT and after every package's propagation  1e-9 relative
  undistorted points                                          2 fp32 units of the coordinate
  per-scan pose (free-running, both sides from the same data)  1e-4 m, 1e-5 rad
  posterior P                                                 1e-6 relative to max |P|
  map keys / LRU order                                        bit-exact (margin cases counted)
All calls go through the C ABI.
"""
import sys

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from mapcmp import compare_maps, state_err, ulp_err
from oracle import Oracle
from voxelmapplus_fastlio2_amd import Context, ImuConfig, LIOBuilder, LIOConfig
from voxelmapplus_fastlio2_amd.synth import ROOM, SEED0, STREET, ImuSequence

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def oracle_for(cfg, imu_cfg):
    o = Oracle(cfg)
    o.imu_config(imu_cfg.na, imu_cfg.ng, imu_cfg.nba, imu_cfg.nbg, imu_cfg.imu_init_num, imu_cfg.gravity_align,
                 np.asarray(imu_cfg.r_il).reshape(3, 3), np.asarray(imu_cfg.p_il))
    o.set_scan_resolution(cfg.scan_resolution)
    return o


def run_pair(ctx, scene, seed, n, nscans, cfg, imu_cfg=None, check=None):
    imu_cfg = imu_cfg or ImuConfig()
    seq = ImuSequence(scene, seed, n)
    orc = oracle_for(cfg, imu_cfg)
    worst = {"pos": 0.0, "rot": 0.0, "P": 0.0, "pts": 0.0, "pts_ulp": 0.0}
    with LIOBuilder(ctx, cfg, imu_cfg) as lb:
        for k in range(nscans):
            x, y, z, t, imu, cs, ce = seq.package(k)
            mapping = lb.status == 2  # this package is registered (LIO_MAPPING)
            st_o, it_o = orc.process(x, y, z, t, imu, cs, ce)
            st_g = lb.process(x, y, z, t, imu, cs, ce)
            assert st_o == st_g, (k, st_o, st_g)
            so, Po = orc.get_state()
            sg, Pg = lb.reg.get_state()
            ep, er = state_err(so, sg)
            eP = np.abs(Po - Pg).max() / np.abs(Po).max()
            line = f"scan {k}: status {st_g} pos {ep:.2e} rot {er:.2e} P {eP:.2e}"
            if mapping:
                pts_o = orc.scan_points()
                pts_g = lb.reg.export_scan()
                assert pts_o.shape == pts_g.shape
                epts = float(np.abs(pts_o - pts_g).max()) if pts_o.size else 0.0
                worst["pts"] = max(worst["pts"], epts)
                worst["pts_ulp"] = max(worst["pts_ulp"], ulp_err(pts_o, pts_g))
                line += f" pts {epts:.1e} exact {np.mean(pts_o == pts_g):.4f}"
                if lb.last_obs is not None:
                    line += f" iters o/g {it_o}/{lb.last_obs.iterations}"
                    if it_o == lb.last_obs.iterations:
                        worst["P"] = max(worst["P"], eP)
            worst["pos"] = max(worst["pos"], ep)
            worst["rot"] = max(worst["rot"], er)
            log(line)
            assert ep <= 1e-4 and er <= 1e-5, line
            if check is not None:
                check(k, orc, lb)
        r = compare_maps(orc.export_map(), lb.reg.export_map(), report=lambda o: log("  final map", o))
        assert r["plane_flag_diff"] <= max(2, r["n"] // 1000)
    log("worst", worst)
    assert worst["pts_ulp"] <= 2.0
    assert worst["P"] <= 1e-6
    return worst


def test_imu_init_and_propagation_match_the_oracle(ctx):
    """IMU_INIT, then the first MAP_INIT package: the state and P after the
    device's propagation equal the oracle's."""
    cfg = LIOConfig(max_scan_points=8000)
    seq = ImuSequence(ROOM, SEED0 + 51, 8000)
    orc = oracle_for(cfg, ImuConfig())
    with LIOBuilder(ctx, cfg) as lb:
        for k in range(3):
            pkg = seq.package(k)
            orc.process(*pkg)
            lb.process(*pkg)
            so, Po = orc.get_state()
            sg, Pg = lb.reg.get_state()
            es = np.abs(so - sg).max() / np.abs(so).max()
            eP = np.abs(Po - Pg).max() / np.abs(Po).max()
            log(f"package {k}: status {lb.status} state rel {es:.2e} P rel {eP:.2e}")
            assert es <= 1e-9 and eP <= 1e-9
            if k == 0:
                R = sg[3:12].reshape(3, 3).T
                g = sg[33:36]
                assert abs(np.linalg.norm(g) - 9.81) < 1e-12 and g[2] < -9.80
                assert np.abs(R).max() <= 1.0 + 1e-12
        # MAP_INIT used the undistorted cloud: same points on both sides
        po = orc.export_map()
        pg = lb.reg.export_map()
        assert np.array_equal(po["keys"], pg["keys"])


def test_imu_room_sequence(ctx):
    """C1's room, 20k points per scan, 40 packages from rest."""
    cfg = LIOConfig(max_scan_points=20000)
    run_pair(ctx, ROOM, SEED0 + 52, 20000, 40, cfg)


def test_imu_street_sequence(ctx):
    """The street at 60k points per scan, 25 packages."""
    cfg = LIOConfig(max_scan_points=60000, map_capacity=2000000)
    run_pair(ctx, STREET, SEED0 + 53, 60000, 25, cfg)


def test_odom_and_body_cloud_exports(ctx):
    """PointCloudWithOdom (lio_node.cpp:200-229): pose, quaternion and the
    6x6 covariance block; the body cloud is the undistorted cloud through
    lidarToBody's fp32 extrinsic transform."""
    r_il = Rotation.from_rotvec([0.0, 0.0, 0.02]).as_matrix()
    imu_cfg = ImuConfig(r_il=tuple(r_il.reshape(-1)), p_il=(0.05, -0.02, 0.1))
    cfg = LIOConfig(max_scan_points=10000)

    def check(k, orc, lb):
        if k < 2:  # IMU_INIT, MAP_INIT
            return
        so, Po = orc.get_state()
        pos, q, cov = lb.reg.odom()
        np.testing.assert_allclose(pos, so[0:3], atol=1e-4)
        R = so[3:12].reshape(3, 3).T
        qs = Rotation.from_matrix(R).as_quat()  # x y z w
        qs = np.array([qs[3], qs[0], qs[1], qs[2]])
        if qs[0] * q[0] < 0:
            qs = -qs
        np.testing.assert_allclose(q, qs, atol=1e-5)
        np.testing.assert_allclose(cov, Po[0:6, 0:6], rtol=1e-6, atol=1e-12)
        body = lb.reg.export_cloud_body()
        pts = orc.scan_points()  # no filter: the registered points are the undistorted cloud
        Re = np.asarray(imu_cfg.r_il, np.float64).reshape(3, 3).astype(np.float32)
        te = np.asarray(imu_cfg.p_il, np.float32)
        exp = np.empty_like(pts)
        for r in range(3):  # m0 x + (m1 y + (m2 z + t)) in fp32
            a = Re[r, 2] * pts[2] + te[r]
            b = Re[r, 1] * pts[1] + a
            exp[r] = Re[r, 0] * pts[0] + b
        assert body.shape == exp.shape
        assert np.abs(body - exp).max() <= 1e-5

    run_pair(ctx, ROOM, SEED0 + 54, 10000, 8, cfg, imu_cfg, check)


def test_imu_sequence_with_scan_filter(ctx):
    """scan_resolution = 0.1 (the reference yaml): every registered scan
    goes through the VoxelGrid filter on the device; its output (count,
    order, fp32 means) equals the oracle's."""
    cfg = LIOConfig(max_scan_points=20000, scan_resolution=0.1)

    def check(k, orc, lb):
        if k >= 2:
            po, pg = orc.scan_points(), lb.reg.export_scan()
            assert pg.shape[1] < 20000, "the filter merged points"
            assert po.shape == pg.shape, (po.shape, pg.shape)

    run_pair(ctx, ROOM, SEED0 + 55, 20000, 12, cfg, check=check)
