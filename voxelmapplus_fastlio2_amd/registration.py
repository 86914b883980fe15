"""Python mirror of the reference's registration interfaces over the C ABI.

Names follow voxel_plus (paths relative to the reference root,
voxel_plus/src/map_builder/):

  LIOConfig             lio_builder.h:17-42 (subset on the registration path)
  Registrar.set_points  residual-slot prep, lio_builder.cpp:225-229
  Registrar.observe     LIOBuilder::sharedUpdateFunc, lio_builder.cpp:250-311
  Registrar.update      kf::IESKF::update, ieskf.cpp:125-156
  Registrar.build       VoxelMap::build via MAP_INIT, lio_builder.cpp:185-211
  Registrar.map_update  VoxelMap::update via lio_builder.cpp:231-246
  Registrar.register    the LIO_MAPPING branch after undistortion
  Registrar.imu_init    LIOBuilder::initializeImu, lio_builder.cpp:28-63
  Registrar.set_raw     a SyncPackage (raw points + IMU) whose undistortCloud
                        (lio_builder.cpp:65-153) the next call runs on the device
  LIOBuilder.process    LIOBuilder::process, lio_builder.cpp:175-248

State vectors use one flat layout of 36 float64 values:
pos[3], rot[9], rot_ext[9], pos_ext[3], vel[3], bg[3], ba[3], g[3], with the
3x3 blocks column-major (Eigen's memory order); P is 23x23.
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib


class VmpError(RuntimeError):
    pass


@dataclass
class LIOConfig:
    """Effective reference parameters (lio_builder.h:17-42, mapping.yaml:11-17).

    update_size_thresh is 10 because the yaml key is misspelled
    (``update_size_threshes``) and the node falls back to its default;
    merge_thresh_for_angle is the yaml's 0.2.
    """

    voxel_size: float = 0.5
    plane_thresh: float = 0.01
    merge_thresh_for_angle: float = 0.2
    merge_thresh_for_distance: float = 0.04
    ranging_cov: float = 0.04
    angle_cov: float = 0.1
    max_point_thresh: int = 100
    update_size_thresh: int = 10
    map_capacity: int = 100000
    max_scan_points: int = 20000
    max_iter: int = 5
    estimate_ext: int = 0
    scan_resolution: float = 0.0

    def to_c(self):
        return _lib.MapCfg(
            self.voxel_size,
            self.plane_thresh,
            self.merge_thresh_for_angle,
            self.merge_thresh_for_distance,
            self.ranging_cov,
            self.angle_cov,
            self.max_point_thresh,
            self.update_size_thresh,
            self.map_capacity,
            self.max_scan_points,
            self.max_iter,
            self.estimate_ext,
            self.scan_resolution,
        )


@dataclass
class ImuConfig:
    """IMU side of LIOConfig (lio_builder.h:19-28, lio_node.cpp:76-88)."""

    na: float = 0.01
    ng: float = 0.01
    nba: float = 0.0001
    nbg: float = 0.0001
    imu_init_num: int = 20
    gravity_align: bool = True
    r_il: tuple = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)  # row-major, as the yaml lists it
    p_il: tuple = (0.0, 0.0, 0.0)

    def to_c(self):
        c = _lib.ImuCfg()
        c.na, c.ng, c.nba, c.nbg = self.na, self.ng, self.nba, self.nbg
        c.imu_init_num = self.imu_init_num
        c.gravity_align = 1 if self.gravity_align else 0
        R = np.asarray(self.r_il, np.float64).reshape(3, 3)
        for i, v in enumerate(R.T.reshape(-1)):  # column-major
            c.r_il[i] = float(v)
        for i in range(3):
            c.p_il[i] = float(self.p_il[i])
        return c


def state_vector(R=None, t=None):
    """Flat 36-value state with the given LiDAR pose, gravity (0,0,-9.81)."""
    s = np.zeros(36)
    s[3:12] = (np.eye(3) if R is None else np.asarray(R)).T.reshape(-1)  # column-major
    s[12:21] = np.eye(3).reshape(-1)
    if t is not None:
        s[0:3] = t
    s[33:36] = [0.0, 0.0, -9.81]
    return s


def state_rot(s):
    return np.asarray(s[3:12]).reshape(3, 3).T


def initial_covariance():
    """P after IMU initialisation (lio_builder.cpp:55-60)."""
    P = np.eye(23)
    for i in range(6, 12):
        P[i, i] = 1e-5
    for i in range(15, 21):
        P[i, i] = 1e-4
    P[21, 21] = P[22, 22] = 1e-5
    return P


def _state_c(s):
    st = _lib.State()
    C.memmove(C.addressof(st), np.ascontiguousarray(s, dtype=np.float64).ctypes.data, 36 * 8)
    return st


def _state_np(st):
    out = np.empty(36)
    C.memmove(out.ctypes.data, C.addressof(st), 36 * 8)
    return out


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Context:
    """One device (vmp_ctx)."""

    def __init__(self, device=0):
        self.lib = _lib.vmp()
        h = C.c_void_p()
        rc = self.lib.vmp_ctx_create(device, C.byref(h))
        if rc != 0:
            raise VmpError(f"vmp_ctx_create failed ({rc})")
        self.h = h

    def close(self):
        if self.h:
            self.lib.vmp_ctx_destroy(self.h)
            self.h = None

    def set_param(self, name, value):
        rc = self.lib.vmp_ctx_set_param(self.h, name.encode(), int(value))
        if rc != 0:
            raise VmpError(f"vmp_ctx_set_param({name}, {value}) failed ({rc})")

    def graph_stats(self):
        """{filter_replays, filter_builds, map_replays, map_builds} so far."""
        out = (C.c_int64 * 4)()
        rc = self.lib.vmp_ctx_graph_stats(self.h, out)
        if rc != 0:
            raise VmpError(f"vmp_ctx_graph_stats failed ({rc})")
        return dict(zip(("filter_replays", "filter_builds", "map_replays", "map_builds"), list(out)))

    def reserve(self, points, sessions):
        """Size the batch workspace for batches of up to `points` points over
        up to `sessions` sessions (no later batch then allocates)."""
        rc = self.lib.vmp_ctx_reserve(self.h, int(points), int(sessions))
        if rc != 0:
            raise VmpError(f"vmp_ctx_reserve({points}, {sessions}) failed ({rc})")

    TIMED_KERNELS = ("k_ins_keys", "k_create_evict", "k_replay_heavy", "k_replay_light", "k_merge_replay",
                     "k_candidates", "k_ring_live", "k_ieskf_step", "k_undistort", "k_imu_predict", "k_spec_eval")

    def kernel_times(self):
        """{kernel: (total ms, launches)} of the per-kernel event timing."""
        ms = (C.c_double * 16)()
        n = (C.c_int64 * 16)()
        k = self.lib.vmp_ctx_kernel_times(self.h, 16, ms, n)
        if k < 0:
            raise VmpError(f"vmp_ctx_kernel_times failed ({k})")
        return {name: (float(ms[i]), int(n[i])) for i, name in enumerate(self.TIMED_KERNELS[:k])}

    def autotune_observe(self, registrar, reps=10):
        """Time the observe pass on the registrar's current scan for 1, 2, 4
        and 8 points per thread; the fastest becomes this context's choice
        for scans of that size bucket.  Returns (points per thread, medians
        in ms for the four candidates)."""
        best = C.c_int32()
        ms = (C.c_float * 4)()
        registrar._check(self.lib.vmp_autotune_observe(registrar.h, int(reps), C.byref(best), ms), "vmp_autotune_observe")
        return best.value, [float(v) for v in ms]


class Registrar:
    """One LiDAR sequence: voxel map, residual slots and filter state in HBM."""

    def __init__(self, ctx, cfg=None):
        self.ctx = ctx
        self.lib = ctx.lib
        self.cfg = cfg or LIOConfig()
        h = C.c_void_p()
        c = self.cfg.to_c()
        rc = self.lib.vmp_session_create(ctx.h, C.byref(c), C.byref(h))
        if rc != 0:
            raise VmpError(f"vmp_session_create failed ({rc})")
        self.h = h
        self.n = 0

    def close(self):
        if self.h:
            self.lib.vmp_session_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.vmp_last_error(self.h)
            raise VmpError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    # ---- filter state
    def set_state(self, s, P=None):
        st = _state_c(s)
        Pc = None if P is None else np.asfortranarray(P, dtype=np.float64)
        self._check(
            self.lib.vmp_state_set(self.h, C.byref(st), None if Pc is None else Pc.ctypes.data_as(C.POINTER(C.c_double))),
            "vmp_state_set",
        )

    def get_state(self):
        st = _lib.State()
        P = np.zeros((23, 23), order="F")
        self._check(self.lib.vmp_state_get(self.h, C.byref(st), P.ctypes.data_as(C.POINTER(C.c_double))), "vmp_state_get")
        return _state_np(st), np.ascontiguousarray(P)

    def apply_motion(self, dR, dp, qdiag):
        dRc = np.asfortranarray(dR, dtype=np.float64)
        dpc = np.ascontiguousarray(dp, dtype=np.float64)
        q = np.ascontiguousarray(qdiag, dtype=np.float64)
        self._check(
            self.lib.vmp_state_apply_motion(self.h, dRc.ctypes.data_as(C.POINTER(C.c_double)), _dptr(dpc), _dptr(q)),
            "vmp_state_apply_motion",
        )

    # ---- scan
    def set_points(self, x, y, z):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        z = np.ascontiguousarray(z, dtype=np.float32)
        self._check(self.lib.vmp_scan_set_points(self.h, _fptr(x), _fptr(y), _fptr(z), len(x)), "vmp_scan_set_points")
        self.n = len(x)

    def set_points_device(self, px, py, pz, n):
        """Device pointers (e.g. torch tensors' data_ptr()) already in HBM."""
        self._check(
            self.lib.vmp_scan_set_points_device(self.h, C.c_void_p(px), C.c_void_p(py), C.c_void_p(pz), int(n)),
            "vmp_scan_set_points_device",
        )
        self.n = int(n)

    # ---- IMU side
    def imu_init(self, imu, cloud_end, imu_cfg=None):
        """Add an IMU_INIT package's samples; True once initialised."""
        t, acc, gyro = (np.ascontiguousarray(a, np.float64) for a in imu)
        done = C.c_int32()
        c = (imu_cfg or ImuConfig()).to_c()
        self._check(
            self.lib.vmp_imu_init(self.h, C.byref(c), _dptr(t), _dptr(acc), _dptr(gyro), len(t), float(cloud_end),
                                  C.byref(done)),
            "vmp_imu_init",
        )
        return bool(done.value)

    def set_raw(self, x, y, z, t_ms, imu, cloud_start, cloud_end):
        """A raw package from host arrays (points in scan time order)."""
        x, y, z, t_ms = (np.ascontiguousarray(a, dtype=np.float32) for a in (x, y, z, t_ms))
        it, acc, gyro = (np.ascontiguousarray(a, np.float64) for a in imu)
        self._check(
            self.lib.vmp_scan_set_raw(self.h, _fptr(x), _fptr(y), _fptr(z), _fptr(t_ms), len(x), _dptr(it),
                                      _dptr(acc), _dptr(gyro), len(it), float(cloud_start), float(cloud_end)),
            "vmp_scan_set_raw",
        )
        self.n = len(x)

    def set_raw_device(self, px, py, pz, pt, n, imu, cloud_start, cloud_end):
        """A raw package whose points and times already live in HBM."""
        it, acc, gyro = (np.ascontiguousarray(a, np.float64) for a in imu)
        self._check(
            self.lib.vmp_scan_set_raw_device(self.h, C.c_void_p(px), C.c_void_p(py), C.c_void_p(pz), C.c_void_p(pt),
                                             int(n), _dptr(it), _dptr(acc), _dptr(gyro), len(it), float(cloud_start),
                                             float(cloud_end)),
            "vmp_scan_set_raw_device",
        )
        self.n = int(n)

    def export_scan(self):
        """The points of the last registration (undistorted, filtered), 3 x n."""
        n = int(self.lib.vmp_scan_export(self.h, 0, None, None, None))
        out = np.zeros((3, max(n, 1)), np.float32)
        k = self.lib.vmp_scan_export(self.h, n, _fptr(out[0]), _fptr(out[1]), _fptr(out[2]))
        if k < 0:
            raise VmpError(f"vmp_scan_export failed ({k})")
        return out[:, :n]

    def export_cloud_body(self):
        """The last package's undistorted cloud in the body frame, 3 x n."""
        n = int(self.lib.vmp_cloud_body_export(self.h, 0, None, None, None))
        out = np.zeros((3, max(n, 1)), np.float32)
        k = self.lib.vmp_cloud_body_export(self.h, n, _fptr(out[0]), _fptr(out[1]), _fptr(out[2]))
        if k < 0:
            raise VmpError(f"vmp_cloud_body_export failed ({k})")
        return out[:, :n]

    def odom(self):
        """(position, quaternion w x y z, 6x6 covariance) of PointCloudWithOdom."""
        pos = np.zeros(3)
        q = np.zeros(4)
        cov = np.zeros(36)
        self._check(self.lib.vmp_odom_export(self.h, _dptr(pos), _dptr(q), _dptr(cov)), "vmp_odom_export")
        return pos, q, cov.reshape(6, 6)

    # ---- path entry points
    def observe(self, s):
        st = _state_c(s)
        H = np.zeros((12, 12), order="F")
        b = np.zeros(12)
        o = _lib.ObsStats()
        self._check(
            self.lib.vmp_observe(self.h, C.byref(st), H.ctypes.data_as(C.POINTER(C.c_double)), _dptr(b), C.byref(o)),
            "vmp_observe",
        )
        return np.ascontiguousarray(H), b, o

    def update(self):
        o = _lib.ObsStats()
        self._check(self.lib.vmp_ieskf_update(self.h, C.byref(o)), "vmp_ieskf_update")
        return o

    def build(self):
        m = _lib.MapStats()
        self._check(self.lib.vmp_map_update(self.h, 1, C.byref(m)), "vmp_map_update(build)")
        return m

    def map_update(self):
        m = _lib.MapStats()
        self._check(self.lib.vmp_map_update(self.h, 0, C.byref(m)), "vmp_map_update")
        return m

    def register(self):
        o = _lib.ObsStats()
        m = _lib.MapStats()
        self._check(self.lib.vmp_register_scan(self.h, C.byref(o), C.byref(m)), "vmp_register_scan")
        return o, m

    def register_noalloc(self):
        """register() through the entry point that allocates nothing (device
        scratch of the library; scans up to 131072 points)."""
        o = _lib.ObsStats()
        m = _lib.MapStats()
        self._check(self.lib.vmp_register_scan_noalloc(self.h, C.byref(o), C.byref(m)), "vmp_register_scan_noalloc")
        return o, m

    def timing(self):
        a = C.c_double()
        b = C.c_double()
        n = C.c_int32()
        self._check(self.lib.vmp_session_timing(self.h, C.byref(a), C.byref(b), C.byref(n)), "vmp_session_timing")
        return a.value, b.value, n.value

    def set_timing(self, on=True):
        self._check(self.lib.vmp_session_set_timing(self.h, 1 if on else 0), "vmp_session_set_timing")

    def timing_totals(self):
        om, mm = C.c_double(), C.c_double()
        on, mn, op, oh = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        self._check(
            self.lib.vmp_session_timing_totals(
                self.h, C.byref(om), C.byref(on), C.byref(mm), C.byref(mn), C.byref(op), C.byref(oh)
            ),
            "vmp_session_timing_totals",
        )
        return {
            "obs_ms": om.value,
            "obs_launches": on.value,
            "map_ms": mm.value,
            "map_calls": mn.value,
            "obs_points": op.value,
            "obs_hits": oh.value,
        }

    # ---- inspection
    def map_size(self):
        return int(self.lib.vmp_map_size(self.h))

    def export_map(self):
        n = self.map_size()
        keys = np.zeros((max(n, 1), 3), np.int64)
        ints = np.zeros((max(n, 1), 5), np.int32)
        group = np.zeros(max(n, 1), np.uint64)
        dbl = np.zeros((max(n, 1), 8))
        cnt = self.lib.vmp_map_export(
            self.h,
            n,
            keys.ctypes.data_as(C.POINTER(C.c_int64)),
            ints.ctypes.data_as(C.POINTER(C.c_int32)),
            group.ctypes.data_as(C.POINTER(C.c_uint64)),
            _dptr(dbl),
        )
        if cnt < 0:
            raise VmpError("vmp_map_export failed")
        return {"keys": keys[:cnt], "ints": ints[:cnt], "group": group[:cnt], "dbl": dbl[:cnt]}

    def export_slots(self, n):
        valid = np.zeros(n, np.uint8)
        res = np.zeros(n, np.float32)
        nrm = np.zeros((n, 3), np.float32)
        self._check(
            self.lib.vmp_slots_export(self.h, n, valid.ctypes.data_as(C.POINTER(C.c_uint8)), _fptr(res), _fptr(nrm)),
            "vmp_slots_export",
        )
        return valid.astype(bool), res, nrm


def register_batch(registrars):
    lib = registrars[0].lib
    n = len(registrars)
    arr = (C.c_void_p * n)(*[r.h.value for r in registrars])
    o = (_lib.ObsStats * n)()
    m = (_lib.MapStats * n)()
    rc = lib.vmp_register_batch(arr, n, o, m)
    if rc != 0:
        raise VmpError(f"vmp_register_batch failed ({rc})")
    return o, m


def apply_motion_batch(registrars, dRs, dps, qdiag):
    """Motion prior for many sessions of one context in one launch."""
    lib = registrars[0].lib
    n = len(registrars)
    arr = (C.c_void_p * n)(*[r.h.value for r in registrars])
    dR = np.ascontiguousarray(np.stack([np.asarray(R, np.float64).T.reshape(-1) for R in dRs]))  # column-major
    dp = np.ascontiguousarray(np.stack([np.asarray(p, np.float64) for p in dps]))
    q = np.asarray(qdiag, np.float64)
    q = np.ascontiguousarray(np.broadcast_to(q, (n, 23)) if q.ndim == 1 else q)
    rc = lib.vmp_state_apply_motion_batch(arr, n, _dptr(dR), _dptr(dp), _dptr(q))
    if rc != 0:
        raise VmpError(f"vmp_state_apply_motion_batch failed ({rc})")


def set_points_device_batch(registrars, xs, ys, zs, counts):
    """Bind device point buffers (addresses) of many sessions in one call."""
    lib = registrars[0].lib
    n = len(registrars)
    arr = (C.c_void_p * n)(*[r.h.value for r in registrars])
    ax = np.ascontiguousarray(xs, dtype=np.uint64)
    ay = np.ascontiguousarray(ys, dtype=np.uint64)
    az = np.ascontiguousarray(zs, dtype=np.uint64)
    cn = np.ascontiguousarray(counts, dtype=np.int32)
    u64 = C.POINTER(C.c_uint64)
    rc = lib.vmp_scan_set_points_device_batch(
        arr, n, ax.ctypes.data_as(u64), ay.ctypes.data_as(u64), az.ctypes.data_as(u64),
        cn.ctypes.data_as(C.POINTER(C.c_int32)),
    )
    if rc != 0:
        raise VmpError(f"vmp_scan_set_points_device_batch failed ({rc})")
    for r, c in zip(registrars, cn):
        r.n = int(c)


def set_raw_device_batch(registrars, xs, ys, zs, ts, counts, imus, cloud_starts, cloud_ends):
    """Bind raw packages of many sessions (device addresses of points and
    times; imus: one (t, acc, gyro) per session) in one call."""
    lib = registrars[0].lib
    n = len(registrars)
    arr = (C.c_void_p * n)(*[r.h.value for r in registrars])
    u64 = C.POINTER(C.c_uint64)
    a = [np.ascontiguousarray(v, dtype=np.uint64) for v in (xs, ys, zs, ts)]
    cn = np.ascontiguousarray(counts, dtype=np.int32)
    it = np.ascontiguousarray(np.concatenate([np.asarray(i[0], np.float64) for i in imus]))
    ia = np.ascontiguousarray(np.concatenate([np.asarray(i[1], np.float64).reshape(-1, 3) for i in imus]))
    ig = np.ascontiguousarray(np.concatenate([np.asarray(i[2], np.float64).reshape(-1, 3) for i in imus]))
    ic = np.ascontiguousarray([len(i[0]) for i in imus], dtype=np.int32)
    cs = np.ascontiguousarray(cloud_starts, dtype=np.float64)
    ce = np.ascontiguousarray(cloud_ends, dtype=np.float64)
    i32 = C.POINTER(C.c_int32)
    rc = lib.vmp_scan_set_raw_device_batch(
        arr, n, *(v.ctypes.data_as(u64) for v in a), cn.ctypes.data_as(i32), _dptr(it), _dptr(ia), _dptr(ig),
        ic.ctypes.data_as(i32), _dptr(cs), _dptr(ce),
    )
    if rc != 0:
        raise VmpError(f"vmp_scan_set_raw_device_batch failed ({rc})")
    for r, c in zip(registrars, cn):
        r.n = int(c)


def register_batch_submit(registrars):
    """Launch a batch without waiting (the next submit or a wait completes it)."""
    lib = registrars[0].lib
    n = len(registrars)
    arr = (C.c_void_p * n)(*[r.h.value for r in registrars])
    rc = lib.vmp_register_batch_submit(arr, n)
    if rc != 0:
        raise VmpError(f"vmp_register_batch_submit failed ({rc})")


def autotune_observe_batch(registrars, reps=5):
    """Tune the observe launch shape of every size bucket on the batch as it
    runs (one observe launch over all the registrars per timing, all of one
    context).  Returns (table of points per thread for the four buckets,
    4 x 4 medians in ms)."""
    lib = registrars[0].lib
    n = len(registrars)
    arr = (C.c_void_p * n)(*[r.h.value for r in registrars])
    table = (C.c_int32 * 4)()
    ms = (C.c_float * 16)()
    rc = lib.vmp_autotune_observe_batch(arr, n, int(reps), table, ms)
    if rc != 0:
        raise VmpError(f"vmp_autotune_observe_batch failed ({rc})")
    return list(table), [list(ms[4 * k : 4 * k + 4]) for k in range(4)]


def register_batch_wait(ctx, n):
    """Complete the batch in flight; statistics of the last completed batch."""
    o = (_lib.ObsStats * n)()
    m = (_lib.MapStats * n)()
    rc = ctx.lib.vmp_register_batch_wait(ctx.h, n, o, m)
    if rc != 0:
        raise VmpError(f"vmp_register_batch_wait failed ({rc})")
    return o, m


IMU_INIT, MAP_INIT, LIO_MAPPING = 0, 1, 2  # LIOStatus (lio_builder.h:10-16)


class LIOBuilder:
    """lio::LIOBuilder over one device session: process() takes one
    SyncPackage (raw points with per-point times, the IMU samples up to the
    scan's end, cloud start/end times) through the reference's state machine
    (lio_builder.cpp:175-248): IMU_INIT averages IMU samples, MAP_INIT
    undistorts and builds the map, LIO_MAPPING undistorts, filters (with
    scan_resolution > 0), updates the IESKF and the map.  Everything after
    IMU_INIT runs in one device call per package."""

    def __init__(self, ctx, cfg=None, imu_cfg=None):
        self.cfg = cfg or LIOConfig()
        self.imu_cfg = imu_cfg or ImuConfig()
        self.reg = Registrar(ctx, self.cfg)
        self.status = IMU_INIT
        self.last_obs = None
        self.last_map = None

    def close(self):
        self.reg.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def process(self, x, y, z, t_ms, imu, cloud_start, cloud_end, device=None):
        """device = (px, py, pz, pt) HBM addresses of the same points, if any."""
        if self.status == IMU_INIT:
            if self.reg.imu_init(imu, cloud_end, self.imu_cfg):
                self.status = MAP_INIT
            return self.status
        if device is None:
            self.reg.set_raw(x, y, z, t_ms, imu, cloud_start, cloud_end)
        else:
            self.reg.set_raw_device(*device, len(x), imu, cloud_start, cloud_end)
        if self.status == MAP_INIT:
            self.last_map = self.reg.build()
            self.status = LIO_MAPPING
        else:
            self.last_obs, self.last_map = self.reg.register()
        return self.status
