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
        )


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

    def reserve(self, points, sessions):
        """Size the batch workspace for batches of up to `points` points over
        up to `sessions` sessions (no later batch then allocates)."""
        rc = self.lib.vmp_ctx_reserve(self.h, int(points), int(sessions))
        if rc != 0:
            raise VmpError(f"vmp_ctx_reserve({points}, {sessions}) failed ({rc})")

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
