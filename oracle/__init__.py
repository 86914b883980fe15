"""ORACLE — test infrastructure only.

ctypes driver of the FP64 CPU restatement of the reference registration path
(oracle/vmp_oracle.cpp).  Imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; the product never uses it.

Pinning status (each function's checks live in tests/test_oracle_*.py):
  eigh3, invert, so3_exp/log, left_jacobian  numpy / scipy fixtures + identities
  calc_body_cov                              closed-form eigen-structure
  key / index                                brute-force floor, face cases
  VoxelMap build/update/LRU/eviction         brute-force Python replica (tiny)
  pushPoint/updatePlane                      hand cases (10/100 points, freeze)
  merge                                      hand-computed fusion incl. the
                                             operator-precedence quirk
  observe (H, b)                             brute-force numpy sum, symmetry
  IESKF update                               zero-observation identities
Not pinned by anything beyond the above (no reference build or golden data
exists): Eigen/Sophus/PCL rounding and op order — see the choices listed in
vmp_oracle.hpp.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "_build", "liborc.so")


class Cfg(C.Structure):
    _fields_ = [
        (n, C.c_double)
        for n in (
            "voxel_size",
            "plane_thresh",
            "merge_thresh_for_angle",
            "merge_thresh_for_distance",
            "ranging_cov",
            "angle_cov",
        )
    ] + [
        (n, C.c_int32)
        for n in ("max_point_thresh", "update_size_thresh", "map_capacity", "max_scan_points", "max_iter", "estimate_ext")
    ] + [("scan_resolution", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            import subprocess

            subprocess.run(["make", "-C", _HERE], check=True, stdout=subprocess.DEVNULL)
        L = C.CDLL(LIB)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(Cfg)]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_map_size.restype = C.c_int64
        L.orc_map_size.argtypes = [C.c_void_p]
        L.orc_map_export.restype = C.c_int64
        L.orc_map_fit.restype = C.c_int64
        L.orc_map_fit.argtypes = [C.c_void_p, C.c_int64, C.c_void_p]
        L.orc_map_export.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        for name in (
            "orc_set_state",
            "orc_get_state",
            "orc_map_init",
            "orc_set_scan",
            "orc_observe",
            "orc_ieskf_update",
            "orc_map_update",
            "orc_register",
            "orc_apply_motion",
            "orc_slots_export",
            "orc_stats",
            "orc_map_push",
            "orc_key",
            "orc_lidar_to_world_f32",
            "orc_imu_config",
            "orc_set_scan_resolution",
            "orc_undistort",
            "orc_predict",
            "orc_imu_state",
            "orc_scan_points",
            "orc_voxel_grid",
            "orc_rot_from_two_vectors",
            "orc_imu_poses",
        ):
            getattr(L, name).argtypes = None
        L.orc_process.restype = C.c_int
        L.orc_process.argtypes = [C.c_void_p] * 5 + [C.c_int] + [C.c_void_p] * 3 + [C.c_int, C.c_double, C.c_double,
                                                                                     C.c_void_p]
        L.orc_initialize_imu.restype = C.c_int
        L.orc_initialize_imu.argtypes = [C.c_void_p] * 4 + [C.c_int]
        _lib = L
    return _lib


def _imu_arrays(imu):
    """(t[m], acc[m,3], gyro[m,3]) as contiguous float64."""
    t, acc, gyro = imu
    return (np.ascontiguousarray(t, np.float64), np.ascontiguousarray(acc, np.float64).reshape(-1, 3),
            np.ascontiguousarray(gyro, np.float64).reshape(-1, 3))


def P(a):
    """Address of a numpy array as a call argument; the argument keeps the
    array alive for the call (temporaries made in the argument list would
    otherwise be freed before the C function reads them)."""
    v = C.c_void_p(a.ctypes.data)
    v._keep = a
    return v


def cfg_from(lio_cfg):
    c = Cfg()
    for f, _ in Cfg._fields_:
        setattr(c, f, getattr(lio_cfg, f))
    return c


class Oracle:
    """Registration driver with the same method names as the product mirror."""

    def __init__(self, lio_cfg):
        self.L = lib()
        self.c = cfg_from(lio_cfg)
        self.h = C.c_void_p(self.L.orc_create(C.byref(self.c)))
        self.n = 0
        self.scan_resolution = float(getattr(lio_cfg, "scan_resolution", 0.0))

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, s, Pm=None):
        s = np.ascontiguousarray(s, np.float64)
        Pc = None if Pm is None else np.asfortranarray(Pm, np.float64)
        self.L.orc_set_state(self.h, P(s), None if Pc is None else C.c_void_p(Pc.ctypes.data))

    def get_state(self):
        s = np.zeros(36)
        Pm = np.zeros((23, 23), order="F")
        self.L.orc_get_state(self.h, P(s), C.c_void_p(Pm.ctypes.data))
        return s, np.ascontiguousarray(Pm)

    def apply_motion(self, dR, dp, qdiag):
        dRc = np.asfortranarray(dR, np.float64)
        self.L.orc_apply_motion(
            self.h, C.c_void_p(dRc.ctypes.data), P(np.ascontiguousarray(dp, np.float64)), P(np.ascontiguousarray(qdiag, np.float64))
        )

    def build(self, x, y, z):
        x, y, z = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))
        self.L.orc_map_init(self.h, P(x), P(y), P(z), C.c_int(len(x)))

    def set_points(self, x, y, z):
        x, y, z = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))
        self.n = len(x)
        self.L.orc_set_scan(self.h, P(x), P(y), P(z), C.c_int(len(x)))

    def observe(self, s):
        H = np.zeros((12, 12), order="F")
        b = np.zeros(12)
        eff = C.c_int()
        self.L.orc_observe(self.h, P(np.ascontiguousarray(s, np.float64)), C.c_void_p(H.ctypes.data), P(b), C.byref(eff))
        return np.ascontiguousarray(H), b, eff.value

    def update(self):
        it = C.c_int()
        self.L.orc_ieskf_update(self.h, C.byref(it))
        return it.value

    def map_update(self):
        self.L.orc_map_update(self.h)

    def register(self, x, y, z):
        x, y, z = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))
        it = C.c_int()
        self.n = len(x)
        self.L.orc_register(self.h, P(x), P(y), P(z), C.c_int(len(x)), C.byref(it))
        return it.value

    def map_size(self):
        return int(self.L.orc_map_size(self.h))

    def export_map(self):
        n = self.map_size()
        keys = np.zeros((max(n, 1), 3), np.int64)
        ints = np.zeros((max(n, 1), 5), np.int32)
        group = np.zeros(max(n, 1), np.uint64)
        dbl = np.zeros((max(n, 1), 8))
        cnt = self.L.orc_map_export(self.h, n, P(keys), P(ints), P(group), P(dbl))
        fit = np.zeros((max(n, 1), 6))
        self.L.orc_map_fit(self.h, n, P(fit))
        return {"keys": keys[:cnt], "ints": ints[:cnt], "group": group[:cnt], "dbl": dbl[:cnt], "fit": fit[:cnt]}

    def export_slots(self, n):
        valid = np.zeros(n, np.uint8)
        res = np.zeros(n)
        nrm = np.zeros((n, 3))
        self.L.orc_slots_export(self.h, C.c_int(n), P(valid), P(res), P(nrm))
        return valid.astype(bool), res, nrm

    def stats(self):
        s = np.zeros(6, np.int64)
        self.L.orc_stats(self.h, P(s))
        return dict(zip(("new_voxels", "evicted", "refits", "merge_calls", "merges", "dropped"), s.tolist()))

    def push(self, pts, covs, is_build=False):
        pts = np.ascontiguousarray(pts, np.float64)
        covs = np.ascontiguousarray(np.transpose(covs, (0, 2, 1)), np.float64)  # column-major 3x3
        self.L.orc_map_push(self.h, P(pts), P(covs), C.c_int(len(pts)), C.c_int(1 if is_build else 0))

    def lidar_to_world_f32(self, x, y, z):
        """Insertion points at the current state (fp32 PCL transform)."""
        x, y, z = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))
        out = np.zeros((3, len(x)), np.float32)
        self.L.orc_lidar_to_world_f32(self.h, P(x), P(y), P(z), C.c_int(len(x)), P(out[0]), P(out[1]), P(out[2]))
        return out.T

    # ---- IMU side of LIOBuilder (lio_builder.cpp:5-153)
    def imu_config(self, na=0.01, ng=0.01, nba=0.0001, nbg=0.0001, imu_init_num=20, gravity_align=True,
                   r_il=None, p_il=None):
        r = np.asfortranarray(np.eye(3) if r_il is None else r_il, np.float64)
        p = np.ascontiguousarray(np.zeros(3) if p_il is None else p_il, np.float64)
        noise = np.array([na, ng, nba, nbg], np.float64)
        self.L.orc_imu_config(self.h, P(noise), C.c_int(imu_init_num), C.c_int(1 if gravity_align else 0),
                              C.c_void_p(r.ctypes.data), P(p))

    def set_scan_resolution(self, leaf):
        self.scan_resolution = float(leaf)
        self.L.orc_set_scan_resolution(self.h, C.c_double(leaf))

    def process(self, x, y, z, t_ms, imu, cloud_start, cloud_end):
        """One SyncPackage through LIOBuilder::process; returns (status, iterations)."""
        x, y, z, t_ms = (np.ascontiguousarray(a, np.float32) for a in (x, y, z, t_ms))
        it, acc, gyro = _imu_arrays(imu)
        iters = C.c_int()
        st = self.L.orc_process(self.h, P(x), P(y), P(z), P(t_ms), len(x), P(it), P(acc), P(gyro), len(it),
                                float(cloud_start), float(cloud_end), C.byref(iters))
        return st, iters.value

    def undistort(self, x, y, z, t_ms, imu, cloud_start, cloud_end):
        x, y, z, t_ms = (np.array(a, np.float32) for a in (x, y, z, t_ms))
        it, acc, gyro = _imu_arrays(imu)
        self.L.orc_undistort(self.h, P(x), P(y), P(z), P(t_ms), C.c_int(len(x)), P(it), P(acc), P(gyro),
                             C.c_int(len(it)), C.c_double(cloud_start), C.c_double(cloud_end))
        return x, y, z, t_ms

    def initialize_imu(self, imu):
        it, acc, gyro = _imu_arrays(imu)
        return bool(self.L.orc_initialize_imu(self.h, P(it), P(acc), P(gyro), len(it)))

    def predict(self, acc, gyro, dt):
        a, g = np.ascontiguousarray(acc, np.float64), np.ascontiguousarray(gyro, np.float64)
        self.L.orc_predict(self.h, P(a), P(g), C.c_double(dt))

    def imu_state(self):
        o = np.zeros(10)
        self.L.orc_imu_state(self.h, P(o))
        return {"last_acc": o[0:3], "last_gyro": o[3:6], "last_cloud_end_time": o[6], "gravity_norm": o[7],
                "last_imu_t": o[8], "status": int(o[9])}

    def imu_poses(self):
        """IMU pose table of the last undistortion: list of dicts."""
        n = self.L.orc_imu_poses(self.h, C.c_int(0), None)
        o = np.zeros((max(n, 1), 22))
        self.L.orc_imu_poses(self.h, C.c_int(n), P(o))
        return [{"offset": r[0], "acc": r[1:4], "gyro": r[4:7], "vel": r[7:10], "pos": r[10:13],
                 "rot": r[13:22].reshape(3, 3).T} for r in o[:n]]

    def scan_points(self, cap=1 << 24):
        n = self.L.orc_scan_points(self.h, C.c_int(0), None, None, None)
        out = np.zeros((3, max(n, 1)), np.float32)
        self.L.orc_scan_points(self.h, C.c_int(n), P(out[0]), P(out[1]), P(out[2]))
        return out[:, :n]

    def key(self, p):
        k = np.zeros(3, np.int64)
        self.L.orc_key(self.h, P(np.ascontiguousarray(p, np.float64)), P(k))
        return k


# ---- math entry points (column-major 3x3 in/out, row-major n x n inverse)
def eigh3(A):
    L = lib()
    Ac = np.asfortranarray(A, np.float64)
    ev = np.zeros(3)
    V = np.zeros((3, 3), order="F")
    L.orc_eigh3(C.c_void_p(Ac.ctypes.data), P(ev), C.c_void_p(V.ctypes.data))
    return ev, np.ascontiguousarray(V)


def invert(A):
    L = lib()
    A = np.ascontiguousarray(A, np.float64)
    out = np.zeros_like(A)
    rc = L.orc_invert(P(A), P(out), C.c_int(A.shape[0]))
    if rc != 0:
        raise np.linalg.LinAlgError("singular")
    return out


def _m3(fn, v):
    L = lib()
    out = np.zeros((3, 3), order="F")
    getattr(L, fn)(P(np.ascontiguousarray(v, np.float64)), C.c_void_p(out.ctypes.data))
    return np.ascontiguousarray(out)


def so3_exp(w):
    return _m3("orc_so3_exp", w)


def left_jacobian(w):
    return _m3("orc_left_jacobian", w)


def so3_log(R):
    L = lib()
    Rc = np.asfortranarray(R, np.float64)
    out = np.zeros(3)
    L.orc_so3_log(C.c_void_p(Rc.ctypes.data), P(out))
    return out


def calc_body_cov(p, ranging=0.04, angle_deg=0.1):
    L = lib()
    p = np.array(p, np.float64)
    out = np.zeros((3, 3), order="F")
    L.orc_calc_body_cov(P(p), C.c_double(ranging), C.c_double(angle_deg), C.c_void_p(out.ctypes.data))
    return p, np.ascontiguousarray(out)


def boxplus(s, d):
    L = lib()
    out = np.zeros(36)
    L.orc_boxplus(P(np.ascontiguousarray(s, np.float64)), P(np.ascontiguousarray(d, np.float64)), P(out))
    return out


def boxminus(a, b):
    L = lib()
    out = np.zeros(23)
    L.orc_boxminus(P(np.ascontiguousarray(a, np.float64)), P(np.ascontiguousarray(b, np.float64)), P(out))
    return out


def voxel_grid(x, y, z, leaf):
    """PCL VoxelGrid restated (oracle): the downsampled cloud, 3 x n_out."""
    L = lib()
    x, y, z = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))
    out = np.zeros((3, max(len(x), 1)), np.float32)
    n = L.orc_voxel_grid(P(x), P(y), P(z), C.c_int(len(x)), C.c_float(leaf), P(out[0]), P(out[1]), P(out[2]))
    return out[:, :n]


def rot_from_two_vectors(a, b):
    L = lib()
    R = np.zeros((3, 3), order="F")
    av, bv = np.ascontiguousarray(a, np.float64), np.ascontiguousarray(b, np.float64)
    L.orc_rot_from_two_vectors(P(av), P(bv), C.c_void_p(R.ctypes.data))
    return np.ascontiguousarray(R)
