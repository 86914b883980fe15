"""Seeded synthetic sequences (room / street) — see include/vmp_synth.h."""
import ctypes as C

import numpy as np

from . import _lib

ROOM, STREET = 0, 1
SEED0 = 0x5EED0000

# Motion-prior process noise added to P between scans (stand-in for IMU
# propagation, which is outside the registration path): 2 cm and 0.2 deg
# per axis match the generator's odometry noise.
QDIAG = np.zeros(23)
QDIAG[0:3] = 4e-4
QDIAG[3:6] = 1.2e-5
QDIAG[6:12] = 1e-10
QDIAG[12:15] = 1e-4
QDIAG[15:23] = 1e-8


class Sequence:
    def __init__(self, scene=ROOM, seed=SEED0, n=20000, origin_offset=None):
        self.lib = _lib.synth()
        self.n = int(n)
        self.scene = scene
        off = None
        if origin_offset is not None:
            self._off = np.ascontiguousarray(origin_offset, dtype=np.float64)
            off = self._off.ctypes.data_as(C.POINTER(C.c_double))
        h = C.c_void_p()
        rc = self.lib.vmp_synth_create(scene, seed, self.n, off, C.byref(h))
        if rc != 0:
            raise RuntimeError(f"vmp_synth_create failed ({rc})")
        self.h = h

    def close(self):
        if self.h:
            self.lib.vmp_synth_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def scan(self, k, n=None):
        n = self.n if n is None else int(n)
        x = np.empty(n, np.float32)
        y = np.empty(n, np.float32)
        z = np.empty(n, np.float32)
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        rc = self.lib.vmp_synth_scan_n(self.h, k, n, fp(x), fp(y), fp(z), None)
        if rc != 0:
            raise RuntimeError(f"vmp_synth_scan failed ({rc})")
        return x, y, z

    def pose(self, k):
        R = np.zeros(9)
        t = np.zeros(3)
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        rc = self.lib.vmp_synth_pose(self.h, k, dp(R), dp(t))
        if rc != 0:
            raise RuntimeError("vmp_synth_pose failed")
        return R.reshape(3, 3).T.copy(), t

    def motion(self, k):
        dR = np.zeros(9)
        dpv = np.zeros(3)
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        rc = self.lib.vmp_synth_motion(self.h, k, dp(dR), dp(dpv))
        if rc != 0:
            raise RuntimeError("vmp_synth_motion failed")
        return dR.reshape(3, 3).T.copy(), dpv


IMU_DT = 0.005  # 200 Hz
SCAN_DT = 0.1   # 10 Hz


class ImuSequence:
    """A sequence in IMU mode (include/vmp_synth.h): raw scans whose points
    are taken along each 0.1 s sweep, and the 200 Hz IMU stream.  package(k)
    forms the k-th SyncPackage the way the reference node does
    (lio_node.cpp:137-165): the cloud with its per-point times, cloud_start
    = 0.1 k, cloud_end = cloud_start + the last point's time, and the IMU
    samples not yet handed out whose time is below cloud_end."""

    def __init__(self, scene=ROOM, seed=SEED0, n=20000, origin_offset=None):
        self.lib = _lib.synth()
        self.n = int(n)
        self.scene = scene
        off = None
        if origin_offset is not None:
            self._off = np.ascontiguousarray(origin_offset, dtype=np.float64)
            off = self._off.ctypes.data_as(C.POINTER(C.c_double))
        h = C.c_void_p()
        rc = self.lib.vmp_synth_create_imu(scene, seed, self.n, off, C.byref(h))
        if rc != 0:
            raise RuntimeError(f"vmp_synth_create_imu failed ({rc})")
        self.h = h
        self.next_imu = 0

    def close(self):
        if self.h:
            self.lib.vmp_synth_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def raw_scan(self, k, n=None):
        n = self.n if n is None else int(n)
        out = np.empty((4, n), np.float32)
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        rc = self.lib.vmp_synth_scan_raw(self.h, k, n, fp(out[0]), fp(out[1]), fp(out[2]), fp(out[3]))
        if rc != 0:
            raise RuntimeError(f"vmp_synth_scan_raw failed ({rc})")
        return out[0], out[1], out[2], out[3]

    def imu(self, j0, count):
        t = np.empty(count)
        acc = np.empty((count, 3))
        gyro = np.empty((count, 3))
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        rc = self.lib.vmp_synth_imu(self.h, j0, count, dp(t), dp(acc), dp(gyro))
        if rc != 0:
            raise RuntimeError(f"vmp_synth_imu failed ({rc})")
        return t, acc, gyro

    def pose_at(self, t):
        R = np.zeros(9)
        p = np.zeros(3)
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        if self.lib.vmp_synth_pose_at(self.h, float(t), dp(R), dp(p)) != 0:
            raise RuntimeError("vmp_synth_pose_at failed")
        return R.reshape(3, 3).T.copy(), p

    def package(self, k, n=None):
        """(x, y, z, t_ms, (imu_t, acc, gyro), cloud_start, cloud_end) of scan
        k; IMU samples are handed out once, in order (call k = 0, 1, ...)."""
        x, y, z, t_ms = self.raw_scan(k, n)
        start = SCAN_DT * k
        end = start + float(t_ms[-1]) / 1000.0
        j1 = self.next_imu
        while IMU_DT * j1 < end:
            j1 += 1
        imu = self.imu(self.next_imu, j1 - self.next_imu)
        self.next_imu = j1
        return x, y, z, t_ms, imu, start, end
