"""Continuous-batching server (include/vmp_serve.h) from Python.

The server and its scheduler are C++ inside libvmp.so; this module only
builds request tables and reads results.  Sessions keep their voxel maps and
filter states in HBM under a byte budget, with least-recently-used session
eviction; `Server.run` replays requests in real time by arrival.
"""
import ctypes as C

import numpy as np

from . import _lib
from .registration import Registrar, VmpError, _state_c


class ServerCfg(C.Structure):
    _fields_ = [
        ("device", C.c_int32),
        ("executors", C.c_int32),
        ("max_batch_sessions", C.c_int32),
        ("max_batch_points", C.c_int64),
        ("hbm_budget", C.c_int64),
    ]


class Request(C.Structure):
    _fields_ = [
        ("session", C.c_int32),
        ("n", C.c_int32),
        ("x", C.c_void_p),
        ("y", C.c_void_p),
        ("z", C.c_void_p),
        ("dR", C.c_double * 9),
        ("dp", C.c_double * 3),
        ("arrival", C.c_double),
        ("start", C.c_double),
        ("done", C.c_double),
        ("status", C.c_int32),
        ("kind", C.c_int32),
        ("batch", C.c_int32),
        ("batch_sessions", C.c_int32),
        ("iterations", C.c_int32),
        ("pad", C.c_int32),
        ("rot", C.c_double * 9),
        ("pos", C.c_double * 3),
    ]


def _lib_serve():
    L = _lib.vmp()
    if not hasattr(L, "_serve_ready"):
        P = C.c_void_p
        _lib._proto(L, "vmp_server_create", C.c_int, [C.POINTER(ServerCfg), C.POINTER(_lib.MapCfg), C.c_int32, C.POINTER(P)])
        _lib._proto(L, "vmp_server_destroy", C.c_int, [P])
        _lib._proto(L, "vmp_server_session_init", C.c_int, [P, C.c_int32, C.POINTER(_lib.State), C.POINTER(C.c_double)])
        _lib._proto(L, "vmp_server_run", C.c_int, [P, C.POINTER(Request), C.c_int32, C.POINTER(C.c_double)])
        _lib._proto(L, "vmp_server_session", P, [P, C.c_int32])
        _lib._proto(L, "vmp_server_stats", C.c_int, [P, C.POINTER(C.c_int64)])
        _lib._proto(L, "vmp_server_last_error", C.c_char_p, [P])
        L._serve_ready = True
    return L


class _Borrowed(Registrar):
    """Read-only view of a server-owned session (map / state export)."""

    def __init__(self, lib, handle, cfg):
        self.lib = lib
        self.h = C.c_void_p(handle)
        self.cfg = cfg
        self.ctx = None

    def close(self):  # owned by the server
        self.h = None


class Server:
    def __init__(self, cfg, nsessions, executors=2, max_batch_sessions=128, max_batch_points=8_000_000,
                 hbm_budget=200 << 30, device=0):
        self.L = _lib_serve()
        self.cfg = cfg
        sc = ServerCfg(device, executors, max_batch_sessions, max_batch_points, int(hbm_budget))
        h = C.c_void_p()
        rc = self.L.vmp_server_create(C.byref(sc), C.byref(cfg.to_c()), int(nsessions), C.byref(h))
        if rc != 0:
            raise VmpError(f"vmp_server_create failed ({rc})")
        self.h = h
        self.n = int(nsessions)

    def close(self):
        if self.h:
            self.L.vmp_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def session_init(self, i, s, P):
        Pc = np.asfortranarray(P, dtype=np.float64)
        rc = self.L.vmp_server_session_init(
            self.h, int(i), C.byref(_state_c(s)), Pc.ctypes.data_as(C.POINTER(C.c_double))
        )
        if rc != 0:
            raise VmpError(f"vmp_server_session_init failed ({rc})")

    def run(self, reqs, qdiag):
        """reqs: ctypes array of Request sorted by arrival; filled in place."""
        q = np.ascontiguousarray(qdiag, dtype=np.float64)
        rc = self.L.vmp_server_run(self.h, reqs, len(reqs), q.ctypes.data_as(C.POINTER(C.c_double)))
        if rc != 0:
            raise VmpError(f"vmp_server_run failed ({rc}): {self.L.vmp_server_last_error(self.h).decode()}")

    def session(self, i):
        p = self.L.vmp_server_session(self.h, int(i))
        return None if not p else _Borrowed(self.L, p, self.cfg)

    def stats(self):
        out = (C.c_int64 * 8)()
        self.L.vmp_server_stats(self.h, out)
        keys = ("batches", "requests", "evictions", "admissions", "resident", "resident_bytes", "max_batch", "executors")
        return dict(zip(keys, list(out)))


def make_requests(items):
    """items: dicts with session, n, x/y/z device addresses, dR (3x3), dp, arrival."""
    arr = (Request * len(items))()
    for r, it in zip(arr, items):
        r.session = it["session"]
        r.n = it["n"]
        r.x, r.y, r.z = it["x"], it["y"], it["z"]
        dR = np.asarray(it.get("dR", np.eye(3)), np.float64).T.reshape(-1)  # column-major
        for k in range(9):
            r.dR[k] = dR[k]
        dp = it.get("dp", (0.0, 0.0, 0.0))
        for k in range(3):
            r.dp[k] = dp[k]
        r.arrival = it["arrival"]
    return arr
