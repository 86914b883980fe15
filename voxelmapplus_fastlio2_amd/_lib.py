"""ctypes bindings of the C ABI (include/vmp.h, include/vmp_synth.h).

The HIP library is mandatory: importing the registration API without a built
``_build/libvmp.so`` raises instead of falling back to any CPU path.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
BUILD_DIR = os.path.join(_HERE, "_build")
LIBVMP = os.path.join(BUILD_DIR, "libvmp.so")
LIBSYNTH = os.path.join(BUILD_DIR, "libvmp_synth.so")


class MapCfg(C.Structure):
    """vmp_map_cfg — LIOConfig subset (lio_builder.h:17-42) + device limits."""

    _fields_ = [
        ("voxel_size", C.c_double),
        ("plane_thresh", C.c_double),
        ("merge_thresh_for_angle", C.c_double),
        ("merge_thresh_for_distance", C.c_double),
        ("ranging_cov", C.c_double),
        ("angle_cov", C.c_double),
        ("max_point_thresh", C.c_int32),
        ("update_size_thresh", C.c_int32),
        ("map_capacity", C.c_int32),
        ("max_scan_points", C.c_int32),
        ("max_iter", C.c_int32),
        ("estimate_ext", C.c_int32),
        ("scan_resolution", C.c_double),
    ]


class ImuCfg(C.Structure):
    """vmp_imu_cfg — the IMU side of LIOConfig (lio_builder.h:19-28)."""

    _fields_ = [
        ("na", C.c_double),
        ("ng", C.c_double),
        ("nba", C.c_double),
        ("nbg", C.c_double),
        ("imu_init_num", C.c_int32),
        ("gravity_align", C.c_int32),
        ("r_il", C.c_double * 9),
        ("p_il", C.c_double * 3),
    ]


class State(C.Structure):
    """vmp_state — kf::State (ieskf.h:31-62), rotations column-major."""

    _fields_ = [
        ("pos", C.c_double * 3),
        ("rot", C.c_double * 9),
        ("rot_ext", C.c_double * 9),
        ("pos_ext", C.c_double * 3),
        ("vel", C.c_double * 3),
        ("bg", C.c_double * 3),
        ("ba", C.c_double * 3),
        ("g", C.c_double * 3),
    ]


class ObsStats(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("points", "hits", "plane_hits", "effective", "key_fallback", "iterations")]


class MapStats(C.Structure):
    _fields_ = [
        (n, C.c_int32)
        for n in (
            "points",
            "touched",
            "created",
            "evicted",
            "refits",
            "merge_calls",
            "merge_evals",
            "merges",
            "dropped",
            "map_size",
            "maintenance",
            "merge_windows",
            "merge_requeued",
            "merge_max_acting",
        )
    ]


_vmp = None
_synth = None


def _proto(lib, name, restype, argtypes):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = argtypes
    return f


def vmp():
    """The HIP registration library; raises if it has not been built."""
    global _vmp
    if _vmp is not None:
        return _vmp
    if not os.path.exists(LIBVMP):
        raise RuntimeError(
            f"HIP library missing: {LIBVMP}. Build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` (hipcc, gfx950)."
        )
    lib = C.CDLL(LIBVMP)
    P = C.c_void_p
    pp = C.POINTER(C.c_void_p)
    dp = C.POINTER(C.c_double)
    fp = C.POINTER(C.c_float)
    _proto(lib, "vmp_version", C.c_char_p, [])
    _proto(lib, "vmp_ctx_create", C.c_int, [C.c_int, pp])
    _proto(lib, "vmp_ctx_destroy", C.c_int, [P])
    _proto(lib, "vmp_session_create", C.c_int, [P, C.POINTER(MapCfg), pp])
    _proto(lib, "vmp_session_destroy", C.c_int, [P])
    _proto(lib, "vmp_session_bytes", C.c_int64, [C.POINTER(MapCfg)])
    _proto(lib, "vmp_last_error", C.c_char_p, [P])
    _proto(lib, "vmp_state_set", C.c_int, [P, C.POINTER(State), dp])
    _proto(lib, "vmp_state_get", C.c_int, [P, C.POINTER(State), dp])
    _proto(lib, "vmp_state_apply_motion", C.c_int, [P, dp, dp, dp])
    _proto(lib, "vmp_state_apply_motion_batch", C.c_int, [pp, C.c_int32, dp, dp, dp])
    _proto(lib, "vmp_scan_set_points", C.c_int, [P, fp, fp, fp, C.c_int32])
    _proto(lib, "vmp_scan_set_points_device", C.c_int, [P, P, P, P, C.c_int32])
    u64p = C.POINTER(C.c_uint64)
    _proto(lib, "vmp_scan_set_points_device_batch", C.c_int, [pp, C.c_int32, u64p, u64p, u64p, C.POINTER(C.c_int32)])
    _proto(lib, "vmp_observe", C.c_int, [P, C.POINTER(State), dp, dp, C.POINTER(ObsStats)])
    _proto(lib, "vmp_map_update", C.c_int, [P, C.c_int, C.POINTER(MapStats)])
    _proto(lib, "vmp_ieskf_update", C.c_int, [P, C.POINTER(ObsStats)])
    _proto(lib, "vmp_register_scan", C.c_int, [P, C.POINTER(ObsStats), C.POINTER(MapStats)])
    _proto(lib, "vmp_register_scan_noalloc", C.c_int, [P, C.POINTER(ObsStats), C.POINTER(MapStats)])
    _proto(lib, "vmp_register_batch", C.c_int, [pp, C.c_int32, C.POINTER(ObsStats), C.POINTER(MapStats)])
    _proto(lib, "vmp_register_batch_submit", C.c_int, [pp, C.c_int32])
    _proto(lib, "vmp_register_batch_wait", C.c_int, [P, C.c_int32, C.POINTER(ObsStats), C.POINTER(MapStats)])
    _proto(
        lib,
        "vmp_map_export",
        C.c_int64,
        [P, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_uint64), dp],
    )
    _proto(lib, "vmp_map_size", C.c_int64, [P])
    _proto(lib, "vmp_slots_export", C.c_int, [P, C.c_int32, C.POINTER(C.c_uint8), fp, fp])
    _proto(lib, "vmp_session_timing", C.c_int, [P, dp, dp, C.POINTER(C.c_int32)])
    _proto(lib, "vmp_session_set_timing", C.c_int, [P, C.c_int])
    _proto(lib, "vmp_ctx_set_param", C.c_int, [P, C.c_char_p, C.c_int64])
    _proto(lib, "vmp_ctx_graph_stats", C.c_int, [P, C.POINTER(C.c_int64)])
    _proto(lib, "vmp_ctx_reserve", C.c_int, [P, C.c_int64, C.c_int32])
    _proto(lib, "vmp_autotune_observe", C.c_int, [P, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_float)])
    _proto(lib, "vmp_autotune_observe_batch", C.c_int,
           [pp, C.c_int32, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_float)])
    i64p = C.POINTER(C.c_int64)
    _proto(lib, "vmp_session_timing_totals", C.c_int, [P, dp, i64p, dp, i64p, i64p, i64p])
    i32p = C.POINTER(C.c_int32)
    _proto(lib, "vmp_imu_init", C.c_int, [P, C.POINTER(ImuCfg), dp, dp, dp, C.c_int32, C.c_double, i32p])
    _proto(lib, "vmp_scan_set_raw", C.c_int,
           [P, fp, fp, fp, fp, C.c_int32, dp, dp, dp, C.c_int32, C.c_double, C.c_double])
    _proto(lib, "vmp_scan_set_raw_device", C.c_int,
           [P, P, P, P, P, C.c_int32, dp, dp, dp, C.c_int32, C.c_double, C.c_double])
    _proto(lib, "vmp_scan_set_raw_device_batch", C.c_int,
           [pp, C.c_int32, u64p, u64p, u64p, u64p, i32p, dp, dp, dp, i32p, dp, dp])
    _proto(lib, "vmp_scan_export", C.c_int32, [P, C.c_int32, fp, fp, fp])
    _proto(lib, "vmp_cloud_body_export", C.c_int32, [P, C.c_int32, fp, fp, fp])
    _proto(lib, "vmp_odom_export", C.c_int, [P, dp, dp, dp])
    _proto(lib, "vmp_ctx_kernel_times", C.c_int32, [P, C.c_int32, dp, i64p])
    _vmp = lib
    return lib


def synth():
    """The seeded synthetic sequence generator (host C++)."""
    global _synth
    if _synth is not None:
        return _synth
    if not os.path.exists(LIBSYNTH):
        raise RuntimeError(f"generator library missing: {LIBSYNTH}; run __graft_entry__.build()")
    lib = C.CDLL(LIBSYNTH)
    P = C.c_void_p
    dp = C.POINTER(C.c_double)
    fp = C.POINTER(C.c_float)
    _proto(lib, "vmp_synth_create", C.c_int, [C.c_int, C.c_uint64, C.c_int32, dp, C.POINTER(C.c_void_p)])
    _proto(lib, "vmp_synth_destroy", None, [P])
    _proto(lib, "vmp_synth_pose", C.c_int, [P, C.c_int32, dp, dp])
    _proto(lib, "vmp_synth_scan", C.c_int, [P, C.c_int32, fp, fp, fp, fp])
    _proto(lib, "vmp_synth_scan_n", C.c_int, [P, C.c_int32, C.c_int32, fp, fp, fp, fp])
    _proto(lib, "vmp_synth_motion", C.c_int, [P, C.c_int32, dp, dp])
    _proto(lib, "vmp_synth_create_imu", C.c_int, [C.c_int, C.c_uint64, C.c_int32, dp, C.POINTER(C.c_void_p)])
    _proto(lib, "vmp_synth_pose_at", C.c_int, [P, C.c_double, dp, dp])
    _proto(lib, "vmp_synth_imu", C.c_int, [P, C.c_int32, C.c_int32, dp, dp, dp])
    _proto(lib, "vmp_synth_scan_raw", C.c_int, [P, C.c_int32, C.c_int32, fp, fp, fp, fp])
    _synth = lib
    return lib
