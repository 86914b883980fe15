"""MI355X registration path of voxel_plus (VoxelMapPlus + FAST-LIO2 IESKF).

The product is the HIP library behind include/vmp.h; this package holds its
kernels (csrc/), the in-tree build (build.py), ctypes bindings and a Python
mirror of the reference interface (registration.py), and the seeded synthetic
sequence generator (synth.py).
"""
from .registration import (
    Context,
    ImuConfig,
    LIOBuilder,
    LIOConfig,
    Registrar,
    VmpError,
    apply_motion_batch,
    autotune_observe_batch,
    initial_covariance,
    register_batch,
    register_batch_submit,
    register_batch_wait,
    set_points_device_batch,
    set_raw_device_batch,
    state_vector,
)

__all__ = [
    "Context",
    "ImuConfig",
    "LIOBuilder",
    "LIOConfig",
    "Registrar",
    "VmpError",
    "apply_motion_batch",
    "autotune_observe_batch",
    "initial_covariance",
    "register_batch",
    "register_batch_submit",
    "register_batch_wait",
    "set_points_device_batch",
    "set_raw_device_batch",
    "state_vector",
]
