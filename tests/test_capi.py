"""The C ABI library on the CPU (no device calls): it loads, exports every
function include/*.h declares, reports its version and session footprint, and
the Python mirror refuses to run without a GPU or without the library."""
import ctypes as C
import os
import re

import pytest

from voxelmapplus_fastlio2_amd import LIOConfig, _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = {"vmp.h": _lib.LIBVMP, "vmp_serve.h": _lib.LIBVMP, "vmp_synth.h": _lib.LIBSYNTH}


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = re.findall(r"\b(vmp_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


@pytest.mark.parametrize("header", sorted(HEADERS))
def test_library_exports_every_declared_function(header):
    names = declared(header)
    assert len(names) >= 5
    lib = C.CDLL(HEADERS[header])
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"{os.path.basename(HEADERS[header])} lacks {missing}"


def test_version_names_the_target():
    v = _lib.vmp().vmp_version().decode()
    assert "gfx950" in v


def test_session_bytes_grows_with_capacity():
    L = _lib.vmp()
    small = L.vmp_session_bytes(C.byref(LIOConfig(map_capacity=10000, max_scan_points=20000).to_c()))
    big = L.vmp_session_bytes(C.byref(LIOConfig(map_capacity=2000000, max_scan_points=120000).to_c()))
    assert 0 < small < big
    # a 2M-voxel street session (FP64 moments and the pending-point lists
    # included) stays under 3 GiB of HBM
    assert big < 3 << 30


def test_context_without_gpu_raises():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from voxelmapplus_fastlio2_amd import Context

    with pytest.raises(RuntimeError):
        Context(0)


def test_missing_library_raises(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "LIBVMP", str(tmp_path / "libvmp.so"))
    monkeypatch.setattr(_lib, "_vmp", None)
    with pytest.raises(RuntimeError, match="libvmp.so"):
        _lib.vmp()


def test_config_defaults_are_the_reference_effective_values():
    c = LIOConfig()
    assert (c.voxel_size, c.plane_thresh, c.max_point_thresh, c.update_size_thresh) == (0.5, 0.01, 100, 10)
    assert (c.merge_thresh_for_angle, c.merge_thresh_for_distance) == (0.2, 0.04)
    assert (c.ranging_cov, c.angle_cov, c.max_iter, c.estimate_ext) == (0.04, 0.1, 5, 0)
    assert c.map_capacity == 100000
