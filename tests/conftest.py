import os
import sys

# torch bundles its own HIP runtime; a process that uses both torch device
# memory and libvmp must load torch first, so that both share that runtime
# (loading libvmp first makes torch find no GPU).
try:
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU test")
