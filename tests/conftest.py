import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (ROOT, HERE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libeif_gpu.so")
    config.addinivalue_line("markers", "slow: longer CPU test")


def _have_gpu():
    try:
        import torch  # noqa: F401  (device count only; no context created)
        return torch.cuda.device_count() > 0
    except Exception:
        return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library; GPU tests fail loudly if it is missing."""
    import ellipsesinflows_amd as eif
    lib = eif.load_library()
    return lib
