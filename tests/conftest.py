import ctypes
import os
import sys

import pytest

# torch (used by the device-pointer tests) and liblshgpu.so share one HIP
# runtime per process (same soname); torch must be the one that loads it.
try:
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (runs the HIP kernels)")


_GPU = None


def gpu_count() -> int:
    global _GPU
    if _GPU is None:
        if torch is not None:
            _GPU = torch.cuda.device_count()
        else:
            try:
                hip = ctypes.CDLL("libamdhip64.so.7")
                n = ctypes.c_int(0)
                _GPU = n.value if hip.hipGetDeviceCount(ctypes.byref(n)) == 0 else 0
            except OSError:
                _GPU = 0
    return _GPU


def pytest_runtest_setup(item):
    if "gpu" in item.keywords and gpu_count() == 0:
        pytest.skip("no GPU visible")


@pytest.fixture(scope="session")
def lsh():
    import lshongpu_amd

    lshongpu_amd.load()
    return lshongpu_amd


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.load()
    return oracle
