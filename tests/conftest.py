import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (runs the HIP path)")


def gpu_available():
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def lut_lib():
    """Host build of the kernel's lower-star table code (tests/native/lut_gradient.cpp)."""
    import ctypes

    src = os.path.join(ROOT, "tests", "native", "lut_gradient.cpp")
    out = os.path.join(ROOT, "tests", "native", "liblut_gradient.so")
    if not os.path.exists(out) or os.path.getmtime(out) < max(
        os.path.getmtime(src), os.path.getmtime(os.path.join(ROOT, "dmsc_amd", "csrc", "dmsc_common.hpp"))
    ):
        subprocess.check_call([
            "/opt/rocm/bin/hipcc", "-x", "hip", "--offload-arch=gfx950", "-O2", "-std=c++17", "-fPIC", "-shared",
            "-I", os.path.join(ROOT, "dmsc_amd", "csrc"), src, "-o", out,
        ], stderr=subprocess.DEVNULL)
    lib = ctypes.CDLL(out)
    lib.lut_gradient.restype = None
    lib.lut_gradient.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    return lib
