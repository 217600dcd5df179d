"""In-tree build of the gfx950 library and its PyTorch binding.

* ``dmsc_amd/lib/libdmsc_hip.so`` -- every HIP translation unit under dmsc_amd/csrc, compiled with
  hipcc for gfx950 only and linked into one C-ABI shared library (include/dmsc_hip.h).
* ``dmsc_amd/csrc/dmsc_gpu<EXT_SUFFIX>`` -- the pybind11 / torch binding, compiled with g++ against
  torch's ROCm headers and linked to libdmsc_hip.so through an $ORIGIN rpath.

Both land inside the package so they travel with the repository snapshot; nothing is written to a
JIT cache.
"""

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
BUILD = os.path.join(PKG, "_build")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_SOURCES = ["k_grid.hip", "k_prims.hip", "k_topo.hip", "k_arcs.hip", "k_diff.hip", "dmsc_hip.cpp"]
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed fp32 instructions are disabled for every kernel: some of their forms have produced
    # silently wrong results on this hardware
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fno-gpu-flush-denormals-to-zero", "-ffp-contract=off", "-Wno-unused-command-line-argument",
    "-I", INCLUDE, "-I", CSRC,
]

LIB_PATH = os.path.join(LIBDIR, "libdmsc_hip.so")
EXT_PATH = os.path.join(CSRC, "dmsc_gpu" + sysconfig.get_config_var("EXT_SUFFIX"))


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and "unknown target feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("build step failed: " + " ".join(cmd) + "\n" + out)
    return out


def build_hip_library(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)

    def one(src):
        obj = os.path.join(BUILD, src + ".o")
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        _run([HIPCC, *HIP_FLAGS, *lang, "-c", os.path.join(CSRC, src), "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(one, HIP_SOURCES))
    tmp = LIB_PATH + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


def build_torch_binding():
    import torch
    import torch.utils.cpp_extension as ce

    inc = []
    for p in ce.include_paths():
        inc += ["-I", p]
    py_inc = sysconfig.get_paths()["include"]
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    tlib = ce.library_paths()[0]
    cmd = [
        "g++", "-O2", "-std=c++17", "-fPIC", "-shared", os.path.join(CSRC, "dmsc_gpu.cpp"),
        *inc, "-I", py_inc, "-I", INCLUDE, "-I", "/opt/rocm/include",
        "-D__HIP_PLATFORM_AMD__=1", "-DUSE_ROCM=1", f"-D_GLIBCXX_USE_CXX11_ABI={abi}",
        "-DTORCH_EXTENSION_NAME=dmsc_gpu", "-DTORCH_API_INCLUDE_EXTENSION_H",
        "-L", tlib, "-lc10", "-lc10_hip", "-ltorch", "-ltorch_cpu", "-ltorch_hip", "-ltorch_python",
        "-L", LIBDIR, "-ldmsc_hip", "-L", "/opt/rocm/lib", "-lamdhip64",
        "-Wl,-rpath,$ORIGIN/../lib", "-Wl,-rpath," + tlib, "-Wl,-rpath,/opt/rocm/lib",
        "-o", EXT_PATH + ".tmp",
    ]
    _run(cmd)
    os.replace(EXT_PATH + ".tmp", EXT_PATH)
    return EXT_PATH


def build(jobs=8):
    """Compile the HIP library and the torch binding (always rebuilds)."""
    lib = build_hip_library(jobs)
    ext = build_torch_binding()
    return lib, ext


if __name__ == "__main__":
    print(build(int(sys.argv[1]) if len(sys.argv) > 1 else 8))
