"""CPU oracle for the compute_dmsc path -- TEST INFRASTRUCTURE ONLY.

A from-scratch C++ restatement of the reference CPU (TBB) pipeline in canonical single-thread
order (``dmsc_oracle.cpp``; each stage cites the reference file:line it restates). It is the
checker that the HIP path is compared against, and the CPU baseline that bench.py times. Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg may import it; the
product package ``dmsc_amd`` never does.

Parity pins: the reference's 13 fixtures (tests/golden), exhaustive lower-star enumeration, an
independent Kruskal/elder-rule persistence, brute-force V-path labels and invariants
(tests/test_oracle.py); rank invariance of the gradient on tied fields, the arc layout (starts,
adjacency, offsets) and an independent list-based arc merging with thresholded cancellations
(tests/test_oracle_arcs.py, tests/arc_reference.py).
"""

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SRC = os.path.join(_HERE, "dmsc_oracle.cpp")
_LIB_PATH = os.path.join(_HERE, "liborc.so")
_lib = None

FLAG_RETURN_GRADIENT = 1
FLAG_DUAL = 2
FLAG_TRACE_MAX_ARCS = 4
FLAG_TRACE_MIN_ARCS = 8
FLAG_TRACE_MAX_GROUPS = 16
FLAG_TRACE_MIN_GROUPS = 32

FIELDS = (
    "shape", "max_pts", "min_pts", "sad_pts", "e_max", "e_min", "p_max", "p_min", "ppairs_max",
    "ppairs_min", "grad_indices", "peaks", "ridges", "ridge_offsets", "ridge_arc_offsets", "basins",
    "valleys", "valley_offsets", "valley_arc_offsets",
)


class OracleError(RuntimeError):
    pass


def build():
    """Compile the oracle into oracle/liborc.so (plain g++, no torch, no fast-math)."""
    cmd = [
        "g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-fno-fast-math",
        "-ffp-contract=off", "-o", _LIB_PATH, _SRC,
    ]
    subprocess.check_call(cmd)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH) or os.path.getmtime(_LIB_PATH) < os.path.getmtime(_SRC):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    lib.orc_extract.restype = ctypes.c_void_p
    lib.orc_extract.argtypes = [
        ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_uint, ctypes.c_int,
        ctypes.POINTER(ctypes.c_int),
    ]
    lib.orc_get.restype = ctypes.c_int
    lib.orc_get.argtypes = [
        ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_int64),
        ctypes.POINTER(ctypes.c_int),
    ]
    lib.orc_free.restype = None
    lib.orc_free.argtypes = [ctypes.c_void_p]
    lib.orc_gradient.restype = None
    lib.orc_gradient.argtypes = [
        ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
    ]
    _lib = lib
    return lib


def _as_field(img):
    a = np.ascontiguousarray(np.asarray(img, dtype=np.float32))
    if a.ndim != 2 or a.shape[0] == 0 or a.shape[1] == 0:
        raise ValueError("oracle expects a non-empty 2-D float32 field")
    return a


def extract(field, persistence_threshold, return_gradient=False, is_dual=False, trace_max_arcs=False,
            trace_min_arcs=False, trace_max_groups=False, trace_min_groups=False, nthreads=1):
    """Reference ``extract_dmsc`` argument meaning (dmsc_cpu.cpp:53-55); returns the 19 outputs as a
    list of numpy arrays in DMSComplex slot order."""
    lib = _load()
    a = _as_field(field)
    H, W = a.shape
    flags = (
        (FLAG_RETURN_GRADIENT if return_gradient else 0) | (FLAG_DUAL if is_dual else 0)
        | (FLAG_TRACE_MAX_ARCS if trace_max_arcs else 0) | (FLAG_TRACE_MIN_ARCS if trace_min_arcs else 0)
        | (FLAG_TRACE_MAX_GROUPS if trace_max_groups else 0)
        | (FLAG_TRACE_MIN_GROUPS if trace_min_groups else 0)
    )
    st = ctypes.c_int(0)
    h = lib.orc_extract(a.ctypes.data, H, W, float(persistence_threshold), flags, int(nthreads), ctypes.byref(st))
    try:
        if st.value != 0:
            raise OracleError("The cancellation tree is too deep to merge path geometries.")
        out = []
        for slot in range(19):
            ptr = ctypes.c_void_p()
            dims = (ctypes.c_int64 * 2)()
            isf = ctypes.c_int(0)
            nd = lib.orc_get(h, slot, ctypes.byref(ptr), dims, ctypes.byref(isf))
            shape = (dims[0],) if nd == 1 else (dims[0], dims[1])
            n = int(np.prod(shape))
            dt = np.float32 if isf.value else np.int32
            if n == 0:
                arr = np.zeros(shape, dtype=dt)
            else:
                ctype = ctypes.c_float if isf.value else ctypes.c_int32
                buf = ctypes.cast(ptr, ctypes.POINTER(ctype * n)).contents
                arr = np.frombuffer(buf, dtype=dt, count=n).reshape(shape).copy()
            out.append(arr)
        return out
    finally:
        lib.orc_free(h)


def compute_dmsc(img, persistence_threshold, return_gradient=True, is_dual=False, trace_valleys=True,
                 trace_ridges=True, trace_peaks=True, trace_basins=True, nthreads=1):
    """Python-API flag mapping of the reference (dmsc.py:230-276) -> dict of numpy arrays."""
    out = extract(img, persistence_threshold, return_gradient, is_dual, trace_max_arcs=trace_ridges,
                  trace_min_arcs=trace_valleys, trace_max_groups=trace_peaks, trace_min_groups=trace_basins,
                  nthreads=nthreads)
    return dict(zip(FIELDS, out))


def gradient(field, is_dual=False, nthreads=1):
    """Raw paired_with array (reference gradient.hxx:19-196) as int32[4(H+1)(W+1)]."""
    lib = _load()
    a = _as_field(field)
    H, W = a.shape
    pw = np.empty(4 * (H + 1) * (W + 1), dtype=np.int32)
    lib.orc_gradient(a.ctypes.data, H, W, int(bool(is_dual)), int(nthreads), pw.ctypes.data)
    return pw
