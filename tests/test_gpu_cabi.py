"""The C-ABI driven the way INTEGRATION.md shows a non-PyTorch caller doing it: ctypes only, device
buffers from hipMalloc through the allocator callback, dmsc_extract on one field and
dmsc_extract_batch on fields of different shapes and thresholds (whose group outputs are shared
buffers, allocated with field == -1 and sliced per field), every output read back with hipMemcpy
and compared with the oracle (reference binding: dmsc/csrc/dmsc_gpu.cpp:128-254)."""

import ctypes

import numpy as np
import pytest

from dmsc_amd import synthetic
from tests.conftest import ROOT
from tests.parity_util import diff_outputs, oracle_extract

pytestmark = pytest.mark.gpu

HIP_H2D, HIP_D2H = 1, 2
ALLOC = ctypes.CFUNCTYPE(ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                         ctypes.POINTER(ctypes.c_int64), ctypes.c_int32)


class Outputs(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p * 19), ("dims", (ctypes.c_int64 * 2) * 19),
                ("ndim", ctypes.c_int32 * 19), ("dtype", ctypes.c_int32 * 19), ("shape", ctypes.c_int32 * 2)]


class Hip:
    def __init__(self):
        self.rt = ctypes.CDLL("libamdhip64.so")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.live = []

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(int(nbytes), 4))) == 0
        self.live.append(p.value)
        return p.value

    def to_device(self, a):
        a = np.ascontiguousarray(a)
        p = self.malloc(a.nbytes)
        assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, HIP_H2D) == 0
        return p

    def to_host(self, p, shape, dtype):
        a = np.empty(shape, dtype=dtype)
        if a.size:
            assert self.rt.hipMemcpy(a.ctypes.data, p, a.nbytes, HIP_D2H) == 0
        return a

    def free_all(self):
        for p in self.live:
            self.rt.hipFree(p)
        self.live.clear()


def load_lib():
    lib = ctypes.CDLL(ROOT + "/dmsc_amd/lib/libdmsc_hip.so")
    lib.dmsc_context_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    lib.dmsc_context_destroy.argtypes = [ctypes.c_void_p]
    lib.dmsc_last_error.argtypes = [ctypes.c_void_p]
    lib.dmsc_last_error.restype = ctypes.c_char_p
    lib.dmsc_extract.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                                 ctypes.c_uint32, ctypes.c_void_p, ALLOC, ctypes.c_void_p, ctypes.POINTER(Outputs)]
    lib.dmsc_extract_batch.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ALLOC,
                                       ctypes.c_void_p, ctypes.POINTER(Outputs)]
    return lib


def read_outputs(hip, out, H, W):
    res = [np.array([H, W], np.int32)]
    for k in range(1, 19):
        shape = tuple(out.dims[k][i] for i in range(out.ndim[k]))
        dt = np.float32 if out.dtype[k] == 1 else np.int32
        res.append(hip.to_host(out.data[k], shape, dt) if out.data[k] else np.zeros(shape, dt))
    return res


@pytest.mark.parametrize("is_dual", [False, True])
def test_ctypes_binding_single_and_batch(is_dual):
    hip, lib = Hip(), load_lib()
    requests = []

    def alloc(user, field, slot, ndim, dims, dtype):
        n = dims[0] * (dims[1] if ndim == 2 else 1)
        requests.append((field, slot))
        return hip.malloc(4 * n)

    alloc_cb = ALLOC(alloc)
    flags = 1 | (2 if is_dual else 0) | 4 | 8 | 16 | 32
    ctx = ctypes.c_void_p()
    assert lib.dmsc_context_create(0, ctypes.byref(ctx)) == 0
    try:
        f = synthetic.gaussian_mixture(96, 130, 12, (3.0, 12.0), 0.05, seed=41).numpy()
        out = Outputs()
        st = lib.dmsc_extract(ctx, hip.to_device(f), 96, 130, ctypes.c_float(0.05), flags, None, alloc_cb, None,
                              ctypes.byref(out))
        assert st == 0, lib.dmsc_last_error(ctx)
        assert tuple(out.shape) == (96, 130)
        assert sorted(s for _, s in requests) == list(range(1, 19)) and all(fl == 0 for fl, _ in requests)
        bad = diff_outputs(read_outputs(hip, out, 96, 130), oracle_extract(f, 0.05, is_dual=is_dual))
        assert not bad, bad

        shapes = [(40, 33), (1, 7), (77, 64), (5, 5)]
        thrs = [0.0, 0.2, 0.05, -1.0]
        fields = [synthetic.gaussian_mixture(h, w, 6, (2.0, 8.0), 0.05, seed=50 + i).numpy()
                  for i, (h, w) in enumerate(shapes)]
        B = len(fields)
        ptrs = (ctypes.c_void_p * B)(*[hip.to_device(x) for x in fields])
        hs = (ctypes.c_int32 * B)(*[h for h, _ in shapes])
        ws = (ctypes.c_int32 * B)(*[w for _, w in shapes])
        th = (ctypes.c_float * B)(*thrs)
        outs = (Outputs * B)()
        requests.clear()
        st = lib.dmsc_extract_batch(ctx, B, ptrs, hs, ws, th, flags, None, alloc_cb, None, outs)
        assert st == 0, lib.dmsc_last_error(ctx)
        # one group: each output slot is one shared buffer holding the fields' arrays back to back
        assert sorted(s for _, s in requests) == list(range(1, 19)) and all(fl == -1 for fl, _ in requests)
        for b in range(B):
            got = read_outputs(hip, outs[b], *shapes[b])
            bad = diff_outputs(got, oracle_extract(fields[b], thrs[b], is_dual=is_dual))
            assert not bad, f"field {b} {shapes[b]}: {bad}"
    finally:
        lib.dmsc_context_destroy(ctx)
        hip.free_all()


def test_ctypes_errors_are_reported():
    hip, lib = Hip(), load_lib()
    alloc_cb = ALLOC(lambda *a: None)  # an allocator that always fails
    ctx = ctypes.c_void_p()
    assert lib.dmsc_context_create(0, ctypes.byref(ctx)) == 0
    try:
        out = Outputs()
        f = np.random.default_rng(3).random((16, 16)).astype(np.float32)
        d = hip.to_device(f)
        assert lib.dmsc_extract(ctx, d, 0, 16, ctypes.c_float(0.1), 1, None, alloc_cb, None, ctypes.byref(out)) == 1
        assert b"non-empty" in lib.dmsc_last_error(ctx)
        assert lib.dmsc_extract(ctx, d, 16, 16, ctypes.c_float(0.1), 1, None, alloc_cb, None, ctypes.byref(out)) == 4
        assert b"allocator" in lib.dmsc_last_error(ctx)
    finally:
        lib.dmsc_context_destroy(ctx)
        hip.free_all()
