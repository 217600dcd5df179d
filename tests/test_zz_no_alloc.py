"""dmsc_persistence_backward_noalloc (no device allocation; fixed per-device scratch) against
dmsc_persistence_backward through the C-ABI: bit-identical gradients on two calls in a row, the
second with a smaller input, and a field too large for the fixed scratch rejected before anything
is queued."""

import ctypes

import numpy as np
import pytest

from tests.conftest import ROOT

pytestmark = pytest.mark.gpu

LIB = ROOT + "/dmsc_amd/lib/libdmsc_hip.so"
DMSC_OK, DMSC_ERR_TOO_LARGE = 0, 2
SCRATCH_FLOATS = 1 << 24


class Pairs(ctypes.Structure):
    _fields_ = [("n_sad", ctypes.c_int64), ("n_min", ctypes.c_int64), ("n_max", ctypes.c_int64),
                ("sad_pts", ctypes.c_void_p), ("p_min", ctypes.c_void_p), ("ppairs_min", ctypes.c_void_p),
                ("min_pts", ctypes.c_void_p), ("p_max", ctypes.c_void_p), ("ppairs_max", ctypes.c_void_p),
                ("max_pts", ctypes.c_void_p)]


def load():
    lib = ctypes.CDLL(LIB)
    args = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
            ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    for name in ("dmsc_persistence_backward", "dmsc_persistence_backward_noalloc"):
        getattr(lib, name).argtypes = args
        getattr(lib, name).restype = ctypes.c_int
    lib.dmsc_context_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    lib.dmsc_context_create.restype = ctypes.c_int
    lib.dmsc_context_destroy.argtypes = [ctypes.c_void_p]
    lib.dmsc_last_error.argtypes = [ctypes.c_void_p]
    lib.dmsc_last_error.restype = ctypes.c_char_p
    return lib


def fields_and_pairs(shapes, seed, is_dual):
    """Device fields (tie-free Gaussian mixtures) and the tensors of their unsimplified pairs."""
    import torch

    from dmsc_amd import synthetic
    from dmsc_amd.csrc import dmsc_gpu

    fields = [synthetic.break_ties(synthetic.gaussian_mixture(h, w, 12, (3.0, 12.0), 0.05, seed=seed + i)).cuda()
              for i, (h, w) in enumerate(shapes)]
    outs = dmsc_gpu.extract_dmsc_list(fields, [-1.0] * len(fields), False, is_dual, False, False, False, False)
    pairs = [[r[k].contiguous() for k in (3, 7, 9, 2, 6, 8, 1)] for r in outs]
    torch.cuda.synchronize()
    return fields, pairs


def call(lib, ctx, entry, fields, pairs, scale, is_dual, grads):
    import torch

    B = len(fields)
    ptrs = (ctypes.c_void_p * B)(*[f.data_ptr() for f in fields])
    hs = (ctypes.c_int32 * B)(*[f.shape[0] for f in fields])
    ws = (ctypes.c_int32 * B)(*[f.shape[1] for f in fields])
    pv = (Pairs * B)()
    for b, p in enumerate(pairs):
        pv[b] = Pairs(p[0].numel(), p[3].numel(), p[6].numel(), *[t.data_ptr() if t.numel() else None for t in p])
    gptrs = (ctypes.c_void_p * B)(*[g.data_ptr() for g in grads])
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    st = getattr(lib, entry)(ctx, B, ptrs, hs, ws, 2 if is_dual else 0, pv, ctypes.c_float(0.1), scale.data_ptr(),
                             gptrs, stream)
    torch.cuda.synchronize()
    return st


@pytest.mark.parametrize("is_dual", [False, True])
def test_noalloc_backward_matches_allocating_backward(is_dual):
    import torch

    lib = load()
    ctx = ctypes.c_void_p()
    assert lib.dmsc_context_create(0, ctypes.byref(ctx)) == DMSC_OK
    try:
        # first call: 20 fields, more than one chunk of the fixed scratch (9 H W floats per field);
        # second call: a smaller input, on the scratch the first call left behind
        for shapes, seed in (([(512, 512)] * 16 + [(700, 650), (333, 517), (1, 1), (2, 9)], 10),
                             ([(64, 80), (31, 17)], 90)):
            print(f"[noalloc] {len(shapes)} fields dual={is_dual}", flush=True)
            fields, pairs = fields_and_pairs(shapes, seed, is_dual)
            scale = torch.linspace(0.5, 2.0, len(fields), device="cuda")
            ref = [torch.empty_like(f) for f in fields]
            got = [torch.full_like(f, float("nan")) for f in fields]
            st = call(lib, ctx, "dmsc_persistence_backward", fields, pairs, scale, is_dual, ref)
            assert st == DMSC_OK, lib.dmsc_last_error(ctx)
            st = call(lib, ctx, "dmsc_persistence_backward_noalloc", fields, pairs, scale, is_dual, got)
            assert st == DMSC_OK, lib.dmsc_last_error(ctx)
            for b in range(len(fields)):
                assert torch.equal(got[b], ref[b]), f"field {b} {tuple(fields[b].shape)}"
            assert sum(int(torch.count_nonzero(g)) for g in got) > 0
    finally:
        lib.dmsc_context_destroy(ctx)


def test_noalloc_backward_rejects_a_field_beyond_its_scratch():
    import torch

    lib = load()
    ctx = ctypes.c_void_p()
    assert lib.dmsc_context_create(0, ctypes.byref(ctx)) == DMSC_OK
    try:
        h, w = 1400, 1400  # 9 h w > 2^24 floats
        assert 9 * h * w > SCRATCH_FLOATS
        f = torch.zeros((h, w), device="cuda")
        empty_i = torch.empty(0, dtype=torch.int32, device="cuda")
        empty_f = torch.empty(0, dtype=torch.float32, device="cuda")
        pairs = [[empty_i, empty_f, empty_i, empty_i, empty_f, empty_i, empty_i]]
        g = torch.full_like(f, 7.0)
        st = call(lib, ctx, "dmsc_persistence_backward_noalloc", [f], pairs, torch.ones(1, device="cuda"), False, [g])
        assert st == DMSC_ERR_TOO_LARGE
        assert b"scratch" in lib.dmsc_last_error(ctx)
        assert bool((g == 7.0).all())  # nothing was queued
        # the allocating entry takes the same field
        st = call(lib, ctx, "dmsc_persistence_backward", [f], pairs, torch.ones(1, device="cuda"), False, [g])
        assert st == DMSC_OK and int(torch.count_nonzero(g)) == 0
    finally:
        lib.dmsc_context_destroy(ctx)
