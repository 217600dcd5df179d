"""The gradient kernel's lower-star formulation (per-vertex comparisons + compile-time queue table,
host build of dmsc_common.hpp) against the oracle's literal FIFO-queue restatement
(reference cpu/gradient.hxx:19-196). Exhaustive on 3x3 ternary fields (every order pattern of one
interior vertex, with ties), plus random small fields with heavy ties, +-0, infinities and NaN."""

import itertools

import numpy as np
import pytest

import oracle


def lut_grad(lib, f, dual):
    f = np.ascontiguousarray(f, dtype=np.float32)
    H, W = f.shape
    pw = np.empty(4 * (H + 1) * (W + 1), dtype=np.int32)
    lib.lut_gradient(f.ctypes.data, H, W, int(dual), pw.ctypes.data)
    return pw


@pytest.mark.parametrize("dual", [False, True])
def test_exhaustive_ternary_3x3(lut_lib, dual):
    vals = np.array(list(itertools.product([0.0, 1.0, 2.0], repeat=9)), dtype=np.float32)
    for v in vals:
        f = v.reshape(3, 3)
        assert np.array_equal(lut_grad(lut_lib, f, dual), oracle.gradient(f, dual)), f


@pytest.mark.parametrize("dual", [False, True])
def test_all_orders_of_a_3x3_window(lut_lib, dual):
    # every strict order of the 9 values (sampled: 20000 of the 362880 permutations) + border sizes
    rng = np.random.default_rng(7)
    for _ in range(20000):
        f = rng.permutation(9).astype(np.float32).reshape(3, 3)
        assert np.array_equal(lut_grad(lut_lib, f, dual), oracle.gradient(f, dual))


@pytest.mark.parametrize("dual", [False, True])
@pytest.mark.parametrize("shape", [(1, 1), (1, 7), (7, 1), (2, 2), (4, 5), (6, 6), (9, 4)])
def test_random_small_fields(lut_lib, dual, shape):
    rng = np.random.default_rng(hash(shape) % 1000)
    specials = np.array([0.0, -0.0, 1.0, np.inf, -np.inf], dtype=np.float32)
    for trial in range(300):
        kind = trial % 4
        if kind == 0:
            f = rng.integers(0, 3, size=shape).astype(np.float32)
        elif kind == 1:
            f = rng.random(shape).astype(np.float32)
        elif kind == 2:
            f = specials[rng.integers(0, len(specials), size=shape)]
        else:
            f = rng.integers(0, 2, size=shape).astype(np.float32)
            f[rng.random(shape) < 0.3] = np.nan
        assert np.array_equal(lut_grad(lut_lib, f, dual), oracle.gradient(f, dual)), f
