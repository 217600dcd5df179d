"""HIP path vs the oracle, bit-exact on all 19 outputs (integer outputs exactly, float32 persistence
values bit for bit), on the reference fixture input and on seeded fields covering tile borders,
degenerate shapes, heavy ties and non-finite values."""

import itertools

import numpy as np
import pytest

from dmsc_amd import synthetic
from tests.conftest import GOLDEN
from tests.parity_util import diff_outputs, gpu_extract, oracle_extract

pytestmark = pytest.mark.gpu


def landscape():
    return np.load(f"{GOLDEN}/landscape_32x32.npz")["field"]


def check(field, thr, **kw):
    print(f"[parity] shape={field.shape} thr={thr} {kw}", flush=True)
    got = gpu_extract(field, thr, **kw)
    exp = oracle_extract(field, thr, **kw)
    bad = diff_outputs(got, exp)
    assert not bad, f"shape={field.shape} thr={thr} {kw}: {bad}"


FLAG_SETS = list(itertools.product([True, False], repeat=4))


@pytest.mark.parametrize("is_dual", [False, True])
@pytest.mark.parametrize("thr", [-1.0, 0.15])
@pytest.mark.parametrize("flags", FLAG_SETS)
def test_fixture_landscape_all_flags(is_dual, thr, flags):
    tmax, tmin, gmax, gmin = flags
    check(landscape(), thr, return_gradient=thr < 0, is_dual=is_dual, tmax=tmax, tmin=tmin, gmax=gmax, gmin=gmin)


SHAPES = [(1, 1), (1, 2), (2, 1), (1, 9), (9, 1), (2, 2), (3, 5), (7, 11), (16, 64), (17, 65), (31, 63), (64, 129),
          (100, 37)]


@pytest.mark.parametrize("is_dual", [False, True])
@pytest.mark.parametrize("shape", SHAPES)
def test_random_shapes(is_dual, shape):
    rng = np.random.default_rng(abs(hash(shape)) % 10000)
    f = rng.random(shape).astype(np.float32)
    for thr in (-1.0, 0.0, 0.1, 0.5):
        check(f, thr, is_dual=is_dual)


@pytest.mark.parametrize("is_dual", [False, True])
@pytest.mark.parametrize("levels", [2, 3, 16, 256])
def test_heavy_ties(is_dual, levels):
    f = synthetic.quantized_noise(70, 90, levels, seed=levels).numpy()
    for thr in (-1.0, 0.0, 0.05, 0.3):
        check(f, thr, is_dual=is_dual)


@pytest.mark.parametrize("is_dual", [False, True])
@pytest.mark.parametrize("value", [0.0, float("inf"), float("-inf"), float("nan")])
def test_constant_fields(is_dual, value):
    check(np.full((6, 5), value, dtype=np.float32), 0.0, is_dual=is_dual)


@pytest.mark.parametrize("is_dual", [False, True])
def test_gaussian_mixture_256(is_dual):
    f = synthetic.gaussian_mixture(256, 256, 24, (8, 32), 0.05, seed=1).numpy()
    for thr in (0.0, 0.1, 0.3):
        check(f, thr, is_dual=is_dual)


def test_infinities_mixed():
    rng = np.random.default_rng(3)
    f = rng.random((40, 50)).astype(np.float32)
    f[rng.random(f.shape) < 0.05] = np.inf
    f[rng.random(f.shape) < 0.05] = -np.inf
    for thr in (-1.0, 0.2):
        check(f, thr)
