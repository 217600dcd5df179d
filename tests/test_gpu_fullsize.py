"""Parity at the benchmark sizes: the HIP path against the oracle, all 19 outputs exact, on the exact
inputs the benchmark times (same generators and seeds) -- the 256^2 Gaussian mixture at thr 0.1, the
4096^2 Gaussian mixture primal and dual at every threshold of {0, 0.05, 0.2}, and the 8192^2
quantised noise at thr 0.05 -- plus size-independent invariants of the outputs (Euler characteristic
of the simplified complex, CSR offsets, survivors' persistence above the threshold, edge rows in
range). The oracle runs on all host cores here; its output is independent of the thread count."""

import functools
import os

import numpy as np
import pytest

from dmsc_amd import synthetic
from tests.parity_util import diff_outputs, gpu_extract, oracle_extract

pytestmark = pytest.mark.gpu


def check_invariants(out, thr):
    max_pts, min_pts, sad_pts = out[1], out[2], out[3]
    # every cancellation removes one saddle and one extremum: chi stays 1
    assert max_pts.size - sad_pts.size + min_pts.size == 1
    ns = sad_pts.size
    for p in (out[6], out[7]):
        assert p.shape == (ns,)
        if thr >= 0:
            assert np.all((p == -1.0) | (p > thr))
    for cells, offs, arc_offs in ((out[12], out[13], out[14]), (out[16], out[17], out[18])):
        assert offs.shape == (ns + 1,) and offs[0] == 0 and offs[-1] == cells.size
        assert np.all(np.diff(offs) >= 0)
        assert np.all((arc_offs >= offs[:-1]) & (arc_offs <= offs[1:]))
    for e, n_ext in ((out[4], max_pts.size), (out[5], min_pts.size)):
        assert e.ndim == 2 and e.shape[1] == 2
        assert np.all((e[:, 0] >= 0) & (e[:, 0] < ns) & (e[:, 1] >= 0) & (e[:, 1] < n_ext))


def run_case(field, thr, is_dual):
    print(f"[fullsize] {field.shape} thr={thr} dual={is_dual}", flush=True)
    got = gpu_extract(field, thr, is_dual=is_dual)
    check_invariants(got, thr)
    exp = oracle_extract(field, thr, is_dual=is_dual, nthreads=min(os.cpu_count() or 1, 16))
    bad = diff_outputs(got, exp)
    assert not bad, f"{field.shape} thr={thr} dual={is_dual}: {bad}"
    return got


def test_gaussian_256_thr_0_1():
    f = synthetic.gaussian_mixture(256, 256, 24, (8.0, 32.0), 0.05, seed=256).numpy()
    run_case(f, 0.1, False)


@functools.lru_cache(maxsize=1)
def gauss4096():
    return synthetic.gaussian_mixture(4096, 4096, 256, (16.0, 256.0), 0.02, seed=4096).numpy()


@pytest.mark.parametrize("is_dual", [False, True])
@pytest.mark.parametrize("thr", [0.0, 0.05, 0.2])
def test_gaussian_4096(is_dual, thr):
    run_case(gauss4096(), thr, is_dual)


def test_noise_8192():
    f = synthetic.quantized_noise(8192, 8192, 256, seed=8192).numpy()
    run_case(f, 0.05, False)
