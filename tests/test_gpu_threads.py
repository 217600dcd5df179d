"""Two Python threads extracting on one GPU share its context (one per device, as the binding keeps
it); the context lock makes them take turns, so each thread's results equal a sequential run."""

import threading

import pytest
import torch

from dmsc_amd import compute_dmsc, compute_dmsc_batch
from dmsc_amd.synthetic import gaussian_mixture
from tests.test_api import assert_complex_equal


@pytest.mark.gpu
def test_two_threads_one_device_match_sequential():
    fields = [gaussian_mixture(300 + 37 * i, 260 - 21 * i, 24, (6, 40), 0.05, seed=900 + i).cuda() for i in range(4)]
    thr = [0.02, 0.1, 0.05, 0.2]
    expected = [compute_dmsc(f, t) for f, t in zip(fields, thr)]
    expected_batch = compute_dmsc_batch(fields, thr)
    torch.cuda.synchronize()

    results = {}
    errors = []

    def worker(k):
        try:
            stream = torch.cuda.Stream()  # each thread on its own stream
            with torch.cuda.stream(stream):
                out = []
                for rep in range(3):
                    for i in range(k, len(fields), 2):
                        out.append((i, compute_dmsc(fields[i], thr[i])))
                    out.append((-1, compute_dmsc_batch(fields, thr)))
                stream.synchronize()
            results[k] = out
        except Exception as e:  # reported below, on the main thread
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads), "an extraction thread did not finish"
    assert not errors, errors
    for k in range(2):
        for i, ms in results[k]:
            if i < 0:
                for a, e in zip(ms, expected_batch, strict=True):
                    assert_complex_equal(a, e)
            else:
                assert_complex_equal(ms, expected[i])
