"""Parity on the two batch workloads at their full benchmark size.

* The 512-field mixed batch (H, W uniform over {64, 96, ..., 1024}, 8-64 bumps, per-field threshold
  in [0.02, 0.2], seed 512; the bench's default workload): one batched call, every field's 19
  outputs exactly equal to the oracle's and to the per-field plain path (the batched == sequential
  criterion of the reference's tests/test_api.py:71-80).
* The persistence loss over 64 tie-free 512x512 fields (seeds 5000-5063; the bench's loss
  workload): loss and gradient of every field against the gradient assembled from the oracle's
  pairs, and central finite differences (eps 1e-3) at pixels whose pairing is stable within +-eps
  on two of the fields.
* A field mixing -0.0 and +0.0: the persistence subtraction can yield either sign of zero and the
  float outputs are compared bit for bit.
"""

import numpy as np
import pytest

from dmsc_amd import synthetic
from tests.diff_util import loss_and_grad, pairing_signature
from tests.parity_util import diff_outputs, gpu_extract, oracle_extract, oracle_many, pool_workers

pytestmark = pytest.mark.gpu

MAX_PERS = 0.1


def test_mixed_batch_512_fields():
    import torch

    import dmsc_amd

    fields, thrs = synthetic.mixed_batch(512, seed=512)
    print(f"[workload] mixed batch: {len(fields)} fields, {sum(f.numel() for f in fields)} pixels", flush=True)
    dev = [f.cuda() for f in fields]
    batched = dmsc_amd.compute_dmsc_batch(dev, thrs, return_gradient=True)
    torch.cuda.synchronize()
    cases = [(f.numpy(), t) for f, t in zip(fields, thrs)]
    for j, exp in enumerate(oracle_many(cases, return_gradient=True)):
        got = [t.cpu().numpy() for t in batched[j]]
        bad = diff_outputs(got, exp)
        assert not bad, f"field {j} shape={tuple(fields[j].shape)} thr={thrs[j]}: {bad}"
    # the plain per-field path gives the same bytes (checked on a spread of the fields)
    for j in range(0, 512, 37):
        plain = dmsc_amd.compute_dmsc(dev[j], thrs[j], return_gradient=True)
        for a, b in zip(plain, batched[j]):
            assert torch.equal(a.cpu(), b.cpu()), f"field {j}"


def gpu_loss_and_grad(stack, is_dual):
    import torch

    import dmsc_amd

    x = stack.cuda().requires_grad_(True)
    losses = dmsc_amd.persistence_loss(x, MAX_PERS, is_dual)
    losses.sum().backward()
    torch.cuda.synchronize()
    return losses.detach().cpu().numpy().astype(np.float64), x.grad.cpu().numpy()


@pytest.mark.parametrize("is_dual", [False, True])
def test_loss_batch_64_fields_against_cpu_pairs(is_dual):
    from concurrent.futures import ThreadPoolExecutor

    stack = synthetic.loss_batch(64)
    fields = stack.numpy()
    assert all(np.unique(f).size == f.size for f in fields), "the loss workload must be tie-free"
    print(f"[workload] loss batch 64 x 512^2 dual={is_dual}", flush=True)
    losses, grads = gpu_loss_and_grad(stack, is_dual)
    with ThreadPoolExecutor(pool_workers()) as ex:
        refs = list(ex.map(lambda f: loss_and_grad(f, MAX_PERS, is_dual), fields))
    for j, (loss, g) in enumerate(refs):
        assert np.count_nonzero(g) > 100
        np.testing.assert_allclose(losses[j], loss, rtol=1e-5, atol=1e-7, err_msg=f"field {j}")
        np.testing.assert_allclose(grads[j], g, rtol=1e-5, atol=1e-6, err_msg=f"field {j}")


def test_loss_batch_finite_differences():
    from concurrent.futures import ThreadPoolExecutor

    eps = np.float32(1e-3)
    stack = synthetic.loss_batch(64)
    _, grads = gpu_loss_and_grad(stack, False)
    rng = np.random.default_rng(64)
    jobs = []
    for j in (0, 37):
        f, g = stack[j].numpy(), grads[j].reshape(-1)
        nz = np.flatnonzero(g)
        for p in np.concatenate([rng.choice(nz, 40, replace=False), rng.choice(f.size, 8, replace=False)]):
            jobs.append((j, int(p)))
    sig = {j: pairing_signature(stack[j].numpy(), MAX_PERS, False) for j in (0, 37)}

    def fd(job):
        j, p = job
        f = stack[j].numpy()
        fp, fm = f.copy().reshape(-1), f.copy().reshape(-1)
        fp[p] += eps
        fm[p] -= eps
        fp, fm = fp.reshape(f.shape), fm.reshape(f.shape)
        if pairing_signature(fp, MAX_PERS, False) != sig[j] or pairing_signature(fm, MAX_PERS, False) != sig[j]:
            return None  # the pairing changes within +-eps: no derivative to compare
        lp, _ = loss_and_grad(fp, MAX_PERS, False)
        lm, _ = loss_and_grad(fm, MAX_PERS, False)
        return (lp - lm) / (float(fp.reshape(-1)[p]) - float(fm.reshape(-1)[p]))

    with ThreadPoolExecutor(pool_workers()) as ex:
        results = list(ex.map(fd, jobs))
    checked = 0
    for (j, p), d in zip(jobs, results):
        if d is None:
            continue
        assert abs(d - grads[j].reshape(-1)[p]) <= 2e-4, f"field {j} pixel {p}: fd {d} vs {grads[j].reshape(-1)[p]}"
        checked += 1
    assert checked >= 30, f"only {checked} pixels had a stable pairing"


@pytest.mark.parametrize("is_dual", [False, True])
@pytest.mark.parametrize("thr", [-1.0, 0.0, 0.3])
def test_signed_zeros(thr, is_dual):
    rng = np.random.default_rng(77)
    f = (rng.integers(0, 4, size=(160, 150)) / 4).astype(np.float32)
    zero = f == 0
    f[zero & (rng.random(f.shape) < 0.5)] = np.float32(-0.0)
    assert np.signbit(f[zero]).any() and (~np.signbit(f[zero])).any()
    print(f"[workload] signed zeros thr={thr} dual={is_dual}", flush=True)
    got = gpu_extract(f, thr, is_dual=is_dual)
    exp = oracle_extract(f, thr, is_dual=is_dual)
    bad = diff_outputs(got, exp)
    assert not bad, bad
    # the cases above do produce persistence values of both zero signs
    p = np.concatenate([exp[6], exp[7]])
    if thr < 0:
        assert np.signbit(p[p == 0]).any() and (~np.signbit(p[p == 0])).any()
