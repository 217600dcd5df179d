"""Persistence-loss gradient: the hand-written HIP backward against the gradient assembled on the CPU
from the oracle's pairs, and against central finite differences (eps = 1e-3) of the oracle's loss
on tie-free fields at pixels where the pairing does not change within +-eps."""

import numpy as np
import pytest

from tests.diff_util import loss_and_grad, pairing_signature, tie_free_field

pytestmark = pytest.mark.gpu

MAX_PERS = 0.1


def gpu_loss_grad(fields, is_dual, weights=None):
    import torch

    import dmsc_amd

    x = torch.from_numpy(np.stack(fields)).cuda().requires_grad_(True)
    losses = dmsc_amd.persistence_loss(x, MAX_PERS, is_dual)
    w = torch.ones(len(fields), device="cuda") if weights is None else torch.tensor(weights, device="cuda")
    (losses * w).sum().backward()
    torch.cuda.synchronize()
    return losses.detach().cpu().numpy().astype(np.float64), x.grad.cpu().numpy()


@pytest.mark.parametrize("is_dual", [False, True])
def test_gradient_matches_cpu_pairs(is_dual):
    fields = [tie_free_field(40, 33, s) for s in (1, 2, 3)]
    weights = [1.0, -0.5, 2.0]
    losses, grads = gpu_loss_grad(fields, is_dual, weights)
    for j, f in enumerate(fields):
        loss, g = loss_and_grad(f, MAX_PERS, is_dual)
        print(f"[diff] field {j} {f.shape} dual={is_dual} loss={loss:.6f} nonzero={np.count_nonzero(g)}", flush=True)
        assert np.count_nonzero(g) > 0
        np.testing.assert_allclose(losses[j], loss, rtol=1e-5, atol=1e-7)
        np.testing.assert_allclose(grads[j], weights[j] * g, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("is_dual", [False, True])
def test_gradient_matches_finite_differences(is_dual):
    eps = 1e-3
    f = tie_free_field(36, 30, 11)
    _, grads = gpu_loss_grad([f], is_dual)
    g = grads[0]
    rng = np.random.default_rng(5)
    nz = np.flatnonzero(g)
    cand = np.concatenate([rng.choice(nz, size=min(64, nz.size), replace=False), rng.choice(f.size, 12, replace=False)])
    sig = pairing_signature(f, MAX_PERS, is_dual)
    checked = 0
    for p in cand:
        fp, fm = f.copy().reshape(-1), f.copy().reshape(-1)
        fp[p] += np.float32(eps)
        fm[p] -= np.float32(eps)
        fp, fm = fp.reshape(f.shape), fm.reshape(f.shape)
        if pairing_signature(fp, MAX_PERS, is_dual) != sig or pairing_signature(fm, MAX_PERS, is_dual) != sig:
            continue  # the pairing changes within +-eps: no derivative to compare
        lp, _ = loss_and_grad(fp, MAX_PERS, is_dual)
        lm, _ = loss_and_grad(fm, MAX_PERS, is_dual)
        fd = (lp - lm) / (float(fp.reshape(-1)[p]) - float(fm.reshape(-1)[p]))
        assert abs(fd - g.reshape(-1)[p]) <= 2e-4, f"pixel {p}: fd {fd} vs backward {g.reshape(-1)[p]}"
        checked += 1
    assert checked >= 20, f"only {checked} pixels had a stable pairing"


@pytest.mark.parametrize("shape", [(1, 1), (1, 7), (9, 1), (25, 61)])
def test_single_fields_of_other_shapes(shape):
    import torch

    import dmsc_amd

    f = tie_free_field(*shape, seed=sum(shape)) if shape[0] * shape[1] > 1 else np.zeros(shape, np.float32)
    x = torch.from_numpy(f).cuda().requires_grad_(True)
    loss = dmsc_amd.persistence_loss(x, MAX_PERS)
    loss.backward()
    ref_loss, ref_g = loss_and_grad(f, MAX_PERS, False)
    np.testing.assert_allclose(float(loss), ref_loss, rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(x.grad.cpu().numpy(), ref_g, rtol=1e-5, atol=1e-6)


def test_single_field_and_batch_agree():
    import torch

    import dmsc_amd

    f = tie_free_field(31, 27, 8)
    x2 = torch.from_numpy(f).cuda().requires_grad_(True)
    l2 = dmsc_amd.persistence_loss(x2, MAX_PERS)
    l2.backward()
    losses, grads = gpu_loss_grad([f, tie_free_field(31, 27, 9)], False)
    assert float(l2) == float(losses[0])
    assert np.array_equal(x2.grad.cpu().numpy(), grads[0])


def test_backward_is_deterministic():
    from dmsc_amd import synthetic

    fields = [synthetic.gaussian_mixture(64, 64, 8, (2.0, 8.0), 0.05, seed=s).numpy() for s in (21, 22)]
    _, g1 = gpu_loss_grad(fields, False)
    _, g2 = gpu_loss_grad(fields, False)
    assert np.array_equal(g1, g2)
