"""The CPU reference gradient of the persistence loss (assembled from the oracle's pairs) against
central finite differences of the oracle's loss, on tie-free fields: pins the reference that the
GPU backward is compared with."""

import numpy as np
import pytest

from tests.diff_util import loss_and_grad, pairing_signature, tie_free_field


@pytest.mark.parametrize("is_dual", [False, True])
def test_reference_gradient_matches_finite_differences(is_dual):
    eps, max_pers = 1e-3, 0.1
    f = tie_free_field(24, 21, 4)
    _, g = loss_and_grad(f, max_pers, is_dual)
    sig = pairing_signature(f, max_pers, is_dual)
    rng = np.random.default_rng(1)
    nz = np.flatnonzero(g)
    checked = 0
    for p in np.concatenate([rng.choice(nz, size=min(16, nz.size), replace=False), rng.choice(f.size, 8)]):
        fp, fm = f.copy().reshape(-1), f.copy().reshape(-1)
        fp[p] += np.float32(eps)
        fm[p] -= np.float32(eps)
        fp, fm = fp.reshape(f.shape), fm.reshape(f.shape)
        if pairing_signature(fp, max_pers, is_dual) != sig or pairing_signature(fm, max_pers, is_dual) != sig:
            continue
        fd = (loss_and_grad(fp, max_pers, is_dual)[0] - loss_and_grad(fm, max_pers, is_dual)[0]) / (
            float(fp.reshape(-1)[p]) - float(fm.reshape(-1)[p]))
        assert abs(fd - g.reshape(-1)[p]) <= 2e-4, (p, fd, g.reshape(-1)[p])
        checked += 1
    assert checked >= 12


def test_dual_of_negated_field_mirrors_primal():
    # on tie-free fields the dual complex of -f is the primal complex of f with the roles of
    # minima and maxima exchanged: same loss, gradient negated
    f = tie_free_field(24, 21, 4)
    loss, g = loss_and_grad(f, 0.1, False)
    loss_d, gd = loss_and_grad(-f, 0.1, True)
    assert loss > 0 and abs(loss - loss_d) <= 1e-12 * max(1.0, loss)
    np.testing.assert_allclose(g, -gd, rtol=0, atol=1e-12)


def test_planted_sign_error_is_caught():
    # the finite-difference check rejects a gradient with its sign flipped (a plausible mistake)
    eps, max_pers = 1e-3, 0.1
    f = tie_free_field(24, 21, 4)
    _, g = loss_and_grad(f, max_pers, False)
    bad = -g
    sig = pairing_signature(f, max_pers, False)
    mismatches = 0
    for p in np.flatnonzero(g)[:16]:
        fp, fm = f.copy().reshape(-1), f.copy().reshape(-1)
        fp[p] += np.float32(eps)
        fm[p] -= np.float32(eps)
        fp, fm = fp.reshape(f.shape), fm.reshape(f.shape)
        if pairing_signature(fp, max_pers, False) != sig or pairing_signature(fm, max_pers, False) != sig:
            continue
        fd = (loss_and_grad(fp, max_pers, False)[0] - loss_and_grad(fm, max_pers, False)[0]) / (
            float(fp.reshape(-1)[p]) - float(fm.reshape(-1)[p]))
        if abs(fd - bad.reshape(-1)[p]) > 2e-4:
            mismatches += 1
    assert mismatches > 0
