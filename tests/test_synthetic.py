"""generate_noisy_landscape against the reference's own fixture field.

The reference made tests/data/landscape_32x32.pt with torch.manual_seed(1780627675) and
generate_noisy_landscape(32, 32, with_loop=True) when the helper's z_scale default was 2.0
(reference tests/generate_ground_truth.py:9-16, dmsc/synthetic.py:11-13). Rebuilding it pins the
formula, its coefficients and the noise draw. The values agree to 2.4e-7 absolute: the remaining
differences are last-bit differences of exp between the machine that wrote the fixture and this one.
"""

import numpy as np
import torch

from dmsc_amd.synthetic import generate_noisy_landscape
from tests.conftest import GOLDEN

import os

FIXTURE = np.load(os.path.join(GOLDEN, "landscape_32x32.npz"))["field"]


def test_landscape_rebuilds_the_reference_fixture():
    torch.manual_seed(1780627675)
    z = generate_noisy_landscape(32, 32, with_loop=True, z_scale=2.0).numpy()
    assert z.dtype == np.float32 and z.shape == FIXTURE.shape
    np.testing.assert_allclose(z, FIXTURE, rtol=0, atol=4.8e-7)
    # a wrong coefficient, noise scale or draw order moves values by far more than that
    torch.manual_seed(1780627675)
    z1 = generate_noisy_landscape(32, 32, with_loop=True, z_scale=1.0).numpy()
    assert np.abs(z1 - FIXTURE).max() > 0.1


def test_loop_feature_is_added_as_one_term():
    """reference dmsc/synthetic.py:23-26: z += peak3 + slope, i.e. z + (peak3 + slope); float
    addition does not associate, so (z + peak3) + slope differs in the last bit at many pixels."""
    base = generate_noisy_landscape(48, 40, with_loop=False, noise_scale=0.0)
    loop = generate_noisy_landscape(48, 40, with_loop=True, noise_scale=0.0)
    y, x = torch.meshgrid(torch.linspace(-2, 2, 48), torch.linspace(-2, 2, 40), indexing="ij")
    peak3 = torch.exp(-((x**2 + (y + 1) ** 2) ** 3) / 0.5)
    slope = -y / 4
    assert torch.equal(loop, base + (peak3 + slope))
    assert not torch.equal(loop, (base + peak3) + slope)


def test_noise_comes_from_the_global_generator():
    torch.manual_seed(3)
    a = generate_noisy_landscape(16, 16)
    torch.manual_seed(3)
    b = generate_noisy_landscape(16, 16)
    assert torch.equal(a, b)
    assert not torch.equal(a, generate_noisy_landscape(16, 16))
