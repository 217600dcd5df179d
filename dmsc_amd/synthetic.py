"""Synthetic scalar fields.

``generate_noisy_landscape`` keeps the reference helper's signature and formula (reference
dmsc/synthetic.py:4-31) for API compatibility. The workload generators below produce the fields the
benchmarks and parity tests use; all are built on the CPU from a seeded ``torch.Generator`` so the
same float32 bytes feed the GPU path and the oracle.
"""

import torch


def generate_noisy_landscape(H=20, W=20, with_loop=False, z_scale=1.0, noise_scale=0.1):
    """Four Gaussian bumps (two peaks, two valleys), optional loop feature, plus uniform noise
    drawn from torch's global generator (reference dmsc/synthetic.py:4-31)."""
    y, x = torch.meshgrid(torch.linspace(-2, 2, H), torch.linspace(-2, 2, W), indexing="ij")
    z = torch.exp(-((x - 1) ** 2 + y**2) / 0.5) + torch.exp(-((x + 1) ** 2 + y**2) / 0.5)
    z = z - 0.5 * torch.exp(-(x**2 + (y - 1) ** 2) / 0.5) - 1.0 * torch.exp(-(x**2 + (y + 1) ** 2) / 0.5)
    if with_loop:
        # added as one term, (peak + slope), in the reference's order: float addition does not associate
        peak3 = torch.exp(-((x**2 + (y + 1) ** 2) ** 3) / 0.5)
        slope = -y / 4
        z = z + (peak3 + slope)
    noise = torch.rand(H, W) - 0.5
    return (z * z_scale + noise * noise_scale).to(torch.float32)


def gaussian_mixture(H, W, n_bumps, sigma_range, noise, seed):
    """Sum of ``n_bumps`` isotropic Gaussians (centres uniform over the grid, sigma uniform in
    ``sigma_range`` pixels, amplitude uniform in [-1, 1]) plus i.i.d. uniform noise in
    [-noise, noise], min-max normalised to [0, 1]. Separable bumps, summed as one matrix product."""
    g = torch.Generator().manual_seed(int(seed))
    cy = torch.rand(n_bumps, generator=g, dtype=torch.float64) * H
    cx = torch.rand(n_bumps, generator=g, dtype=torch.float64) * W
    lo, hi = sigma_range
    sig = lo + (hi - lo) * torch.rand(n_bumps, generator=g, dtype=torch.float64)
    amp = 2.0 * torch.rand(n_bumps, generator=g, dtype=torch.float64) - 1.0
    ys = torch.arange(H, dtype=torch.float64)
    xs = torch.arange(W, dtype=torch.float64)
    gy = torch.exp(-((ys[None, :] - cy[:, None]) ** 2) / (2 * sig[:, None] ** 2))  # (K, H)
    gx = torch.exp(-((xs[None, :] - cx[:, None]) ** 2) / (2 * sig[:, None] ** 2))  # (K, W)
    z = (gy * amp[:, None]).T @ gx
    z = z + (2.0 * torch.rand(H, W, generator=g, dtype=torch.float64) - 1.0) * noise
    z = (z - z.min()) / (z.max() - z.min())
    return z.to(torch.float32)


def quantized_noise(H, W, levels, seed):
    """i.i.d. uniform [0, 1) noise quantised to ``levels`` values (heavy ties)."""
    g = torch.Generator().manual_seed(int(seed))
    u = torch.rand(H, W, generator=g, dtype=torch.float64)
    return (torch.floor(u * levels) / levels).to(torch.float32)


def mixed_batch(n_fields, seed, sizes=tuple(range(64, 1025, 32)), bumps=(8, 64), thr_range=(0.02, 0.2), noise=0.05):
    """Fields of independent sizes: H and W uniform over ``sizes``, 8-64 bumps, sigma uniform in
    [min(H,W)/32, min(H,W)/8], noise amplitude ``noise``; per-field threshold uniform in thr_range."""
    g = torch.Generator().manual_seed(int(seed))
    fields, thrs = [], []
    for _ in range(n_fields):
        H = sizes[int(torch.randint(len(sizes), (1,), generator=g))]
        W = sizes[int(torch.randint(len(sizes), (1,), generator=g))]
        k = int(torch.randint(bumps[0], bumps[1] + 1, (1,), generator=g))
        thr = thr_range[0] + (thr_range[1] - thr_range[0]) * float(torch.rand(1, generator=g))
        m = min(H, W)
        sub_seed = int(torch.randint(2**31 - 1, (1,), generator=g))
        fields.append(gaussian_mixture(H, W, k, (m / 32.0, m / 8.0), noise, sub_seed))
        thrs.append(thr)
    return fields, thrs


def break_ties(field):
    """A tie-free float32 field with the same (value, pixel id) order: visiting pixels in that order,
    every value is raised to at least the next float32 above its predecessor. Values move by a few
    ulps at most where they were tied, and the order the complex is built from is unchanged, so the
    discrete gradient and every pairing equal those of the input."""
    import numpy as np

    f = np.ascontiguousarray(torch.as_tensor(field, dtype=torch.float32).numpy()).reshape(-1)
    order = np.lexsort((np.arange(f.size), f))
    v = f[order].copy()
    bump = np.flatnonzero(v[1:] <= v[:-1]) + 1
    while bump.size:  # runs of equal values need one pass per member
        v[bump] = np.nextafter(v[bump - 1], np.float32(np.inf))
        bump = np.flatnonzero(v[1:] <= v[:-1]) + 1
    out = np.empty_like(f)
    out[order] = v
    return torch.from_numpy(out.reshape(tuple(torch.as_tensor(field).shape)))


def loss_batch(n_fields=64, H=512, W=512, seed=5000):
    """Fields of the persistence-loss workload: Gaussian mixtures (24 bumps, sigma 8-32 px, noise
    0.05, seeds seed..seed+n-1) made tie-free by break_ties."""
    return torch.stack([break_ties(gaussian_mixture(H, W, 24, (8.0, 32.0), 0.05, seed=seed + b))
                        for b in range(n_fields)])
