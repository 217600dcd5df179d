"""Differentiable persistence loss on the GPU (no reference counterpart: the reference lists
differentiation as future work, readme.md:237-242).

    loss(field) = sum over persistence pairs with 0 <= pers < max_persistence of pers**2

over both the min-side and the max-side pairs of the unsimplified complex (threshold < 0, so every
pair is present; cycles, persistence -1, are excluded). The forward runs the batched extraction;
the backward is a hand-written HIP kernel (dmsc_persistence_backward) that puts +-2*pers*grad on
each pair's two value pixels, deterministically.
"""

import torch

from .dmsc import _validate

try:
    from .csrc import dmsc_gpu
except ImportError as e:  # fail loudly: there is no CPU fallback
    raise ImportError(f"dmsc_amd: the HIP extension is not built: {e}") from e

# output slots of one complex: sad_pts, p_min, ppairs_min, min_pts, p_max, ppairs_max, max_pts
_PAIR_SLOTS = (3, 7, 9, 2, 6, 8, 1)


class PersistenceLoss(torch.autograd.Function):
    """Per-field loss of a [B, H, W] float32 GPU tensor; returns a [B] tensor."""

    @staticmethod
    def forward(ctx, fields, max_persistence, is_dual):
        fields = fields.detach().contiguous()
        out = dmsc_gpu.extract_dmsc(fields, -1.0, False, bool(is_dual), False, False, False, False)
        pairs = [[r[k] for k in _PAIR_SLOTS] for r in out]
        losses = dmsc_gpu.persistence_loss(pairs, float(max_persistence))  # one kernel for every field
        ctx.save_for_backward(fields)
        ctx.pairs = pairs
        ctx.max_persistence = float(max_persistence)
        ctx.is_dual = bool(is_dual)
        return losses

    @staticmethod
    def backward(ctx, grad_losses):
        (fields,) = ctx.saved_tensors
        g = grad_losses.detach().to(device=fields.device, dtype=torch.float32).reshape(-1).contiguous()
        grads = dmsc_gpu.persistence_backward(list(fields.unbind(0)), ctx.pairs, g, ctx.max_persistence, ctx.is_dual)
        return torch.stack(grads), None, None


def persistence_loss(img, max_persistence=0.1, is_dual=False):
    """sum(pers**2 for pairs with 0 <= pers < max_persistence) of a [H, W] field (scalar) or of each
    field of a [B, H, W] batch ([B] tensor), differentiable with respect to the field values.

    The gradient is exact where the pairing is locally constant (tie-free values); at ties it is the
    gradient of the pairing chosen by the (value, cell id) order."""
    _validate(img)
    if img.device.type != "cuda":
        raise RuntimeError("dmsc_amd computes on AMD Instinct GPUs only; move the field with .to('cuda').")
    if img.dim() == 2:
        return PersistenceLoss.apply(img.unsqueeze(0), max_persistence, is_dual)[0]
    return PersistenceLoss.apply(img, max_persistence, is_dual)
