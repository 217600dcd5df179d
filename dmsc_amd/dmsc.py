"""Python API of the MI355X compute_dmsc path.

Mirrors the reference ``dmsc.compute_dmsc`` / ``dmsc.MSComplex`` (reference dmsc/dmsc.py:9-276):
same argument names, defaults, flag mapping, validation messages and the 19 result fields in the
same order, so code written against the reference runs unchanged. Computation happens in the
gfx950 kernels behind ``dmsc_amd.csrc.dmsc_gpu``; there is no CPU fallback -- a CPU tensor raises.
"""

import time
from dataclasses import dataclass, fields

import torch

try:
    from .csrc import dmsc_gpu
except ImportError as e:  # fail loudly: the product path has no CPU fallback
    raise ImportError(
        "dmsc_amd: the HIP extension is not built (run `python -c \"import __graft_entry__ as g; g.build()\"` "
        f"or `python dmsc_amd/_build.py`): {e}"
    ) from e


@dataclass
class MSComplex:
    """Discrete Morse-Smale complex of a 2-D field (field order of reference dmsc/dmsc.py:9-70).

    Primal (``is_dual=False``): maxima are faces on the (H+1)x(W+1) corner grid, minima are pixels.
    Dual: the roles are exchanged. Points, ridges and valleys are raw cell ids
    (id = type + 4*(y*(W+1) + x); type 0 pixel, 1 horizontal edge, 2 vertical edge, 3 corner).
    """

    shape: torch.Tensor
    max_pts: torch.Tensor
    min_pts: torch.Tensor
    sad_pts: torch.Tensor
    e_max: torch.Tensor
    e_min: torch.Tensor
    p_max: torch.Tensor
    p_min: torch.Tensor
    ppairs_max: torch.Tensor
    ppairs_min: torch.Tensor
    grad_indices: torch.Tensor
    peaks: torch.Tensor
    ridges: torch.Tensor
    ridge_offsets: torch.Tensor
    ridge_arc_offsets: torch.Tensor
    basins: torch.Tensor
    valleys: torch.Tensor
    valley_offsets: torch.Tensor
    valley_arc_offsets: torch.Tensor

    def __str__(self):
        parts = [f" {f.name}:{tuple(getattr(self, f.name).shape)}," for f in fields(self)
                 if hasattr(getattr(self, f.name), "shape")]
        return "MSComplex(" + "".join(parts) + ")"

    @property
    def Nx(self) -> int:
        return int(self.shape[1].item()) + 1

    def to_coordinates_yx(self, cell_tensor: torch.Tensor, staggered: bool = True) -> torch.Tensor:
        """Cell ids -> (N, 2) [y, x]; with ``staggered`` edges/corners get their -0.5 offsets
        (reference dmsc.py:87-118)."""
        if cell_tensor.numel() == 0:
            return torch.empty((0, 2), dtype=torch.float32 if staggered else torch.int32, device=cell_tensor.device)
        kind = cell_tensor % 4
        slot = cell_tensor // 4
        y = slot // self.Nx
        x = slot % self.Nx
        if not staggered:
            return torch.stack([y, x], dim=-1).to(torch.int32)
        y = y.float() - torch.where((kind == 2) | (kind == 3), 0.5, 0.0)
        x = x.float() - torch.where((kind == 1) | (kind == 3), 0.5, 0.0)
        return torch.stack([y, x], dim=-1)

    @property
    def offset_max_pts_yx(self):
        return self.to_coordinates_yx(self.max_pts)

    @property
    def offset_min_pts_yx(self):
        return self.to_coordinates_yx(self.min_pts)

    @property
    def offset_sad_pts_yx(self):
        return self.to_coordinates_yx(self.sad_pts)

    def _csr_slice(self, flat, offsets, arc_offsets, saddle_idx, split_arcs):
        start = offsets[saddle_idx]
        end = offsets[saddle_idx + 1] if saddle_idx + 1 < len(offsets) else len(flat)
        pts = flat[start:end]
        if not split_arcs:
            return pts
        cut = arc_offsets[saddle_idx]
        return pts[: cut - start], pts[cut - start : end - start]

    def get_ridge(self, saddle_idx: int, split_arcs=False):
        """Ridge cells of output saddle ``saddle_idx`` (both arcs, or split; reference dmsc.py:132-143)."""
        return self._csr_slice(self.ridges, self.ridge_offsets, self.ridge_arc_offsets, saddle_idx, split_arcs)

    def get_valley(self, saddle_idx: int, split_arcs=False):
        """Valley cells of output saddle ``saddle_idx`` (reference dmsc.py:145-156)."""
        return self._csr_slice(self.valleys, self.valley_offsets, self.valley_arc_offsets, saddle_idx, split_arcs)

    def __iter__(self):
        # the result tensors themselves, in field order (no copies)
        return iter(tuple(getattr(self, f.name) for f in fields(self)))

    def plot(self, *args, **kwargs):
        raise NotImplementedError("plotting dashboards are not part of dmsc_amd (use the reference's dmsc.plots)")

    plot_gradient = plot_complex = plot_barcode = plot

    @classmethod
    def _wrap(cls, native):
        """Result object over the extension's 19-tensor record (``native[i]`` is field i). Each field's
        Python tensor object is made on its first access and kept: for a batch of hundreds of fields,
        making all 19 up front costs about 18 us per field of host time between GPU calls."""
        obj = cls.__new__(cls)
        obj.__dict__["_native"] = native
        return obj

    def __getattr__(self, name):
        # only reached for a field not yet taken from the record (and for unknown names)
        native = self.__dict__.get("_native")
        i = _FIELD_INDEX.get(name)
        if native is None or i is None:
            raise AttributeError(f"{type(self).__name__!r} object has no attribute {name!r}")
        value = native[i]
        self.__dict__[name] = value
        return value


_FIELD_INDEX = {f.name: i for i, f in enumerate(fields(MSComplex))}


def _tuned(device):
    """Saved scheduling parameters of this GPU's architecture (autotune.py), applied once per device."""
    from . import autotune

    autotune.apply_saved(device.index if device.index is not None else torch.cuda.current_device())


def _validate(img):
    # reference dmsc.py:245-254 (same exception types and messages)
    if img.dim() not in (2, 3):
        raise ValueError(f"Input tensor must be 2D [H, W] or 3D [B, H, W]. Got {img.dim()}D.")
    if img.layout != torch.strided:
        raise ValueError(f"Input tensor must use strided layout. Got {img.layout}.")
    if img.dtype != torch.float32:
        raise TypeError(f"Input tensor must have dtype torch.float32. Got {img.dtype}.")
    if img.dim() == 3 and img.shape[0] == 0:
        raise ValueError("Batch dimension must be non-empty.")
    if img.shape[-2] == 0 or img.shape[-1] == 0:
        raise ValueError(f"Spatial dimensions must be non-empty. Got {tuple(img.shape[-2:])}.")


def _compute_dmsc_on_device(img, persistence_threshold, return_gradient, is_dual, trace_max_arcs, trace_min_arcs,
                            trace_max_groups, trace_min_groups):
    """Device dispatch (reference dmsc.py:196-227). Only GPU tensors are accepted."""
    if img.device.type != "cuda":
        raise RuntimeError(
            f"dmsc_amd computes on AMD Instinct GPUs only; got a tensor on '{img.device}'. Move it with .to('cuda')."
        )
    _tuned(img.device)
    return dmsc_gpu.extract_dmsc(img, persistence_threshold, return_gradient, is_dual, trace_max_arcs, trace_min_arcs,
                                 trace_max_groups, trace_min_groups)


def compute_dmsc(img, persistence_threshold, return_gradient=True, is_dual=False, trace_valleys=True,
                 trace_ridges=True, trace_peaks=True, trace_basins=True, verbose=False):
    """Discrete Morse-Smale complex of a [H, W] field (-> MSComplex) or a [B, H, W] batch (-> list).

    Flag mapping of the reference (dmsc.py:257-266): ridges -> max arcs, valleys -> min arcs,
    peaks -> max groups, basins -> min groups.
    """
    _validate(img)
    t0 = time.perf_counter()
    out = _compute_dmsc_on_device(img, persistence_threshold, return_gradient, is_dual, trace_max_arcs=trace_ridges,
                                  trace_min_arcs=trace_valleys, trace_max_groups=trace_peaks,
                                  trace_min_groups=trace_basins)
    if verbose:
        print(f"Time for MS-complex computation: {(time.perf_counter() - t0) * 1000:.3f} ms")
    if img.dim() == 2:
        return MSComplex._wrap(out)
    return [MSComplex._wrap(r) for r in out]


def compute_dmsc_batch(fields_list, persistence_thresholds, return_gradient=True, is_dual=False, trace_valleys=True,
                       trace_ridges=True, trace_peaks=True, trace_basins=True):
    """Fields of independent shapes with one threshold each (no reference counterpart: the reference
    batch requires one shape and one threshold). Returns a list of MSComplex, each identical to
    ``compute_dmsc(fields_list[b], persistence_thresholds[b], ...)``."""
    if len(fields_list) == 0:
        raise ValueError("Batch dimension must be non-empty.")
    for f in fields_list:
        _validate(f)
        if f.dim() != 2:
            raise ValueError("compute_dmsc_batch takes a list of 2D [H, W] fields.")
    if isinstance(persistence_thresholds, (int, float)):
        persistence_thresholds = [float(persistence_thresholds)] * len(fields_list)
    if fields_list[0].device.type != "cuda":
        raise RuntimeError("dmsc_amd computes on AMD Instinct GPUs only; move the fields with .to('cuda').")
    _tuned(fields_list[0].device)
    out = dmsc_gpu.extract_dmsc_list(list(fields_list), [float(t) for t in persistence_thresholds], return_gradient,
                                     is_dual, trace_ridges, trace_valleys, trace_peaks, trace_basins)
    return [MSComplex._wrap(r) for r in out]
