"""API behaviour of dmsc_amd.compute_dmsc, mirroring the reference's tests/test_api.py: argument
validation messages, exact equality of batched vs sequential calls (including device placement),
non-contiguous inputs, determinism on tied / non-finite fields; plus the mixed-size list batch."""

import pytest
import torch

from dmsc_amd import compute_dmsc, compute_dmsc_batch
from dmsc_amd import synthetic


def assert_complex_equal(actual, expected):
    for a, e in zip(actual, expected, strict=True):
        assert a.device == e.device
        a, e = a.cpu(), e.cpu()
        if a.is_floating_point():
            torch.testing.assert_close(a, e, rtol=0, atol=0, equal_nan=True)
        else:
            assert torch.equal(a, e)


@pytest.mark.parametrize("shape", [(0, 4), (4, 0), (0, 4, 4)])
def test_rejects_empty_dimensions(shape):
    with pytest.raises(ValueError, match="non-empty"):
        compute_dmsc(torch.empty(shape), 0.0)


@pytest.mark.parametrize("dtype", [torch.float16, torch.float64, torch.int32, torch.int64])
def test_rejects_unsupported_dtype(dtype):
    with pytest.raises(TypeError, match="float32"):
        compute_dmsc(torch.zeros((4, 4), dtype=dtype), 0.0)


def test_rejects_wrong_rank():
    with pytest.raises(ValueError, match="2D"):
        compute_dmsc(torch.zeros(4), 0.0)


def test_cpu_tensor_fails_loudly():
    with pytest.raises(RuntimeError, match="GPU"):
        compute_dmsc(torch.zeros((4, 4)), 0.0)


@pytest.mark.gpu
def test_accepts_noncontiguous_input():
    field = torch.randn((7, 11), generator=torch.Generator().manual_seed(4)).transpose(0, 1).cuda()
    assert not field.is_contiguous()
    assert_complex_equal(compute_dmsc(field, 0.1), compute_dmsc(field.contiguous(), 0.1))


@pytest.mark.gpu
def test_batch_matches_sequential_calls():
    fields = torch.randn((3, 9, 7), generator=torch.Generator().manual_seed(8)).cuda()
    batched = compute_dmsc(fields, 0.15)
    sequential = [compute_dmsc(f, 0.15) for f in fields]
    assert len(batched) == len(sequential)
    for a, e in zip(batched, sequential, strict=True):
        assert_complex_equal(a, e)


@pytest.mark.gpu
@pytest.mark.parametrize("value", [0.0, float("inf"), float("-inf"), float("nan")])
def test_tied_and_nonfinite_fields_are_deterministic(value):
    field = torch.full((6, 5), value).cuda()
    assert_complex_equal(compute_dmsc(field, 0.0), compute_dmsc(field, 0.0))


@pytest.mark.gpu
@pytest.mark.parametrize("is_dual", [False, True])
def test_mixed_size_list_batch_matches_plain_calls(is_dual):
    fields, thrs = synthetic.mixed_batch(24, seed=3, sizes=tuple(range(16, 200, 16)))
    fields = [f.cuda() for f in fields]
    batched = compute_dmsc_batch(fields, thrs, is_dual=is_dual)
    for f, t, b in zip(fields, thrs, batched, strict=True):
        assert_complex_equal(b, compute_dmsc(f, t, is_dual=is_dual))


@pytest.mark.gpu
def test_msc_helpers():
    ms = compute_dmsc(synthetic.gaussian_mixture(40, 30, 6, (3, 8), 0.05, seed=2).cuda(), 0.05)
    assert ms.Nx == 31
    yx = ms.to_coordinates_yx(ms.max_pts, staggered=False)
    assert yx.shape == (ms.max_pts.numel(), 2)
    for i in range(ms.sad_pts.numel()):
        r0, r1 = ms.get_ridge(i, split_arcs=True)
        assert torch.equal(torch.cat([r0, r1]), ms.get_ridge(i))
        v0, v1 = ms.get_valley(i, split_arcs=True)
        assert torch.equal(torch.cat([v0, v1]), ms.get_valley(i))


@pytest.mark.parametrize("batched", [False, True])
def test_cell_id_limit_of_the_binding(batched):
    """The binding's 32-bit cell-id check (reference input_validation.hxx:24-26: H + 1 <= INT32_MAX / 4 /
    (W + 1), integer division) runs before anything touches the device: the largest square that fits is
    23169^2 (refused only as a CPU tensor), 23170^2 is refused as too large. Expanded views, no memory."""
    from dmsc_amd.csrc import dmsc_gpu

    def view(h, w):
        t = torch.zeros(1, 1).expand(h, w)
        return t.unsqueeze(0) if batched else t

    with pytest.raises(RuntimeError, match="too large for 32-bit cell IDs"):
        dmsc_gpu.extract_dmsc(view(23170, 23170), 0.1)
    with pytest.raises(RuntimeError, match="too large for 32-bit cell IDs"):
        dmsc_gpu.extract_dmsc(view(1, (2**31 - 1) // 8), 0.1)  # H + 1 = 2 > INT32_MAX / 4 / (W + 1) = 1
    with pytest.raises(RuntimeError, match="GPU backend requires"):
        dmsc_gpu.extract_dmsc(view(23169, 23169), 0.1)


def test_result_objects_made_from_the_native_record():
    """compute_dmsc wraps the extension's 19-tensor record (MSComplex._wrap): every field, iteration,
    equality, printing and dataclass helpers behave as for an object built from the 19 tensors."""
    import dataclasses

    from dmsc_amd import MSComplex

    record = [torch.arange(k + 1, dtype=torch.int32) for k in range(19)]
    lazy, eager = MSComplex._wrap(record), MSComplex(*record)
    names = [f.name for f in dataclasses.fields(MSComplex)]
    assert all(getattr(lazy, n) is record[i] for i, n in enumerate(names))
    assert all(a is b for a, b in zip(lazy, eager, strict=True))
    assert str(MSComplex._wrap(record)) == str(eager)
    assert all(torch.equal(a, b) for a, b in zip(dataclasses.astuple(MSComplex._wrap(record)), record, strict=True))
    assert dataclasses.replace(MSComplex._wrap(record), shape=record[0]).ridges is record[12]
    assert not hasattr(lazy, "no_such_field")
    with pytest.raises(AttributeError):
        lazy.no_such_field
