"""Batched extraction (all fields of a call share every launch and every persistence / geometry
round) against the oracle, field by field: mixed shapes and thresholds, all flag combinations,
heavy ties across field boundaries, the 3-D batch entry, and batches split into several groups."""

import itertools
import os

import numpy as np
import pytest

from dmsc_amd import synthetic
from tests.parity_util import diff_outputs, oracle_extract

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1), (1, 2), (2, 1), (1, 9), (9, 1), (2, 2), (3, 5), (7, 11), (16, 64), (17, 65), (31, 63), (64, 129),
          (100, 37), (40, 33), (5, 200)]
THRS = (-1.0, 0.0, 0.1, 0.5)


def run_list(fields, thrs, **kw):
    import torch

    from dmsc_amd.csrc import dmsc_gpu

    args = [kw.get(k, True) for k in ("return_gradient",)] + [kw.get("is_dual", False)] + \
        [kw.get(k, True) for k in ("tmax", "tmin", "gmax", "gmin")]
    res = dmsc_gpu.extract_dmsc_list([torch.from_numpy(np.ascontiguousarray(f)).cuda() for f in fields],
                                     [float(t) for t in thrs], *args)
    torch.cuda.synchronize()
    return [[r[i].cpu().numpy() for i in range(19)] for r in res]


def check_batch(fields, thrs, **kw):
    print(f"[batch] {len(fields)} fields {kw}", flush=True)
    got = run_list(fields, thrs, **kw)
    assert len(got) == len(fields)
    for j, (f, t, g) in enumerate(zip(fields, thrs, got)):
        bad = diff_outputs(g, oracle_extract(f, t, **kw))
        assert not bad, f"field {j} shape={f.shape} thr={t} {kw}: {bad}"


@pytest.mark.parametrize("is_dual", [False, True])
def test_mixed_shapes_and_thresholds(is_dual):
    rng = np.random.default_rng(11)
    fields = [rng.random(s).astype(np.float32) for s in SHAPES]
    thrs = [THRS[j % len(THRS)] for j in range(len(fields))]
    check_batch(fields, thrs, is_dual=is_dual)


@pytest.mark.parametrize("is_dual", [False, True])
@pytest.mark.parametrize("flags", list(itertools.product([True, False], repeat=4)))
def test_flag_combinations(is_dual, flags):
    tmax, tmin, gmax, gmin = flags
    rng = np.random.default_rng(5)
    fields = [rng.random(s).astype(np.float32) for s in [(9, 1), (33, 20), (12, 70), (1, 1), (48, 48)]]
    thrs = [0.0, 0.2, -1.0, 0.1, 0.05]
    check_batch(fields, thrs, return_gradient=not gmin, is_dual=is_dual, tmax=tmax, tmin=tmin, gmax=gmax, gmin=gmin)


@pytest.mark.parametrize("is_dual", [False, True])
def test_ties_across_fields(is_dual):
    # identical values in every field: the batch-wide sorts must keep each field's (value, id) order
    fields = [synthetic.quantized_noise(h, w, lv, seed=lv).numpy() for h, w, lv in
              [(30, 41, 2), (30, 41, 2), (17, 9, 3), (64, 64, 2), (5, 5, 1)]]
    fields.append(np.zeros((6, 7), np.float32))
    check_batch(fields, [0.0, 0.3, 0.05, 0.0, 0.0, 0.0], is_dual=is_dual)


def test_gaussian_fields_with_per_field_thresholds():
    fields, thrs = synthetic.mixed_batch(12, seed=9, sizes=tuple(range(32, 161, 32)))
    check_batch([f.numpy() for f in fields], thrs)


def test_3d_batch_entry():
    import torch

    from dmsc_amd.csrc import dmsc_gpu

    stack = np.random.default_rng(2).random((5, 33, 47)).astype(np.float32)
    res = dmsc_gpu.extract_dmsc(torch.from_numpy(stack).cuda(), 0.1, True, True, True, True, True, True)
    torch.cuda.synchronize()
    assert len(res) == 5
    for b in range(5):
        bad = diff_outputs([res[b][i].cpu().numpy() for i in range(19)], oracle_extract(stack[b], 0.1, is_dual=True))
        assert not bad, f"field {b}: {bad}"


def test_batch_split_into_groups():
    rng = np.random.default_rng(7)
    fields = [rng.random(s).astype(np.float32) for s in [(30, 40), (1, 1), (50, 50), (20, 90), (3, 3), (45, 31)]]
    thrs = [0.1, 0.0, 0.2, 0.05, 0.0, 0.3]
    os.environ["DMSC_GROUP_SLOTS"] = "2000"  # several groups, one oversize field alone
    try:
        check_batch(fields, thrs)
    finally:
        del os.environ["DMSC_GROUP_SLOTS"]


@pytest.mark.parametrize("is_dual", [False, True])
def test_pipelined_halves(is_dual):
    """A batch run as parts side by side (all but the first on worker threads with their own contexts
    and streams) against the oracle, with each part split into several groups too, and as two, three
    and four parts bit for bit against the same batch as one pipeline."""
    import torch

    from dmsc_amd.csrc import dmsc_gpu

    rng = np.random.default_rng(23)
    fields = [rng.random(s).astype(np.float32) for s in [(40, 33), (1, 1), (64, 129), (7, 11), (90, 20), (3, 80)]]
    fields.append(synthetic.quantized_noise(80, 60, 3, seed=4).numpy())
    thrs = [0.0, 0.1, 0.3, 0.05, 0.2, 0.0, 0.1]
    saved = {k: dmsc_gpu.get_param(0, k) for k in ("pipe_groups", "pipe_min_slots")}
    try:
        dmsc_gpu.set_param(0, "pipe_groups", 2)
        dmsc_gpu.set_param(0, "pipe_min_slots", 0)
        check_batch(fields, thrs, is_dual=is_dual)
        os.environ["DMSC_GROUP_SLOTS"] = "3000"
        try:
            check_batch(fields, thrs, is_dual=is_dual)
        finally:
            del os.environ["DMSC_GROUP_SLOTS"]
        parts = []
        for groups in (2, 3, 4):  # parts side by side, each on its own worker thread and stream
            dmsc_gpu.set_param(0, "pipe_groups", groups)
            parts.append(run_list(fields, thrs, is_dual=is_dual))
        dmsc_gpu.set_param(0, "pipe_groups", 1)
        one = run_list(fields, thrs, is_dual=is_dual)
    finally:
        for k, v in saved.items():
            dmsc_gpu.set_param(0, k, v)
    torch.cuda.synchronize()
    for split in parts:
        for j, (a, b) in enumerate(zip(split, one)):
            for i in range(19):
                assert a[i].shape == b[i].shape and a[i].tobytes() == b[i].tobytes(), (j, i)


def test_scheduling_parameters_do_not_change_results():
    import torch

    from dmsc_amd.csrc import dmsc_gpu

    rng = np.random.default_rng(17)
    fields = [rng.random(s).astype(np.float32) for s in [(40, 33), (64, 129), (7, 11)]]
    fields.append(synthetic.quantized_noise(80, 60, 3, seed=4).numpy())
    thrs = [0.0, 0.1, 0.3, 0.05]
    names = ("round_grid", "small_list", "rounds_first", "rounds_max", "expand_levels", "graph_rounds", "tail_items",
             "grad_rows", "side_rounds")
    saved = {k: dmsc_gpu.get_param(0, k) for k in names}
    try:
        for setting in ({"round_grid": 64, "small_list": 0, "rounds_first": 1, "rounds_max": 1, "expand_levels": 1,
                         "graph_rounds": 0, "tail_items": 0, "grad_rows": 8, "side_rounds": 4096},
                        {"round_grid": 8192, "small_list": 1 << 20, "rounds_first": 64, "rounds_max": 64,
                         "expand_levels": 64, "graph_rounds": 0, "tail_items": 1 << 16, "grad_rows": 32,
                         "side_rounds": 0},
                        {"round_grid": 1024, "small_list": 4096, "rounds_first": 2, "rounds_max": 8, "expand_levels": 4,
                         "graph_rounds": 1, "tail_items": 64, "grad_rows": 16, "side_rounds": 2}):
            for k, v in setting.items():
                dmsc_gpu.set_param(0, k, v)
            check_batch(fields, thrs, is_dual=False)
            check_batch(fields, thrs, is_dual=True)
        with pytest.raises(RuntimeError):
            dmsc_gpu.set_param(0, "round_grid", 1)
        with pytest.raises(RuntimeError):
            dmsc_gpu.set_param(0, "no_such_parameter", 1)
        with pytest.raises(RuntimeError):
            dmsc_gpu.set_param(0, "grad_rows", 12)
    finally:
        for k, v in saved.items():
            dmsc_gpu.set_param(0, k, v)
    torch.cuda.synchronize()


def test_autotune_smoke():
    from dmsc_amd import autotune

    res = autotune.tune(size=256, n_small=4, reps=1, save=False, log=lambda *_: None,
                        candidates={"round_grid": [512], "rounds_max": [8]})
    assert set(res["params"]) == set(autotune.DEFAULTS)
    autotune.apply(autotune.DEFAULTS)


def test_graph_replayed_round_batches():
    """Round batches captured once as a hipGraph and replayed batch after batch within one call (host
    readbacks and eager launches between the replays): every replay must leave exactly the state the
    eager launches leave, which the outputs check field by field."""
    from dmsc_amd.csrc import dmsc_gpu

    fields = [synthetic.quantized_noise(300, 280, 6, seed=31).numpy(), synthetic.quantized_noise(250, 300, 4, seed=32).numpy()]
    thrs = [0.2, 0.5]
    names = ("graph_rounds", "rounds_first", "rounds_max", "tail_items")
    saved = {k: dmsc_gpu.get_param(0, k) for k in names}
    try:
        for k, v in {"graph_rounds": 1, "rounds_first": 2, "rounds_max": 2, "tail_items": 0}.items():
            dmsc_gpu.set_param(0, k, v)
        for rep in range(2):
            check_batch(fields, thrs, is_dual=bool(rep))
            rounds = dmsc_gpu.debug_fetch(0, "rounds", 3).tolist()
            assert rounds[0] >= 6, rounds  # at least three replays of the 2-round persistence batch
    finally:
        for k, v in saved.items():
            dmsc_gpu.set_param(0, k, v)

