"""Autotuner for the scheduling parameters of the HIP pipeline (gradient tile height, round kernel
grid cap, one-item-per-thread threshold, host-check cadence of the round and expansion loops,
hipGraph replay of round batches).

No reference counterpart. The parameters change only how work is scheduled, never the results: the
tuner checks that every candidate gives outputs identical to the defaults' and rejects it otherwise.
It runs a coordinate descent over candidate values on synthetic fields (a large Gaussian mixture and
a batch of small fields), timing each setting on the device, and saves the winners per GPU
architecture under ``dmsc_amd/tuning/<arch>.json``. ``apply_saved(device)`` loads them; the Python
API applies them once per device before its first extraction.

    python -m dmsc_amd.autotune            # tune device 0 and save
"""

import json
import os
import re
import time

import torch

from .csrc import dmsc_gpu

TUNING_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning")
DEFAULTS = {"round_grid": 2048, "small_list": 65536, "rounds_first": 4, "rounds_max": 16, "expand_levels": 8,
            "graph_rounds": 0, "tail_items": 2048, "grad_rows": 16, "side_rounds": 1}
CANDIDATES = {
    "round_grid": [512, 1024, 2048, 4096, 8192, 16384],
    "small_list": [4096, 16384, 65536, 262144],
    "rounds_first": [2, 4, 8],
    "rounds_max": [8, 16, 32, 64],
    "expand_levels": [4, 8, 16],
    "graph_rounds": [0, 1],
    "tail_items": [0, 512, 2048, 8192],
    "grad_rows": [8, 16, 32],
    "side_rounds": [0, 1, 2, 3, 4],
}
_applied = set()


def arch_name(device=0):
    name = torch.cuda.get_device_properties(device).gcnArchName
    return re.sub(r"[^A-Za-z0-9_.-]", "_", name.split(":")[0])


def tuning_path(device=0):
    return os.path.join(TUNING_DIR, arch_name(device) + ".json")


def apply(params, device=0):
    for k, v in params.items():
        dmsc_gpu.set_param(device, k, int(v))


def apply_saved(device=0):
    """Apply the saved parameters of this device's architecture, if any (once per device)."""
    if device in _applied:
        return
    _applied.add(device)
    path = tuning_path(device)
    if os.path.exists(path):
        with open(path) as fh:
            apply(json.load(fh)["params"], device)


def _workloads(device, size, n_small):
    from . import synthetic

    big = synthetic.gaussian_mixture(size, size, max(16, size // 16), (size / 256 + 1, size / 16 + 1), 0.02,
                                     seed=size).to(device)
    fields, thrs = synthetic.mixed_batch(n_small, seed=n_small, sizes=tuple(range(64, 513, 32)))
    return big, [f.to(device) for f in fields], thrs


def _run(big, small, thrs):
    outs = [dmsc_gpu.extract_dmsc(big, thr, True, dual, True, True, True, True)
            for dual, thr in ((False, 0.05), (True, 0.2))]
    outs += list(dmsc_gpu.extract_dmsc_list(small, thrs, True, False, True, True, True, True))
    return outs


def _same(a, b):
    for x, y in zip(a, b):
        for i in range(19):
            tx, ty = x[i], y[i]
            if tx.shape != ty.shape:
                return False
            if tx.is_floating_point():
                if not torch.equal(torch.nan_to_num(tx, nan=-7.0), torch.nan_to_num(ty, nan=-7.0)):
                    return False
            elif not torch.equal(tx, ty):
                return False
    return True


def _time(big, small, thrs, reps):
    best = float("inf")
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _run(big, small, thrs)
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best


def tune(device=0, size=2048, n_small=64, reps=3, save=True, candidates=None, log=print):
    """Coordinate descent from the defaults; returns {"params": ..., "seconds": ...}."""
    candidates = candidates or CANDIDATES
    big, small, thrs = _workloads(device, size, n_small)
    params = dict(DEFAULTS)
    apply(params, device)
    ref = _run(big, small, thrs)
    best_t = _time(big, small, thrs, reps)
    log(f"defaults: {best_t * 1e3:.2f} ms")
    for name, values in candidates.items():
        for v in values:
            if v == params[name]:
                continue
            trial = dict(params, **{name: v})
            apply(trial, device)
            if not _same(_run(big, small, thrs), ref):
                raise RuntimeError(f"outputs changed with {name}={v}: scheduling parameters must not change results")
            t = _time(big, small, thrs, reps)
            log(f"  {name}={v}: {t * 1e3:.2f} ms")
            if t < best_t * 0.99:
                best_t, params = t, trial
        apply(params, device)
    result = {"arch": arch_name(device), "params": params, "seconds": best_t,
              "workload": f"{size}^2 gaussian mixture primal 0.05 + dual 0.2, {n_small} mixed fields 64-512"}
    if save:
        os.makedirs(TUNING_DIR, exist_ok=True)
        with open(tuning_path(device), "w") as fh:
            json.dump(result, fh, indent=1)
    return result


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser(description="pick the scheduling parameters for this GPU")
    ap.add_argument("--size", type=int, default=2048, help="side of the square gaussian-mixture field")
    ap.add_argument("--n-small", type=int, default=64, help="fields in the mixed batch")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--no-save", action="store_true")
    a = ap.parse_args()
    res = tune(size=a.size, n_small=a.n_small, reps=a.reps, save=not a.no_save)
    res["note"] = f"written by dmsc_amd.autotune.tune(size={a.size}, n_small={a.n_small}, reps={a.reps})"
    print(json.dumps(res))
