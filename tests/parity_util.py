"""Shared helpers: run the HIP path and the oracle on the same float32 bytes and compare outputs."""

import numpy as np

import oracle

FIELDS = oracle.FIELDS


def gpu_extract(field, thr, return_gradient=True, is_dual=False, tmax=True, tmin=True, gmax=True, gmin=True):
    import torch

    from dmsc_amd.csrc import dmsc_gpu

    t = torch.as_tensor(np.ascontiguousarray(field, dtype=np.float32)).to("cuda")
    r = dmsc_gpu.extract_dmsc(t, float(thr), return_gradient, is_dual, tmax, tmin, gmax, gmin)
    torch.cuda.synchronize()
    return [r[i].cpu().numpy() for i in range(19)]


def oracle_extract(field, thr, return_gradient=True, is_dual=False, tmax=True, tmin=True, gmax=True, gmin=True,
                   nthreads=1):
    return oracle.extract(field, thr, return_gradient, is_dual, tmax, tmin, gmax, gmin, nthreads=nthreads)


def float_equal(a, b):
    """Bit-exact except that any NaN matches any NaN (x86 and gfx950 produce different NaN payloads)."""
    if a.shape != b.shape:
        return False
    na, nb = np.isnan(a), np.isnan(b)
    if not np.array_equal(na, nb):
        return False
    return np.array_equal(a[~na].view(np.int32), b[~nb].view(np.int32))


def diff_outputs(got, exp):
    """List of (slot name, description) for every output that differs."""
    bad = []
    for i, name in enumerate(FIELDS):
        g, e = np.asarray(got[i]), np.asarray(exp[i])
        if g.shape != e.shape:
            bad.append((name, f"shape {g.shape} vs {e.shape}"))
            continue
        ok = float_equal(g, e) if e.dtype == np.float32 else np.array_equal(g, e)
        if not ok:
            idx = np.flatnonzero(g.reshape(-1) != e.reshape(-1))
            first = int(idx[0]) if len(idx) else -1
            bad.append((name, f"{len(idx)} of {e.size} differ; first at {first}: "
                              f"{g.reshape(-1)[first] if first >= 0 else None} vs {e.reshape(-1)[first] if first >= 0 else None}"))
    return bad


def pool_workers():
    """Worker threads for oracle runs: the host's CPUs, at most 16 (one GPU's CPU share)."""
    import os

    return max(1, min(os.cpu_count() or 1, 16))


def oracle_many(cases, **kw):
    """Oracle outputs of many (field, thr) cases, one single-threaded oracle call per worker thread
    (the oracle releases the GIL and keeps no global state), yielded in case order."""
    from concurrent.futures import ThreadPoolExecutor

    with ThreadPoolExecutor(pool_workers()) as ex:
        yield from ex.map(lambda c: oracle_extract(c[0], c[1], nthreads=1, **kw), cases)
