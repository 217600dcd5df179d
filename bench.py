"""Benchmark of the MI355X compute_dmsc path.

Default workload (one GPU): 512 float32 Gaussian-mixture fields of independent sizes (H, W uniform
over {64, 96, ..., 1024}, 8-64 bumps, per-field persistence threshold uniform in [0.02, 0.2],
seeded), full DMSC (gradient returned, ridges, valleys, peaks, basins). One step = one batched call
over all 512 fields, each step on its own device copies of the fields. Metric: Mpix/s = pixels
processed / wall time of the timed steps; the same fields through the per-field plain path (warmed
up, timed over 3 steps) give plain_loop_Mpix_s and batched_over_plain. The 8192^2 quantised-noise
field (thr 0.05) and the 4096^2 Gaussian mixture (primal + dual x thr {0, 0.05, 0.2}) are measured
after it and reported under the extra keys noise8192 and gauss4096.

Prints one JSON line with the metric, the roofline object of the dominant kernel (algorithmic
bytes / average launch time from HIP events recorded around its launches during the timed steps),
the end-to-end algorithmic HBM rate, and the CPU oracle timed on a bounded sample of the same
workload on this host (all allotted cores, and one thread).

Other workloads (--workload): gauss4096, noise8192 (one line each, same fields as above), diff
(persistence loss over 64 tie-free fields of 512x512: batched forward + hand-written backward per
step). --cpu-table times the CPU oracle on every configuration.
"""

import argparse
import contextlib
import gc
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per launch of each instrumented kernel, per unit it processes:
#   gradient:      4 B field read per pixel + 16 B gradient write per slot
#   pointer_jump:  8 B (terminal, depth) read + 8 B written per forest node
#   labels_*:      8 B terminal read + 4 B label written per cell
KERNEL_MODELS = {
    "gradient": lambda H, W: 4 * H * W + 16 * (H + 1) * (W + 1),
}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--workload", default="mixed", choices=["mixed", "gauss4096", "noise8192", "diff"])
    p.add_argument("--headline-only", action="store_true",
                   help="mixed workload without the 8192^2 and 4096^2 configurations beside it")
    p.add_argument("--no-plain", action="store_true",
                   help="mixed workload without the per-field plain loop (for profiling the batched calls alone)")
    p.add_argument("--cpu-table", action="store_true",
                   help="only time the CPU oracle on every configuration (all allotted cores and 1 thread)")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    p.add_argument("--pipe-groups", type=int, choices=tuple(range(1, 9)), default=None,
                   help="run large batches as one pipeline (1) or as this many parts side by side (the "
                        "library's default is 4) for the whole run, e.g. so that a rocprofv3 run matches the "
                        "roofline pass")
    p.add_argument("--profile", action="store_true", help="print per-kernel event timings to stderr")
    p.add_argument("--copy-calibration", action="store_true",
                   help="only run plain 256 MiB device-to-device copies (reference kernel for PMC byte counters)")
    p.add_argument("--summarize-profile", nargs=5, metavar=("TRACE_CSV", "FETCH_CSV", "WRITE_CSV", "COPY_FETCH_CSV",
                                                            "COPY_WRITE_CSV"),
                   help="print a per-kernel table (duration, calibrated HBM bytes, GB/s vs peak) from rocprofv3 "
                        "kernel-trace and FETCH_SIZE / WRITE_SIZE counter CSVs of the same command")
    return p.parse_args()


@contextlib.contextmanager
def gc_paused():
    """Python's cyclic garbage collector paused over a timed region, as timeit does: a full collection
    of this process's heap (torch, numpy, the result objects) takes about 45 ms, a host pause unrelated
    to the GPU work that would otherwise land in one timed step at random. One collection runs first."""
    gc.collect()
    was_enabled = gc.isenabled()
    gc.disable()
    try:
        yield
    finally:
        if was_enabled:
            gc.enable()


def copy_calibration():
    """Five copies of 256 MiB between two device buffers: each moves exactly 256 MiB in and out, so
    FETCH_SIZE / WRITE_SIZE of these dispatches calibrate the counters' units."""
    import torch

    n = 64 * 1024 * 1024  # float32 elements = 256 MiB
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_per_dispatch": 4 * n, "dispatches": 5}))


def workload(name):
    from dmsc_amd import synthetic

    if name == "gauss4096":
        f = synthetic.gaussian_mixture(4096, 4096, 256, (16.0, 256.0), 0.02, seed=4096)
        runs = [(d, t) for d in (False, True) for t in (0.0, 0.05, 0.2)]
        desc = ("gaussian-mixture 4096x4096 fp32, 256 bumps, sigma 16-256 px, noise 0.02, seed 4096; primal+dual x "
                "thr {0, 0.05, 0.2}; full geometry + gradient")
        return [f], runs, desc
    if name == "noise8192":
        f = synthetic.quantized_noise(8192, 8192, 256, seed=8192)
        desc = "uniform noise 8192x8192 quantised to 256 levels, seed 8192; primal thr 0.05; full geometry + gradient"
        return [f], [(False, 0.05)], desc
    fields, thrs = synthetic.mixed_batch(512, seed=512)
    desc = ("512 gaussian-mixture fields, H,W uniform in {64..1024 step 32}, 8-64 bumps, sigma min(H,W)/32..min(H,W)/8, "
            "noise 0.05, thr U[0.02,0.2] per field, seed 512; primal; full geometry + gradient; one batched call "
            "per step vs the per-field plain loop")
    return fields, list(zip([False] * len(fields), thrs)), desc


def measured_traffic(workload, kernel):
    """HBM bytes per launch from the committed rocprofv3 PMC summary (profiles/traffic.json), if any."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as fh:
            return json.load(fh).get(workload, {}).get(kernel, {}).get("traffic_bytes_per_launch")
    except OSError:
        return None


def out_bytes(ms):
    return sum(t.numel() * t.element_size() for t in ms)


def summarize_profile(trace_csv, fetch_csv, write_csv, cfetch_csv, cwrite_csv):
    """Per-kernel table from rocprofv3 output: average duration from the kernel trace, bytes per
    dispatch from the FETCH_SIZE / WRITE_SIZE counters (KiB), each scaled so that a plain 256 MiB
    copy reads and writes 256 MiB, and the resulting GB/s against the 8 TB/s HBM peak."""
    import csv
    from collections import defaultdict

    def short(name):
        name = name.replace("void ", "").replace("dmsc::", "")
        return name.split("(")[0]

    def counters(path):
        acc = defaultdict(lambda: [0.0, set()])
        for r in csv.DictReader(open(path)):
            a = acc[short(r["Kernel_Name"])]
            a[0] += float(r["Counter_Value"])
            a[1].add(r["Dispatch_Id"])
        return {k: (v[0], len(v[1])) for k, v in acc.items()}

    copy_kib = 4 * 64 * 1024 * 1024 / 1024
    fetch_scale = copy_kib / (counters(cfetch_csv)["__amd_rocclr_copyBuffer"][0] /
                              counters(cfetch_csv)["__amd_rocclr_copyBuffer"][1])
    write_scale = copy_kib / (counters(cwrite_csv)["__amd_rocclr_copyBuffer"][0] /
                              counters(cwrite_csv)["__amd_rocclr_copyBuffer"][1])
    dur = defaultdict(lambda: [0, 0])
    for r in csv.DictReader(open(trace_csv)):
        d = dur[short(r["Kernel_Name"])]
        d[0] += int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
        d[1] += 1
    fetch, write = counters(fetch_csv), counters(write_csv)
    total_ns = sum(v[0] for v in dur.values())
    print(f"calibration: FETCH_SIZE x{fetch_scale:.3f}, WRITE_SIZE x{write_scale:.3f} (256 MiB copy)")
    print("| kernel | calls | total ms | % time | avg us | MB / dispatch (read + write) | GB/s | % of 8 TB/s |")
    print("|---|---:|---:|---:|---:|---:|---:|---:|")
    for k, (ns, n) in sorted(dur.items(), key=lambda kv: -kv[1][0]):
        fb = fetch.get(k, (0.0, 1))
        wb = write.get(k, (0.0, 1))
        mb = (fb[0] / max(fb[1], 1) * fetch_scale + wb[0] / max(wb[1], 1) * write_scale) * 1024 / 1e6
        avg_us = ns / n / 1e3
        gbs = mb * 1e6 / (avg_us * 1e3) if avg_us > 0 else 0.0
        print(f"| `{k}` | {n} | {ns / 1e6:.3f} | {100 * ns / total_ns:.2f} | {avg_us:.1f} | {mb:.2f} | {gbs:.1f} | "
              f"{100 * gbs / HBM_PEAK_GBS:.2f} |")


METRIC = "Mpix/s full DMSC (batched vs plain, 1 GPU) + achieved HBM GB/s, bit-exact"


def measure(name, steps, warmup, plain_steps=0, profile=False):
    """Times one extraction workload; returns the figures of its JSON line. mixed: one batched call
    per step over all fields, and (plain_steps > 0) the same fields through the per-field plain
    path, warmed up and timed over plain_steps steps. Other workloads: every extraction of a step on
    its own device copy of the field."""
    import torch

    import dmsc_amd
    from dmsc_amd.csrc import dmsc_gpu

    dev = torch.device("cuda:0")
    fields, runs, desc = workload(name)
    mixed = name == "mixed"
    n_ext = len(runs)
    total_steps = max(warmup, 1) + steps
    # distinct device copies per extraction and step (no cache reuse across extractions)
    if mixed:
        copies = [[f.to(dev) for f in fields] for _ in range(total_steps)]
    else:
        copies = [[fields[0].to(dev) for _ in range(n_ext)] for _ in range(total_steps)]
    thrs = [t for _, t in runs]
    torch.cuda.synchronize()

    def step(s, plain=False):
        if mixed and not plain:  # one batched call over all fields (per-field thresholds)
            return dmsc_amd.compute_dmsc_batch(copies[s], thrs, return_gradient=True)
        return [dmsc_amd.compute_dmsc(copies[s][j], thr, return_gradient=True, is_dual=dual)
                for j, (dual, thr) in enumerate(runs)]

    # bookkeeping on the first warmup step (untimed): output bytes per field, critical-cell counts per call
    sizes, counts = [], []
    if mixed:
        res = step(0)
        sizes = [(copies[0][j].shape, out_bytes(ms)) for j, ms in enumerate(res)]
        counts.append([int(v) for v in dmsc_gpu.debug_fetch(0, "counts", 3).tolist()])
        del res
    else:
        for j, (dual, thr) in enumerate(runs):
            ms = dmsc_amd.compute_dmsc(copies[0][j], thr, return_gradient=True, is_dual=dual)
            sizes.append((copies[0][j].shape, out_bytes(ms)))
            counts.append([int(v) for v in dmsc_gpu.debug_fetch(0, "counts", 3).tolist()])
    for s in range(1, max(warmup, 1)):
        step(s)
    torch.cuda.synchronize()
    calls_per_step = len(counts)

    def timed(idx, level):
        """Wall time of the steps idx, and the per-scope event times at profile `level` (1: events only
        around the gradient launches and each batch of persistence rounds; 0: none)."""
        dmsc_gpu.profile_enable(0, level)
        torch.cuda.synchronize()
        with gc_paused():
            t0 = time.perf_counter()
            for s in idx:
                step(s)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
        prof = dmsc_gpu.profile_read(0)
        dmsc_gpu.profile_enable(0, 0)
        return dt, prof

    # A large batch runs as several parts side by side (pipe_groups, 4 by default), whose sweeps and
    # gradient launches then share the GPU with the other parts' work. The kernels' roofline figures come from
    # a separate pass over the same steps with the batch as one pipeline (pipe_groups 1), so that each
    # launch's time is its own; the headline value is the timed region's, as the library runs it.
    split_roof = mixed and dmsc_gpu.get_param(0, "pipe_groups") > 1
    elapsed, prof = timed(range(max(warmup, 1), total_steps), 0 if split_roof else 1)
    roof_steps, roof_elapsed = steps, elapsed
    if split_roof:
        pipe = dmsc_gpu.get_param(0, "pipe_groups")
        dmsc_gpu.set_param(0, "pipe_groups", 1)
        try:
            roof_steps = min(steps, 3)
            step(max(warmup, 1))  # warmup of the single pipeline (its scratch)
            roof_elapsed, prof = timed(range(max(warmup, 1), max(warmup, 1) + roof_steps), 1)
        finally:
            dmsc_gpu.set_param(0, "pipe_groups", pipe)

    px_step = sum(int(sh[0]) * int(sh[1]) for sh, _ in sizes)
    value = px_step * steps / elapsed / 1e6
    b_alg_step = sum(4 * int(sh[0]) * int(sh[1]) + ob for sh, ob in sizes)
    e2e_gbs = b_alg_step * steps / elapsed / 1e9

    if profile:  # every stage timed, on one extra untimed step (its events slow the round loops down)
        dmsc_gpu.profile_enable(0, 2)
        step(max(warmup, 1))
        torch.cuda.synchronize()
        full = dmsc_gpu.profile_read(0)
        dmsc_gpu.profile_enable(0, 0)
        print(f"[{name}] per-stage event times of one step:", file=sys.stderr)
        for k, (ms_tot, n) in sorted(full.items(), key=lambda kv: -kv[1][0]):
            print(f"  {k:20s} total {ms_tot:9.3f} ms  launches {n:6d}  avg {ms_tot / max(n, 1) * 1e3:9.1f} us",
                  file=sys.stderr)
    step_ms = elapsed / steps * 1e3
    roof_step_ms = roof_elapsed / roof_steps * 1e3
    n_calls = roof_steps * calls_per_step
    roof = None
    pers_ms, _ = prof.get("persist_sweep", (0.0, 0))
    if pers_ms and counts:
        # Persistence sweep (all rounds of one call). Compulsory bytes per sweep: per saddle and
        # side 20 B (two extremum indices, saddle value in; pair, persistence out), per extremum
        # 5 B (value in, alive flag out).
        b_sweep = sum(2 * 20 * ns + 5 * (nmi + nma) for nmi, ns, nma in counts) / len(counts)
        sweep_s = pers_ms / n_calls / 1e3
        ach = b_sweep / sweep_s / 1e9
        roof = {"bound": "hbm", "kernel": "persistence sweep (k_persist_find + k_persist_apply rounds)",
                "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 5),
                "traffic": measured_traffic(name, "persistence_sweep"), "bytes_per_launch": int(b_sweep),
                "avg_launch_us": round(sweep_s * 1e6, 1), "launch_unit": "one sweep (all rounds) per call",
                "share_of_step_time": round(pers_ms / roof_steps / roof_step_ms, 3)}
        if split_roof:
            roof["measured_on"] = (f"{roof_steps} extra steps with the batch as one pipeline (pipe_groups 1, "
                                   f"{roof_step_ms:.1f} ms per step); the timed steps run it as {pipe} parts side by side")
    grad_ms, grad_n = prof.get("gradient", (0.0, 0))
    if grad_n:
        # gradient kernel: all fields of a launch (one field per plain call, the whole batch per
        # batched call): 4 B field read per pixel + 16 B gradient record per slot
        per_launch = sum(KERNEL_MODELS["gradient"](int(sh[0]), int(sh[1])) for sh, _ in sizes) * roof_steps / grad_n
        avg_s = grad_ms / grad_n / 1e3
        ach = per_launch / avg_s / 1e9
        g = {"achieved": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4), "avg_launch_us": round(avg_s * 1e6, 2),
             "bytes_per_launch": int(per_launch), "traffic": measured_traffic(name, "gradient"),
             "share_of_step_time": round(grad_ms / roof_steps / roof_step_ms, 4)}
        if roof is None:
            roof = {"bound": "hbm", "kernel": "gradient", "peak": HBM_PEAK_GBS, "unit": "GB/s", **g}
        else:
            roof["gradient_kernel"] = g

    out = {"name": name, "desc": desc, "value": value, "steps": steps, "warmup": warmup, "ms_per_step": step_ms,
           "extractions_per_step": n_ext, "pixels_per_step": px_step, "b_alg_step": b_alg_step, "e2e_gbs": e2e_gbs,
           "roofline": roof, "fields": fields, "runs": runs}
    if mixed and plain_steps > 0:  # the same fields through the plain per-field path
        step(0, plain=True)  # warmup
        torch.cuda.synchronize()
        with gc_paused():
            tp = time.perf_counter()
            for s in range(plain_steps):
                step(1 + s % (total_steps - 1), plain=True)
            torch.cuda.synchronize()
            plain = px_step * plain_steps / (time.perf_counter() - tp) / 1e6
        out["plain_loop_Mpix_s"] = plain
        out["plain_steps"] = plain_steps
    del copies
    torch.cuda.empty_cache()
    return out


def side_config(m):
    """Summary of a non-headline workload for the extra keys of the JSON line."""
    r = m["roofline"] or {}
    return {"workload": m["desc"], "Mpix_s": round(m["value"], 2), "ms_per_step": round(m["ms_per_step"], 3),
            "steps": m["steps"], "end_to_end_GBs": round(m["e2e_gbs"], 1),
            "end_to_end_frac": round(m["e2e_gbs"] / HBM_PEAK_GBS, 4),
            "persist_sweep_frac": r.get("frac"), "persist_sweep_us": r.get("avg_launch_us"),
            "gradient_frac": (r.get("gradient_kernel") or {}).get("frac"),
            "gradient_us": (r.get("gradient_kernel") or {}).get("avg_launch_us")}


def main():
    args = parse()
    if args.summarize_profile:
        summarize_profile(*args.summarize_profile)
        return
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    if args.cpu_table:
        cpu_table()
        return
    import torch

    import dmsc_amd
    from dmsc_amd.csrc import dmsc_gpu

    if args.pipe_groups is not None:
        from dmsc_amd import autotune

        autotune.apply_saved(0)  # the saved parameters first, so that they cannot override this one
        dmsc_gpu.set_param(0, "pipe_groups", args.pipe_groups)
    if args.workload == "diff":
        bench_diff(args, torch, dmsc_amd, dmsc_gpu)
        return
    headline = args.workload == "mixed"
    m = measure(args.workload, args.steps, args.warmup, plain_steps=3 if headline and not args.no_plain else 0,
                profile=args.profile)
    config = {"workload": m["desc"], "extractions_per_step": m["extractions_per_step"],
              "pixels_per_step": m["pixels_per_step"]}
    if "plain_loop_Mpix_s" in m:
        config["plain_loop_Mpix_s"] = round(m["plain_loop_Mpix_s"], 2)
        config["plain_loop_steps"] = m["plain_steps"]
        config["batched_over_plain"] = round(m["value"] / m["plain_loop_Mpix_s"], 2)
    extra = {}
    if headline and not args.headline_only:
        # the single-field configurations, a few steps each, reported beside the headline
        side_steps = max(1, min(args.steps, 5))
        for name in ("noise8192", "gauss4096"):
            extra[name] = side_config(measure(name, side_steps, 1, profile=args.profile))
    cpu = None if args.no_cpu else cpu_baseline(m["fields"], m["runs"], headline)
    line = {
        "metric": METRIC,
        "value": round(m["value"], 2),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(m["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": config,
        "end_to_end": {"algorithmic_bytes_per_step": m["b_alg_step"], "achieved_GBs": round(m["e2e_gbs"], 1),
                       "frac": round(m["e2e_gbs"] / HBM_PEAK_GBS, 4)},
        "roofline": m["roofline"],
        "cpu_baseline": cpu,
        **extra,
    }
    print(json.dumps(line))


def bench_diff(args, torch, dmsc_amd, dmsc_gpu):
    """Persistence loss sum(pers^2, 0 <= pers < 0.1) of 64 fields of 512x512 and its gradient: one
    step = the batched forward (unsimplified complex, persistence only) + the backward kernel."""
    from dmsc_amd import synthetic

    B, H, W, maxp = 64, 512, 512, 0.1
    stack = synthetic.loss_batch(B, H, W, seed=5000)  # tie-free (finite differences are checked on these fields)
    total_steps = max(args.warmup, 1) + args.steps
    copies = [stack.cuda() for _ in range(total_steps)]
    torch.cuda.synchronize()

    def step(s):
        x = copies[s].requires_grad_(True)
        loss = dmsc_amd.persistence_loss(x, maxp).sum()
        loss.backward()
        return x

    for s in range(max(args.warmup, 1)):
        step(s)
    torch.cuda.synchronize()
    dmsc_gpu.profile_enable(0, 1)
    with gc_paused():
        t0 = time.perf_counter()
        for s in range(max(args.warmup, 1), total_steps):
            step(s)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
    prof = dmsc_gpu.profile_read(0)
    dmsc_gpu.profile_enable(0, 0)
    px = B * H * W
    bw_ms, bw_n = prof.get("persistence_backward", (0.0, 0))
    # backward kernels per field: contributions zeroed + scattered (16 B per slot) and one gather
    # pass (reads the field, 4 B per pixel, writes the gradient, 4 B per pixel)
    b_bw = B * (16 * (H + 1) * (W + 1) + 8 * H * W)
    roof = None
    if bw_n:
        avg_s = bw_ms / args.steps / 1e3
        ach = b_bw / avg_s / 1e9
        roof = {"bound": "hbm", "kernel": "persistence backward (all fields of a step)", "achieved": round(ach, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                "bytes_per_launch": b_bw, "avg_launch_us": round(avg_s * 1e6, 1)}
    cpu = None if args.no_cpu else cpu_baseline_diff(stack.numpy(), maxp)
    print(json.dumps({
        "metric": "Mpix/s persistence-loss forward + backward (1 GPU)", "value": round(px * args.steps / elapsed / 1e6, 2),
        "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "64 gaussian-mixture fields 512x512 (24 bumps, sigma 8-32 px, noise 0.05, seeds "
                               "5000-5063, ties broken by ulps); loss = sum of pers^2 over pairs with 0 <= pers < 0.1, "
                               "both sides",
                   "backward_ms_per_step": round(bw_ms / args.steps, 3)},
        "roofline": roof, "cpu_baseline": cpu}))


def cpu_info():
    """(CPU model, CPUs of the host, CPUs this process may run on)."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    allowed = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return model, os.cpu_count(), allowed


def cpu_threads():
    """Oracle threads for the CPU baseline: every CPU this process may use, at most 16 (the CPU share
    of one GPU on the GPU host)."""
    return max(1, min(cpu_info()[2], 16))


def oracle_rate(cases, threads, budget_s, flags=(True, True, True, True, True)):
    """Mpix/s of the oracle over (field, thr, dual) cases taken in order until budget_s has passed:
    fields are spread over `threads` worker threads, one single-threaded oracle call each (the
    oracle releases the GIL); a single case runs with the oracle's own `threads` threads."""
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np

    import oracle

    grad, tmax, tmin, gmax, gmin = flags

    def one(c, nt):
        f, thr, dual = c
        oracle.extract(np.ascontiguousarray(f), thr, grad, dual, tmax, tmin, gmax, gmin, nthreads=nt)
        return f.shape[0] * f.shape[1]

    px, done = 0, 0
    t0 = time.perf_counter()
    if len(cases) == 1:
        px = one(cases[0], threads)
        done = 1
    else:
        with ThreadPoolExecutor(threads) as ex:
            while done < len(cases) and time.perf_counter() - t0 < budget_s:
                chunk = cases[done:done + 4 * threads]
                px += sum(ex.map(lambda c: one(c, 1), chunk))
                done += len(chunk)
    dt = time.perf_counter() - t0
    return px / dt / 1e6, done, dt


def cpu_baseline(fields, runs, mixed):
    """The oracle (C++ restatement of the reference CPU path) on a bounded sample of the workload,
    on all allotted cores and on one thread."""
    model, host_cpus, _ = cpu_info()
    cores = cpu_threads()
    if mixed:
        cases = [(fields[j].numpy(), runs[j][1], runs[j][0]) for j in range(len(fields))]
        v, n, dt = oracle_rate(cases, cores, 12.0)
        v1, n1, dt1 = oracle_rate(cases, 1, 6.0) if cores > 1 else (v, n, dt)
        what = (f"first {n} of the 512 fields with their thresholds, one field per thread ({dt:.1f} s); "
                f"1 thread: first {n1} fields ({dt1:.1f} s)")
    else:
        f = fields[0].numpy()
        thr = runs[1][1] if len(runs) > 1 else runs[0][1]
        v, _, dt = oracle_rate([(f, thr, False)], cores, 0)
        H, W = f.shape
        crop = f[H // 4: H // 4 + H // 2, W // 4: W // 4 + W // 2].copy()  # a quarter of the pixels
        v1, _, dt1 = oracle_rate([(crop, thr, False)], 1, 0)
        what = (f"one primal extraction of the {H}x{W} field at thr {thr}, {cores} oracle threads ({dt:.1f} s); "
                f"1 thread: its central {H // 2}x{W // 2} crop ({dt1:.1f} s)")
    return {"value": round(v, 3), "unit": "Mpix/s", "cores": cores, "kind": "port", "sample": what,
            "one_thread_Mpix_s": round(v1, 3), "cpu_model": model, "host_cpus": host_cpus}


def cpu_table():
    """The CPU oracle on every configuration, all allotted cores and one thread; one line per
    measurement (the figures behind the CPU baseline table of the documentation)."""
    import statistics

    import numpy as np

    from dmsc_amd import synthetic

    model, host_cpus, allowed = cpu_info()
    cores = cpu_threads()
    print(json.dumps({"cpu_model": model, "host_cpus": host_cpus, "allowed_cpus": allowed, "threads": cores}),
          flush=True)

    def emit(cfg, case, threads, mpix, runs, seconds):
        print(json.dumps({"config": cfg, "case": case, "threads": threads, "Mpix_s": round(mpix, 3), "runs": runs,
                          "seconds": round(seconds, 2)}), flush=True)

    f1 = synthetic.gaussian_mixture(256, 256, 24, (8.0, 32.0), 0.05, seed=256).numpy()
    for nt in (cores, 1):
        rates = [oracle_rate([(f1, 0.1, False)], nt, 0) for _ in range(10)]
        emit(1, "256^2 thr 0.1 primal (median of 10)", nt, statistics.median(r[0] for r in rates), 10,
             sum(r[2] for r in rates))
    f2 = synthetic.gaussian_mixture(4096, 4096, 256, (16.0, 256.0), 0.02, seed=4096).numpy()
    for dual in (False, True):
        for thr in (0.0, 0.05, 0.2):
            r = oracle_rate([(f2, thr, dual)], cores, 0)
            emit(2, f"4096^2 thr {thr} {'dual' if dual else 'primal'}", cores, r[0], 1, r[2])
    r = oracle_rate([(f2, 0.05, False)], 1, 0)
    emit(2, "4096^2 thr 0.05 primal", 1, r[0], 1, r[2])
    del f2
    f3 = synthetic.quantized_noise(8192, 8192, 256, seed=8192).numpy()
    for nt in (cores, 1):
        r = oracle_rate([(f3, 0.05, False)], nt, 0)
        emit(3, "8192^2 quantised noise thr 0.05 primal", nt, r[0], 1, r[2])
    del f3
    fields, thrs = synthetic.mixed_batch(512, seed=512)
    cases = [(f.numpy(), t, False) for f, t in zip(fields, thrs)]
    r = oracle_rate(cases, cores, 1e9)
    emit(4, "all 512 mixed fields, one field per thread", cores, r[0], r[1], r[2])
    r = oracle_rate(cases[::8], 1, 1e9)
    emit(4, "every 8th of the 512 mixed fields", 1, r[0], r[1], r[2])
    stack = synthetic.loss_batch(64).numpy()
    for nt in (cores, 1):
        v, n, dt = loss_rate(stack, 0.1, nt)
        emit(5, "64 x 512^2 loss: oracle pairs + gradient assembly", nt, v, n, dt)


def loss_rate(stack, maxp, threads):
    """Mpix/s of the CPU persistence-loss gradient: the oracle's unsimplified pairs plus a numpy
    assembly of the gradient, one field per worker thread."""
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np

    import oracle

    def one(f):
        H, W = f.shape
        out = oracle.extract(np.ascontiguousarray(f), -1.0, False, False, False, False, False, False, nthreads=1)
        g = np.zeros(H * W)
        ff = f.reshape(-1)

        def px(cells):  # highest vertex of each cell (primal, (value, pixel id) order)
            t, q = cells & 3, cells >> 2
            y, x = q // (W + 1), q % (W + 1)
            best = np.full(cells.shape, -1)
            for dy, dx, ok in ((0, 0, t == 0), (0, -1, t == 1), (0, 0, t == 1), (-1, 0, t == 2), (0, 0, t == 2),
                               (-1, -1, t == 3), (-1, 0, t == 3), (0, -1, t == 3), (0, 0, t == 3)):
                yy, xx = y + dy, x + dx
                v = ok & (yy >= 0) & (yy < H) & (xx >= 0) & (xx < W)
                cand = np.where(v, yy * W + xx, 0)
                better = v & ((best < 0) | (ff[cand] > ff[np.maximum(best, 0)]) |
                              ((ff[cand] == ff[np.maximum(best, 0)]) & (cand > best)))
                best = np.where(better, cand, best)
            return best

        sad, pmin, ppmin, mins, pmax, ppmax, maxs = (np.asarray(out[k]) for k in (3, 7, 9, 2, 6, 8, 1))
        sp = px(sad)
        for p, pp, ext, sgn in ((pmin, ppmin, mins, 1.0), (pmax, ppmax, maxs, -1.0)):
            m = (pp >= 0) & (p >= 0) & (p < maxp)
            np.add.at(g, sp[m], sgn * 2 * p[m])
            np.add.at(g, px(ext[pp[m]]), -sgn * 2 * p[m])
        return H * W

    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        px_total = sum(ex.map(one, list(stack)))
    dt = time.perf_counter() - t0
    return px_total / dt / 1e6, len(stack), dt


def cpu_baseline_diff(stack, maxp):
    """The CPU persistence-loss gradient on all allotted cores (all 64 fields) and on one thread
    (8 fields)."""
    model, host_cpus, _ = cpu_info()
    cores = cpu_threads()
    v, n, dt = loss_rate(stack, maxp, cores)
    v1, n1, dt1 = loss_rate(stack[:8], maxp, 1)
    return {"value": round(v, 3), "unit": "Mpix/s", "cores": cores, "kind": "port",
            "sample": (f"{n} of the 512x512 fields, oracle pairs + numpy gradient, one field per thread ({dt:.1f} s); "
                       f"1 thread: {n1} fields ({dt1:.1f} s)"),
            "one_thread_Mpix_s": round(v1, 3), "cpu_model": model, "host_cpus": host_cpus}


if __name__ == "__main__":
    main()
