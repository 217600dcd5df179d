"""Benchmark of the MI355X compute_dmsc path.

Default workload (one GPU): a 4096x4096 float32 Gaussian-mixture field (256 bumps, sigma 16-256 px,
noise 0.02, seeded), full DMSC (gradient returned, ridges, valleys, peaks, basins) for primal and
dual at persistence thresholds {0, 0.05, 0.2}. One step = those 6 extractions, each on its own
device copy of the field (no cache reuse between extractions). Metric: Mpix/s = pixels processed /
wall time of the timed steps.

Prints one JSON line with the metric, the roofline object of the dominant kernel (algorithmic
bytes / average launch time from HIP events recorded around its launches during the timed steps),
the end-to-end algorithmic HBM rate, and the CPU oracle timed on a bounded sample of the same
workload on this host.

Other workloads (--workload): noise8192 (8192^2 uniform noise quantised to 256 levels, thr 0.05),
mixed (512 fields of sizes 64..1024 with per-field thresholds: one batched call per step, with the
per-field loop of the plain path timed beside it for the batched/plain ratio), diff (persistence
loss over 64 fields of 512x512: batched forward + hand-written backward per step).
"""

import argparse
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
    p.add_argument("--workload", default="gauss4096", choices=["gauss4096", "noise8192", "mixed", "diff"])
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    p.add_argument("--profile", action="store_true", help="print per-kernel event timings to stderr")
    p.add_argument("--copy-calibration", action="store_true",
                   help="only run plain 256 MiB device-to-device copies (reference kernel for PMC byte counters)")
    p.add_argument("--summarize-profile", nargs=5, metavar=("TRACE_CSV", "FETCH_CSV", "WRITE_CSV", "COPY_FETCH_CSV",
                                                            "COPY_WRITE_CSV"),
                   help="print a per-kernel table (duration, calibrated HBM bytes, GB/s vs peak) from rocprofv3 "
                        "kernel-trace and FETCH_SIZE / WRITE_SIZE counter CSVs of the same command")
    return p.parse_args()


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
    desc = "512 gaussian-mixture fields, H,W in {64..1024 step 32}, 8-64 bumps, thr U[0.02,0.2], seed 512"
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
    import torch

    import dmsc_amd
    from dmsc_amd.csrc import dmsc_gpu

    dev = torch.device("cuda:0")
    if args.workload == "diff":
        bench_diff(args, torch, dmsc_amd, dmsc_gpu)
        return
    fields, runs, desc = workload(args.workload)
    mixed = args.workload == "mixed"
    n_ext = len(runs)  # extractions per step

    def field_of(j):
        return fields[j] if mixed else fields[0]

    # distinct device copies per extraction and step (no cache reuse across extractions)
    total_steps = args.warmup + args.steps
    copies = [[field_of(j).to(dev) if (mixed or s == 0) else None for j in range(n_ext)] for s in range(total_steps)]
    if not mixed:
        for s in range(total_steps):
            copies[s] = [fields[0].to(dev) for _ in range(n_ext)]
    torch.cuda.synchronize()

    def step(s, plain=False):
        if mixed and not plain:  # one batched call over all fields (per-field thresholds)
            return [dmsc_amd.compute_dmsc_batch(copies[s], [t for _, t in runs], return_gradient=True)]
        return [dmsc_amd.compute_dmsc(copies[s][j], thr, return_gradient=True, is_dual=dual)
                for j, (dual, thr) in enumerate(runs)]

    # bookkeeping (output sizes, critical-cell counts per call) on the first warmup step, untimed
    sizes = []   # (shape, output bytes) per extracted field
    counts = []  # (minima, saddles, maxima) per call (a batched call counts all its fields)
    for s in range(max(args.warmup, 1)):
        if mixed and not s:
            res = dmsc_amd.compute_dmsc_batch(copies[s], [t for _, t in runs], return_gradient=True)
            sizes = [(copies[s][j].shape, out_bytes(ms)) for j, ms in enumerate(res)]
            counts.append([int(v) for v in dmsc_gpu.debug_fetch(0, "counts", 3).tolist()])
        elif not s:
            for j, (dual, thr) in enumerate(runs):
                ms = dmsc_amd.compute_dmsc(copies[s][j], thr, return_gradient=True, is_dual=dual)
                sizes.append((copies[s][j].shape, out_bytes(ms)))
                counts.append([int(v) for v in dmsc_gpu.debug_fetch(0, "counts", 3).tolist()])
        else:
            step(s)
    torch.cuda.synchronize()
    calls_per_step = len(counts)

    # level 1: events only around the gradient launches and each batch of persistence rounds
    dmsc_gpu.profile_enable(0, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.warmup, total_steps):
        step(s)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = dmsc_gpu.profile_read(0)
    dmsc_gpu.profile_enable(0, 0)

    elapsed = t1 - t0
    px_step = sum(int(sh[0]) * int(sh[1]) for sh, _ in sizes)
    value = px_step * args.steps / elapsed / 1e6
    b_alg_step = sum(4 * int(sh[0]) * int(sh[1]) + ob for sh, ob in sizes)
    e2e_gbs = b_alg_step * args.steps / elapsed / 1e9

    if args.profile:  # every stage timed, on one extra untimed step (its events slow the round loops)
        dmsc_gpu.profile_enable(0, 2)
        step(args.warmup)
        torch.cuda.synchronize()
        full = dmsc_gpu.profile_read(0)
        dmsc_gpu.profile_enable(0, 0)
        for k, (ms_tot, n) in sorted(full.items(), key=lambda kv: -kv[1][0]):
            print(f"  {k:20s} total {ms_tot:9.3f} ms  launches {n:6d}  avg {ms_tot / max(n, 1) * 1e3:9.1f} us",
                  file=sys.stderr)
    step_ms = elapsed * 1e3
    n_calls = args.steps * calls_per_step
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
                "traffic": measured_traffic(args.workload, "persistence_sweep"), "bytes_per_launch": int(b_sweep),
                "avg_launch_us": round(sweep_s * 1e6, 1),
                "launch_unit": "one sweep (all rounds) per call",
                "share_of_step_time": round(pers_ms / step_ms, 3)}
    grad_ms, grad_n = prof.get("gradient", (0.0, 0))
    if grad_n:
        # gradient kernel: all fields of a launch (one field per plain call, the whole batch per
        # batched call) plus their gradient output
        per_launch = sum(KERNEL_MODELS["gradient"](int(sh[0]), int(sh[1])) for sh, _ in sizes) * args.steps / grad_n
        avg_s = grad_ms / grad_n / 1e3
        ach = per_launch / avg_s / 1e9
        measured = measured_traffic(args.workload, "gradient")
        g = {"achieved": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4), "avg_launch_us": round(avg_s * 1e6, 2),
             "bytes_per_launch": int(per_launch), "traffic": measured, "share_of_step_time": round(grad_ms / step_ms, 4)}
        if roof is None:
            roof = {"bound": "hbm", "kernel": "gradient", "peak": HBM_PEAK_GBS, "unit": "GB/s", **g}
        else:
            roof["gradient_kernel"] = g

    extra = {}
    if mixed:  # the same fields through the plain per-field path, for the batched-vs-plain ratio
        torch.cuda.synchronize()
        tp = time.perf_counter()
        step(args.warmup, plain=True)
        torch.cuda.synchronize()
        plain = px_step / (time.perf_counter() - tp) / 1e6
        extra = {"plain_loop_Mpix_s": round(plain, 2), "batched_over_plain": round(value / plain, 2)}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(fields, runs, mixed)

    line = {
        "metric": "Mpix/s full DMSC (1 GPU) + achieved HBM GB/s, bit-exact",
        "value": round(value, 2),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": desc, "extractions_per_step": n_ext, "pixels_per_step": px_step, **extra},
        "end_to_end": {"algorithmic_bytes_per_step": b_alg_step, "achieved_GBs": round(e2e_gbs, 1),
                       "frac": round(e2e_gbs / HBM_PEAK_GBS, 4)},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


def bench_diff(args, torch, dmsc_amd, dmsc_gpu):
    """Persistence loss sum(pers^2, 0 <= pers < 0.1) of 64 fields of 512x512 and its gradient: one
    step = the batched forward (unsimplified complex, persistence only) + the backward kernel."""
    from dmsc_amd import synthetic

    B, H, W, maxp = 64, 512, 512, 0.1
    stack = torch.stack([synthetic.gaussian_mixture(H, W, 24, (8.0, 32.0), 0.05, seed=5000 + b) for b in range(B)])
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
    cpu = None if args.no_cpu else cpu_baseline_diff(stack[:2].numpy(), maxp)
    print(json.dumps({
        "metric": "Mpix/s persistence-loss forward + backward (1 GPU)", "value": round(px * args.steps / elapsed / 1e6, 2),
        "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "64 gaussian-mixture fields 512x512 (24 bumps, sigma 8-32 px, noise 0.05, seeds "
                               "5000-5063); loss = sum of pers^2 over pairs with 0 <= pers < 0.1, both sides",
                   "backward_ms_per_step": round(bw_ms / args.steps, 3)},
        "roofline": roof, "cpu_baseline": cpu}))


def cpu_baseline_diff(fields, maxp):
    """The oracle's unsimplified pairs plus a numpy gradient assembly, on a sample of fields."""
    import numpy as np

    import oracle

    cores = min(os.cpu_count() or 1, 16)
    t0 = time.perf_counter()
    for f in fields:
        H, W = f.shape
        out = oracle.extract(np.ascontiguousarray(f), -1.0, False, False, False, False, False, False, nthreads=cores)
        g = np.zeros(H * W)
        ff = f.reshape(-1)

        def px(cells):  # highest vertex of each cell (primal)
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
    dt = time.perf_counter() - t0
    px_total = sum(f.size for f in fields)
    return {"value": round(px_total / dt / 1e6, 3), "unit": "Mpix/s", "cores": cores, "kind": "port",
            "sample": f"{len(fields)} of the 512x512 fields: oracle pairs + numpy gradient ({dt:.1f} s)"}


def cpu_baseline(fields, runs, mixed):
    """The oracle (C++ restatement of the reference CPU path) on a bounded sample of the workload."""
    import numpy as np

    import oracle

    cores = min(os.cpu_count() or 1, 16)
    if mixed:
        sample = [(fields[j].numpy(), runs[j][1], runs[j][0]) for j in range(16)]
        what = "first 16 fields of the mixed batch, primal, their thresholds"
    else:
        sample = [(fields[0].numpy(), runs[1][1] if len(runs) > 1 else runs[0][1], False)]
        what = f"one primal extraction of the {fields[0].shape[0]}x{fields[0].shape[1]} field at thr {sample[0][1]}"
    px = 0
    t0 = time.perf_counter()
    for f, thr, dual in sample:
        oracle.extract(np.ascontiguousarray(f), thr, True, dual, True, True, True, True, nthreads=cores)
        px += f.shape[0] * f.shape[1]
    dt = time.perf_counter() - t0
    return {"value": round(px / dt / 1e6, 3), "unit": "Mpix/s", "cores": cores, "kind": "port",
            "sample": what + f" ({dt:.1f} s)"}


if __name__ == "__main__":
    main()
