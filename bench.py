"""Benchmark of the xc_fast hot path on one MI355X.

Workload (the BASELINE precision-form configuration): 1000 stations, Z only,
one day at dt = 0.1 s cut into 24 one-hour windows, output FFT 37044
(nspec = 18523 bins), all 499500 station pairs, max_lag 100 s (2001 lags).
Spectra are synthetic (seeded complex Gaussian, the shape a StepPack payload
holds) and resident in HBM before timing starts.  One step = the full
windowed stack + lag synthesis of every pair (the work xc_fast does for one
timestamp).  Three arithmetic forms are timed: split-bf16 MFMA (headline),
the tiled fp32 contraction, and "plain" -- the reference's own step order
(the stack row of every pair read, updated and written once per window, fp32)
run on the same GPU.  Unit: pair-windows per second (pairs x windows / step
time).  The CPU baseline is the multi-threaded C restatement in the oracle,
timed on a bounded station subset.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision ...]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "pair-windows/s at 1000 sta, FP32 vs split-BF16 MFMA stack; p99 stream latency"
HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TFLOPS = 2500.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nsta", type=int, default=1000)
    ap.add_argument("--nstep", type=int, default=24)
    ap.add_argument("--nspec", type=int, default=18523)
    ap.add_argument("--half-cc", type=int, default=1000)
    ap.add_argument("--precision", default="bf16x3,fp32,plain",
                    help="comma list; the first is the headline value")
    ap.add_argument("--chunk-rows", type=int, default=0)
    ap.add_argument("--cpu-sample-sta", type=int, default=200)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: the cores this process may use, at most 16")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy-gib", type=float, default=0.0,
                    help="first run one device-to-device copy of this size (HBM counter calibration)")
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "xc_hbm_traffic_1000sta.txt"))
    return ap.parse_args()


def run_mode(precision, spec, src, rec, args, out):
    import torch
    from fastxc_amd.ops import XcPlan

    plan = XcPlan(args.nspec, args.nstep, 0.1, args.half_cc, precision, chunk_rows=args.chunk_rows)
    plan.set_pairs(args.nsta, src, rec)
    for _ in range(args.warmup):
        plan.correlate(spec, out)
    torch.cuda.synchronize()
    plan.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.correlate(spec, out)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    tm = plan.get_timing()
    info = plan.pairs_info()
    plan.close()
    npairs = len(src)
    sample = torch.linspace(0, npairs - 1, 257, device=out.device).round().long()
    rows = out[sample].double().cpu()
    ms = wall * 1e3 / args.steps
    return {
        "precision": precision,
        "ms_per_step": ms,
        "pair_windows_per_s": npairs * args.nstep / (ms / 1e3),
        "contract_ms": tm["contract_ms"] / args.steps,
        "lag_ms": tm["lag_ms"] / args.steps,
        "split_ms": tm["split_ms"] / args.steps,
        "contract_launches_per_step": tm["contract_launches"] / args.steps,
        "lag_launches_per_step": tm["lag_launches"] / args.steps,
        "chunks": info["nchunks"],
        "_rows": rows,
    }


def algorithmic_bytes(args, npairs):
    """Path-level algorithmic HBM bytes of one step: every input spectrum read
    once (8 B x nspec x windows x stations) and every correlation written once
    (4 B x lags x pairs) -- 7.56 GB, about 630 B per pair-window, at 1000
    stations x 24 windows x 18523 bins x 2001 lags."""
    return args.nsta * args.nstep * args.nspec * 8 + npairs * (2 * args.half_cc + 1) * 4


def read_traffic(path, precision, kernel):
    """HBM bytes per launch of `kernel` measured with rocprofv3 FETCH_SIZE and
    WRITE_SIZE (each scaled by a plain copy kernel's reading), from the
    committed summary; None when absent or for another workload."""
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 4 and parts[0] == "traffic" and parts[1] == precision \
                        and parts[2] == kernel:
                    return int(float(parts[3]))
    except OSError:
        pass
    return None


def roofline(mode, args, npairs, default_workload):
    """Dominant kernel of the headline form against the HBM roofline.
    achieved = algorithmic bytes of the units one launch processes / the
    launch's average duration (HIP events on the launch stream)."""
    if mode["contract_ms"] >= mode["lag_ms"]:
        name = {"fp32": "contract_fp32_kernel", "plain": "plain_accum_kernel"}.get(
            mode["precision"], "contract_mfma_kernel")
        t = mode["contract_ms"] / max(mode["contract_launches_per_step"], 1)
        launches = max(mode["contract_launches_per_step"], 1)
    else:
        # nfft 37044 runs the compile-time-plan lag kernel
        name = "lag_kernel_fixed" if args.nspec == 18523 else "lag_kernel"
        t = mode["lag_ms"] / max(mode["lag_launches_per_step"], 1)
        launches = max(mode["lag_launches_per_step"], 1)
    byts = algorithmic_bytes(args, npairs) / launches
    achieved = byts / (t / 1e3) / 1e9
    traffic = read_traffic(args.traffic_file, mode["precision"], name) if default_workload else None
    return {"kernel": name, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "launch_ms": round(t, 3), "algorithmic_bytes_per_launch": int(byts)}


def cpu_threads(args):
    if args.cpu_threads > 0:
        return args.cpu_threads
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_baseline(args):
    """The oracle's multi-threaded C restatement (oracle/cpu_xc.c: FP32 window
    stack in the reference's order + its own mixed-radix FFT), timed on a
    bounded subset of stations with the same window count, bins and lags."""
    import numpy as np
    from oracle import cbind

    ns = args.cpu_sample_sta
    th = cpu_threads(args)
    rng = np.random.default_rng(5)
    spec = (rng.standard_normal((args.nstep, ns, args.nspec), dtype=np.float32) +
            1j * rng.standard_normal((args.nstep, ns, args.nspec), dtype=np.float32)
            ).astype(np.complex64)
    src, rec = np.triu_indices(ns, k=1)
    blk = 128 * th
    done = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < args.cpu_seconds:
        i = done % len(src)
        j = min(i + blk, len(src))
        cbind.cpu_xc_correlate(spec, src[i:j], rec[i:j], args.half_cc, 0.1, nthreads=th)
        done += j - i
    dt = time.perf_counter() - t0
    return {"value": round(done * args.nstep / dt, 1), "unit": "pair-windows/s", "cores": th,
            "kind": "port",
            "sample": f"{done} pairs drawn from the {len(src)} pairs of a {ns}-station subset, "
                      f"x {args.nstep} windows, nspec {args.nspec}, {2 * args.half_cc + 1} lags, "
                      f"{dt:.1f} s with {th} threads (oracle/cpu_xc.c, FP32)"}


def calibrate_copy(gib):
    """One device-to-device copy of `gib` GiB: the reference reading for the
    HBM byte counters (FETCH_SIZE / WRITE_SIZE) of a kernel of known traffic."""
    import torch

    n = int(gib * (1 << 30)) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b
    torch.cuda.empty_cache()


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only (this build has no multi-GPU path); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    if args.calibrate_copy_gib > 0:
        calibrate_copy(args.calibrate_copy_gib)
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(1234)
    spec = torch.randn((args.nstep, args.nsta, args.nspec), dtype=torch.complex64, device=dev,
                       generator=g)
    src, rec = np.triu_indices(args.nsta, k=1)
    src = src.astype(np.int32)
    rec = rec.astype(np.int32)
    out = torch.empty((len(src), 2 * args.half_cc + 1), dtype=torch.float32, device=dev)
    modes = [run_mode(p, spec, src, rec, args, out) for p in args.precision.split(",")]
    # every form is checked against the tiled fp32 form on 257 rows spread over
    # all pairs (first and last included): a cheap guard against a form that
    # silently skips work
    ref_rows = next((m["_rows"] for m in modes if m["precision"] == "fp32"), None)
    for m in modes:
        rows = m.pop("_rows")
        if ref_rows is not None:
            m["rel_l2_vs_fp32_rows"] = float(((rows - ref_rows).norm() / ref_rows.norm()).item())
    default_workload = (args.nsta, args.nstep, args.nspec, args.half_cc, args.chunk_rows) == \
        (1000, 24, 18523, 1000, 0)
    head = modes[0]
    res = {
        "metric": METRIC,
        "value": round(head["pair_windows_per_s"], 1),
        "unit": "pair-windows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"bf16x3": "bf16x3-split/f32-acc", "fp32": "f32", "fp16": "f16/f32-acc",
                  "plain": "f32"}[head["precision"]],
        "data": "synthetic",
        "config": {"workload": f"xc_fast one timestamp: {args.nsta} stations x {args.nstep} windows, "
                               f"nspec {args.nspec} (nfft {2 * (args.nspec - 1)}), {len(src)} pairs, "
                               f"{2 * args.half_cc + 1} lags",
                   "stations": args.nsta, "windows": args.nstep, "nspec": args.nspec,
                   "pairs": len(src), "lags": 2 * args.half_cc + 1, "chunk_rows": args.chunk_rows},
        "modes": modes,
        "roofline": roofline(head, args, len(src), default_workload),
    }
    for m in modes[1:]:
        res["ratio_vs_" + m["precision"]] = round(head["pair_windows_per_s"] /
                                                   m["pair_windows_per_s"], 3)
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
