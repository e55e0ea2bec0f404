"""Benchmark of the xc_fast hot path on one MI355X.

Workload (the BASELINE precision-form configuration): 1000 stations, Z only,
one day at dt = 0.1 s cut into 24 one-hour windows, output FFT 37044
(nspec = 18523 bins), all 499500 station pairs, max_lag 100 s (2001 lags).
Spectra are synthetic (seeded complex Gaussian, the shape a StepPack payload
holds) and resident in HBM before timing starts.  One step = the full
windowed stack + lag synthesis of every pair (the work xc_fast does for one
timestamp).  Two arithmetic forms are timed: split-bf16 MFMA (headline) and
plain fp32.  Unit: pair-windows per second (pairs x windows / step time).

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
    ap.add_argument("--precision", default="bf16x3,fp32",
                    help="comma list; the first is the headline value")
    ap.add_argument("--chunk-rows", type=int, default=0)
    ap.add_argument("--cpu-sample-sta", type=int, default=48)
    ap.add_argument("--no-cpu", action="store_true")
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
    }


def roofline(mode, args, npairs):
    """Dominant kernel vs its bound.  Algorithmic bytes per pair:
    contraction writes one cross-spectrum row (8*nspec) and reads the split
    spectra once per launch; lag synthesis reads the row and writes 4*cc."""
    nspec, cc = args.nspec, 2 * args.half_cc + 1
    if mode["contract_ms"] >= mode["lag_ms"]:
        split_bytes = args.nsta * nspec * 2 * ((args.nstep + 7) // 8 * 8) * 2
        split_bytes *= 2 if mode["precision"] == "bf16x3" else (1 if mode["precision"] == "fp16" else 0)
        spec_bytes = args.nsta * nspec * 8 * args.nstep if mode["precision"] == "fp32" else split_bytes
        byts = npairs * nspec * 8 + spec_bytes
        t = mode["contract_ms"] / max(mode["contract_launches_per_step"], 1)
        byts /= max(mode["contract_launches_per_step"], 1)
        name = "contract_" + mode["precision"]
    else:
        byts = npairs * (nspec * 8 + cc * 4) / max(mode["lag_launches_per_step"], 1)
        t = mode["lag_ms"] / max(mode["lag_launches_per_step"], 1)
        name = "lag_kernel"
    achieved = byts / (t / 1e3) / 1e9
    return {"kernel": name, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None}


def cpu_baseline(args):
    """FP64 NumPy oracle on a bounded station sample (single thread)."""
    import numpy as np
    from oracle import xc

    ns = args.cpu_sample_sta
    rng = np.random.default_rng(5)
    spec = (rng.standard_normal((args.nstep, ns, args.nspec)) +
            1j * rng.standard_normal((args.nstep, ns, args.nspec))).astype(np.complex64)
    src, rec = np.triu_indices(ns, k=1)
    nfft = 2 * (args.nspec - 1)
    t0 = time.perf_counter()
    done = 0
    blk = 64
    for i in range(0, len(src), blk):
        xc.correlate(spec, src[i:i + blk], rec[i:i + blk], nfft, args.half_cc, 0.1)
        done += len(src[i:i + blk])
        if time.perf_counter() - t0 > 20.0:
            break
    dt = time.perf_counter() - t0
    return {"value": round(done * args.nstep / dt, 1), "unit": "pair-windows/s", "cores": 1,
            "kind": "port",
            "sample": f"{done} pairs of a {ns}-station subset x {args.nstep} windows, "
                      f"nspec {args.nspec}, FP64 numpy oracle (oracle/xc.py)"}


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only (this build has no multi-GPU path); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

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
        "dtype": {"bf16x3": "bf16x3-split/f32-acc", "fp32": "f32", "fp16": "f16/f32-acc"}[head["precision"]],
        "data": "synthetic",
        "config": {"workload": f"xc_fast one timestamp: {args.nsta} stations x {args.nstep} windows, "
                               f"nspec {args.nspec} (nfft {2 * (args.nspec - 1)}), {len(src)} pairs, "
                               f"{2 * args.half_cc + 1} lags",
                   "stations": args.nsta, "windows": args.nstep, "nspec": args.nspec,
                   "pairs": len(src), "lags": 2 * args.half_cc + 1, "chunk_rows": args.chunk_rows},
        "modes": modes,
        "roofline": roofline(head, args, len(src)),
    }
    if len(modes) > 1:
        res["ratio_first_over_second"] = round(modes[0]["pair_windows_per_s"] /
                                               modes[1]["pair_windows_per_s"], 3)
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
