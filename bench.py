"""Benchmark of the xc_fast hot path on one MI355X.

Workload (the precision-form configuration of the metric): 1000 stations, Z only,
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

The other stages of the path are measured after xc, each at the scale of
the same 1000-station day, under "stages" in the same JSON line: sac2spec (1000 stations x 24
windows), PWS and the float64 linear stack (49950 pairs x 30 timestamps),
TF-PWS (2000 pairs), the ENZ -> RTZ rotation (49950 pairs x 9 components);
--no-stages skips them, --stages-only runs only them.
The streaming loop (one-hour segments of random 10-500-station subnetworks,
Poisson arrivals at a sweep of rates) reports p50/p99 segment-update latency
and pair-windows/s under "stream"; --no-stream skips it.  The SourcePack
indexer (host C++, bin/sourcepack_index) is timed on the xcpack TSV of the
same timestamp (499500 records) under "sourcepack", beside the Python
restatement of the reference builder on a 50000-record sample;
--no-sourcepack skips it.  The inventory path planner's all-pairs loop (host
C++) is timed on 2000 stations under "planner", beside the oracle's Python
loop on 200 of them; --no-planner skips it.  The batched uneven-segment form
(64 segments of 20-400 stations and 6-48 windows, 10 % gaps) is timed eager
and as a replayed HIP graph under "batch"; --no-batch skips it.

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
    ap.add_argument("--precision", default="bf16x3,fp16x3,fp16,fp32,plain",
                    help="comma list; the first is the headline value")
    ap.add_argument("--check-pairs", type=int, default=300,
                    help="pairs of every form checked against an FP64 evaluation")
    ap.add_argument("--chunk-rows", type=int, default=0)
    ap.add_argument("--cpu-sample-sta", type=int, default=200)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: the cores this process may use, at most 16")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy-gib", type=float, default=0.0,
                    help="first run one device-to-device copy of this size (HBM counter calibration)")
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "xc_hbm_traffic_1000sta.txt"))
    ap.add_argument("--no-stages", action="store_true",
                    help="skip the sac2spec / stack stage measurements")
    ap.add_argument("--stages-only", action="store_true", help="measure the stages, not xc")
    ap.add_argument("--stage", default="",
                    help="comma list of stages to measure (sac2spec,pws,tfpws,linear_stack,"
                         "rotate); default all")
    ap.add_argument("--no-stream", action="store_true",
                    help="skip the streaming segment-update measurement")
    ap.add_argument("--stream-only", action="store_true", help="measure the streaming loop only")
    ap.add_argument("--no-sourcepack", action="store_true",
                    help="skip the SourcePack indexer measurement")
    ap.add_argument("--xc-file-to-file", action="store_true",
                    help="time bin/xc_fast file to file at 1000 stations x 24 windows only")
    ap.add_argument("--no-band", action="store_true",
                    help="skip the band-limited spec-mode lines (fxc_xc_set_band)")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched uneven-segment measurement")
    ap.add_argument("--batch-only", action="store_true",
                    help="measure the batched uneven-segment form only")
    ap.add_argument("--no-planner", action="store_true",
                    help="skip the inventory path-planner measurement")
    ap.add_argument("--stream-rates", default="160,192,224",
                    help="Poisson arrival rates (segments/s) of the streaming sweep")
    ap.add_argument("--stream-seconds", type=float, default=30.0,
                    help="arrival time per rate of the streaming sweep")
    return ap.parse_args()


def sample_pairs(nsta, count, tile=32, seed=5):
    """Row indices (into the i < j row-major pair list) checked against FP64:
    the first and last pairs, pairs inside a diagonal and an off-diagonal
    32-station tile and near the end of the array, then random pairs."""
    import numpy as np

    rng = np.random.default_rng(seed)
    pick = {(0, 1), (nsta - 2, nsta - 1), (0, nsta - 1)}
    for a in range(3):
        for b in range(a + 1, 4):
            for p in ((tile + a, tile + b), (a, 5 * tile + b), (nsta - 40 + a, nsta - 5 + b)):
                if 0 <= p[0] < p[1] < nsta:
                    pick.add(p)
    while len(pick) < min(count, nsta * (nsta - 1) // 2):
        i, j = sorted(rng.choice(nsta, 2, replace=False))
        pick.add((int(i), int(j)))
    pairs = sorted(pick)
    row = lambda i, j: i * nsta - i * (i + 1) // 2 + (j - i - 1)
    return pairs, [row(i, j) for i, j in pairs]


def fp64_rows(spec, pairs, half_cc, dt=0.1):
    """FP64 host evaluation of the sampled pairs, in the reference's step
    order: (1/nstep) sum_w conj(S_src) S_rec, x(-1)^k, inverse real FFT of
    length 2(nspec-1), x 1/(nfft dt), lags [nfft/2-H, nfft/2+H]."""
    import numpy as np

    stations = sorted({s for p in pairs for s in p})
    loc = {s: i for i, s in enumerate(stations)}
    sub = spec[:, stations, :].cpu().numpy().astype(np.complex128)
    nstep, _, nspec = sub.shape
    nfft = 2 * (nspec - 1)
    sign = np.where(np.arange(nspec) % 2 == 0, 1.0, -1.0)
    out = np.empty((len(pairs), 2 * half_cc + 1))
    for r, (i, j) in enumerate(pairs):
        X = (np.conj(sub[:, loc[i]]) * sub[:, loc[j]]).mean(axis=0) * sign
        y = np.fft.irfft(X, n=nfft) * nfft / (float(np.float32(nfft)) * dt)
        out[r] = y[nfft // 2 - half_cc: nfft // 2 + half_cc + 1]
    return out


def run_mode(precision, spec, src, rec, args, out, check_rows, band=None):
    import torch
    from fastxc_amd.ops import XcPlan

    plan = XcPlan(args.nspec, args.nstep, 0.1, args.half_cc, precision, chunk_rows=args.chunk_rows)
    plan.set_pairs(args.nsta, src, rec)
    if band is not None:
        plan.set_band(*band)
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
    rows = out[torch.tensor(check_rows, device=out.device)].double().cpu().numpy()
    ms = wall * 1e3 / args.steps
    return {
        "precision": precision if band is None else f"{precision} band {band[0]}-{band[1]}",
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


def band_measure(args, src, rec, out, check_pairs, check_rows):
    """Spec mode fxc_xc_set_band beside parity, on the spectra sac2spec makes
    of seeded noise (1000 stations x 1 day at 0.1 s, a common component plus
    independent noise, RUN-ABS, whitening after on 0.1-2 Hz: the whitening
    taper spans bins [240, 9597) of the 36000-sample window, 0.067-2.67 Hz).
    Bands: the taper's support on the output grid (nfft 37044) with 16 bins of
    margin, and the flat 0.1-2 Hz band (bins 370-7409, ~7040 retained).  Each
    line carries its FP64 error over the sampled pairs, the FP64 reference
    taking every bin."""
    import numpy as np
    import torch
    from fastxc_amd.ops import S2sPlan

    nsta, npts, seg, nfft_f, nfft_o, nstep = args.nsta, 864000, 36000, 36288, 37044, args.nstep
    if (args.nspec, nstep) != (nfft_o // 2 + 1, 24):
        return {"skipped": "band spec mode is measured at nspec 18523 x 24 windows"}
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(99)
    x = torch.randn((nsta, npts), dtype=torch.float32, device=dev, generator=g)
    x += torch.randn((1, npts), dtype=torch.float32, device=dev, generator=g)
    resp = _butter_power(nfft_f, 0.1, 0.1, 2.0)[None]
    idx = [240, 360, 7200, 9597]
    plan = S2sPlan(npts, seg, seg, nfft_f, nfft_o, 1, 0.1, 0.1, idx, resp, [0.1], whiten=2,
                   normalize=3)
    spec = torch.empty((nstep, nsta, nfft_o // 2 + 1), dtype=torch.complex64, device=dev)
    plan.process(x, np.arange(nstep, dtype=np.int32), spec)
    plan.close()
    del x
    torch.cuda.empty_cache()
    scale = nfft_o / seg
    taper = (int(idx[0] * scale) - 16, min(int(np.ceil(idx[3] * scale)) + 16, args.nspec - 1))
    flat = (int(round(0.1 * nfft_o * 0.1)), int(round(2.0 * nfft_o * 0.1)))
    lines = [run_mode("bf16x3", spec, src, rec, args, out, check_rows),
             run_mode("bf16x3", spec, src, rec, args, out, check_rows, band=taper),
             run_mode("bf16x3", spec, src, rec, args, out, check_rows, band=flat)]
    want = fp64_rows(spec, check_pairs, args.half_cc)
    for m, b in zip(lines, (None, taper, flat)):
        rows = m.pop("_rows")
        d = rows - want
        m["rel_l2_vs_fp64"] = float(np.linalg.norm(d) / np.linalg.norm(want))
        m["max_abs_vs_fp64"] = float(np.abs(d).max())
        m["retained_bins"] = args.nspec if b is None else b[1] - b[0] + 1
    del spec
    torch.cuda.empty_cache()
    return {"spectra": "sac2spec of seeded noise (RUN-ABS, whitening after, 0.1-2 Hz)",
            "lines": lines}


def xc_file_to_file(nsta=1000, nstep=24, nspec=18523, half_cc=1000, precision="bf16x3"):
    """bin/xc_fast file to file on one timestamp: a StepPack v3 batch of nsta
    stations x nstep windows x nspec bins (seeded complex Gaussian, 3.56 GB at
    the default size) and an allowed-path table of every pair i < j in, the
    xcpack parts (SAC records + TSV sidecars) out, on the box's scratch disk.
    The wall time of the executable covers the StepPack read, the upload, the
    correlation, the result download and the xcpack writes."""
    import shutil
    import subprocess
    import tempfile

    import numpy as np
    from tests import fxformats as ff

    exe = os.path.join(ROOT, "bin", "xc_fast")
    if not os.path.exists(exe):
        return {"skipped": "bin/xc_fast not built"}
    tmp = tempfile.mkdtemp(prefix="fxc_f2f_")
    try:
        rng = np.random.default_rng(21)
        spec = np.empty((nstep, nsta, nspec), dtype=np.complex64)
        for w in range(nstep):
            spec[w].real = rng.standard_normal((nsta, nspec), dtype=np.float32)
            spec[w].imag = rng.standard_normal((nsta, nspec), dtype=np.float32)
        nodes = [dict(timestamp="2020.123.0000", nsl_id="XX.S%04d." % i, gnsl_id=i + 1,
                      stla=30.0 + i * 0.01, stlo=100.0 - i * 0.01, network="XX",
                      station="S%04d" % i, location="", component="BHZ") for i in range(nsta)]
        t0 = time.perf_counter()
        sp_dir = os.path.join(tmp, "stepack")
        ff.write_stepack(sp_dir, 0, 0, nodes, spec, 0.1, 1.0 / (2 * (nspec - 1) * 0.1),
                         start_group=0, group_count=nsta)
        del spec
        recs = [(a, b, 0.01, 1.0, 10.0, 20.0) for a in range(1, nsta + 1)
                for b in range(a + 1, nsta + 1)]
        paths = os.path.join(tmp, "allowed_paths.tsv")
        ff.write_allowed_paths(paths, recs)
        prep = time.perf_counter() - t0
        out = os.path.join(tmp, "ncf")
        t0 = time.perf_counter()
        r = subprocess.run([exe, "-I", sp_dir, "-P", paths, "-O", out, "-C",
                            str(half_cc * 0.1), "--precision", precision],
                           capture_output=True, text=True, timeout=900)
        wall = time.perf_counter() - t0
        if r.returncode != 0:
            return {"error": r.stderr[-500:]}
        xdir = os.path.join(out, "xcpack")
        nbytes = sum(os.path.getsize(os.path.join(xdir, f)) for f in os.listdir(xdir))
        npairs = len(recs)
        return {"stations": nsta, "windows": nstep, "nspec": nspec, "pairs": npairs,
                "precision": precision, "wall_s": round(wall, 3),
                "pair_windows_per_s": round(npairs * nstep / wall, 1),
                "records_per_s": round(npairs / wall, 1),
                "stepack_bytes": os.path.getsize(os.path.join(sp_dir, "w000.b000000.stepack")),
                "xcpack_dir_bytes": nbytes, "input_write_s": round(prep, 1),
                "detail": "bin/xc_fast file to file on the box's scratch disk: StepPack read, "
                          "upload, correlation, download, xcpack SAC records and TSVs written"}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


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
        # nfft 37044 runs the compile-time-plan lag kernel (pass 0 in registers
        # unless FXC_XC_LAG_R0=0 selects the all-LDS form)
        if args.nspec == 18523:
            name = "lag_kernel_fixed" if os.environ.get("FXC_XC_LAG_R0") == "0" \
                else "lag_kernel_fixed_r0"
        else:
            name = "lag_kernel"
        t = mode["lag_ms"] / max(mode["lag_launches_per_step"], 1)
        launches = max(mode["lag_launches_per_step"], 1)
    byts = algorithmic_bytes(args, npairs) / launches
    achieved = byts / (t / 1e3) / 1e9
    traffic = read_traffic(args.traffic_file, mode["precision"], name) if default_workload else None
    rec = {"kernel": name, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
           "launch_ms": round(t, 3), "algorithmic_bytes_per_launch": int(byts)}
    if traffic:
        # the bandwidth the kernel actually moves (measured bytes / launch time)
        rec["traffic_gbs"] = round(traffic / (t / 1e3) / 1e9, 1)
    lim = read_limiter(args.traffic_file, name) if default_workload else None
    if lim:
        rec["limiter"] = lim
    return rec


def read_limiter(path, kernel):
    """What the kernel's hardware counters say limits it (a 'limiter <kernel>
    <what>' line of the committed traffic summary), e.g. VALU issue."""
    try:
        with open(path) as f:
            for line in f:
                parts = line.split(None, 2)
                if len(parts) == 3 and parts[0] == "limiter" and parts[1] == kernel:
                    return parts[2].strip()
    except OSError:
        pass
    return None


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


# ----------------------------------------------------------------- stages --
# The other stages of the path, each at the 1000-station-day scale, timed as whole
# entry-point calls (allocation of their small workspaces included) between
# device synchronisations.  Work units and algorithmic bytes/flops:
#   sac2spec: station-windows; 4*seg in + 8*nspec out bytes each
#   PWS:      pairs; 4*N*(G+2) bytes each (G group rows, linear stack, output)
#   TF-PWS:   pairs; 5*N*log2(N) flops per complex transform, (G+1)*(N/2+1)
#             voice transforms + 1 final transform per pair
#   linear:   pairs; 4*N*(T+1) bytes each (T records in, one stack out)
FP32_VECTOR_PEAK_TFLOPS = 157.3


def _time_calls(fn, reps):
    import torch

    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


STAGE_TRAFFIC = os.path.join(ROOT, "profiles", "stages_hbm_traffic_1000sta.txt")


def _stage_traffic(name):
    """HBM bytes per call of the stage (its kernels summed), measured with
    rocprofv3 FETCH_SIZE / WRITE_SIZE at these workloads; None when absent."""
    try:
        with open(STAGE_TRAFFIC) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic" and parts[1] == name:
                    return int(parts[2])
    except OSError:
        pass
    return None


def _stage(name, units, unit, sec, work, bound, peak, punit, detail):
    ach = work / sec / (1e9 if punit == "GB/s" else 1e12)
    return {"stage": name, "value": round(units / sec, 1), "unit": unit,
            "ms_per_call": round(sec * 1e3, 3), "detail": detail,
            "roofline": {"bound": bound, "achieved": round(ach, 2), "peak": peak, "unit": punit,
                         "frac": round(ach / peak, 4), "traffic": _stage_traffic(name),
                         "limiter": read_limiter(STAGE_TRAFFIC, name)}}


STAGE_CPU_SECONDS = 3.0
STAGE_CPU_THREADS = 1


def _stage_cpu(fn, per_call, unit, sample, seconds=None, threads=1, kind="port"):
    """A CPU leg timed over a bounded sample: calls of `fn`, each worth
    `per_call` work units, repeated until `seconds` have passed.  `threads`
    is what fn uses (the C restatements take it as an argument; numpy legs
    run with the BLAS and FFT pools limited to one thread)."""
    from threadpoolctl import threadpool_limits

    seconds = STAGE_CPU_SECONDS if seconds is None else seconds
    with threadpool_limits(limits=1):
        fn()
        calls = 0
        t0 = time.perf_counter()
        while True:
            fn()
            calls += 1
            dt = time.perf_counter() - t0
            if dt >= seconds:
                break
    return {"value": round(calls * per_call / dt, 3), "unit": unit, "cores": threads,
            "kind": kind,
            "sample": f"{calls} x {sample}, {dt:.1f} s on {threads} thread"
                      f"{'s' if threads > 1 else ''}"}


def _butter_power(nfft, delta, lo, hi):
    """|B/A|^2 of a 2nd-order Butterworth band-pass (scipy design), k = 0..nfft/2."""
    import numpy as np
    from scipy.signal import butter

    b, a = butter(2, [lo, hi], btype="bandpass", fs=1.0 / delta)
    z = np.exp(-1j * np.pi * np.arange(nfft // 2 + 1) / (nfft // 2))
    return (np.abs(np.polyval(b[::-1], z) / np.polyval(a[::-1], z)) ** 2).astype(np.float32)


def stage_sac2spec(dev, reps, cpu=True):
    """1000 stations x one day at dt 0.1 s: 24 one-hour windows, RUN-ABS,
    whitening after, band 0.1-2 Hz (seg 36000, filter FFT 36288, output FFT 37044)."""
    import numpy as np
    import torch
    from fastxc_amd.ops import S2sPlan

    nsta, npts, seg, nfft_f, nfft_o, nstep = 1000, 864000, 36000, 36288, 37044, 24
    resp = _butter_power(nfft_f, 0.1, 0.1, 2.0)[None]
    plan = S2sPlan(npts, seg, seg, nfft_f, nfft_o, 1, 0.1, 0.1, [240, 360, 7200, 9597], resp,
                   [0.1], whiten=2, normalize=3)
    g = torch.Generator(device=dev)
    g.manual_seed(7)
    x = torch.randn((nsta, npts), dtype=torch.float32, device=dev, generator=g)
    out = torch.empty((nstep, nsta, nfft_o // 2 + 1), dtype=torch.complex64, device=dev)
    steps = np.arange(nstep, dtype=np.int32)
    sec = _time_calls(lambda: plan.process(x, steps, out), reps)
    ok = bool(torch.isfinite(torch.view_as_real(out[:, :: 97])).all().item())
    plan.close()
    units = nsta * nstep
    work = units * (4 * seg + 8 * (nfft_o // 2 + 1))
    del x, out
    torch.cuda.empty_cache()
    res = _stage("sac2spec", units, "station-windows/s", sec, work, "hbm", HBM_PEAK_GBS, "GB/s",
                 {"stations": nsta, "windows": nstep, "seg": seg, "nfft_filter": nfft_f,
                  "nfft_out": nfft_o, "normalize": "RUN-ABS", "whiten": "after",
                  "output_finite": ok})
    if cpu:
        from oracle import cbind
        from oracle import sac2spec as o

        th = STAGE_CPU_THREADS
        oplan = o.Plan(npts=npts, delta=0.1, seg=seg, shift=seg, nstep=nstep, nfft_out=nfft_o,
                       nspec=nfft_o // 2 + 1, nfft_filter=nfft_f, freq_low=0.1, freq_high=2.0,
                       idx=(240, 360, 7200, 9597), responses=[resp[0]], freq_lows=[0.1])
        nc = 2 * th
        xs = np.random.default_rng(7).standard_normal((nc, seg)).astype(np.float32)
        res["cpu_baseline"] = _stage_cpu(
            lambda: cbind.cpu_s2s_process(xs, [0], oplan, 2, 3, 1, nthreads=th), nc,
            "station-windows/s",
            f"{nc} station-windows (oracle/cpu_stages.c cpu_s2s_process, FP32 with FP64 "
            "sums, same plan)", threads=th)
    return res


def _stack_inputs(dev, npairs, T, B, n, kind, seed):
    import numpy as np
    import torch

    G = -(-T // B)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    pre = torch.randn((npairs * G, n), dtype=torch.float32, device=dev, generator=g)
    sizes = [min(B, T - i * B) for i in range(G)]
    w = torch.tensor([s / T for s in sizes] * npairs, dtype=torch.float32, device=dev)
    lin = pre.view(npairs, G, n).mean(1).contiguous()
    off = np.arange(0, npairs * G + 1, G, dtype=np.int64)
    return pre, w, lin, off, G


def _stack_inputs_host(npairs, T, B, n, seed):
    """Host copy of the stack-stage input shapes for the CPU leg: sub-stack
    rows [npairs*G, n], weights [npairs*G], group offsets [npairs+1], linear
    stacks [npairs, n] (float32)."""
    import numpy as np

    G = -(-T // B)
    pre = np.random.default_rng(seed).standard_normal((npairs * G, n)).astype(np.float32)
    w = np.array([min(B, T - i * B) / T for i in range(G)] * npairs, dtype=np.float32)
    off = np.arange(0, npairs * G + 1, G, dtype=np.int64)
    return pre, w, off, pre.reshape(npairs, G, n).mean(axis=1)


def stage_pws(dev, reps, cpu=True):
    """49950 pairs (a tenth of the 1000-station triangle), 30 timestamps in
    sub-stacks of 10 (pre_stack_size), 2001 lags."""
    import torch
    from fastxc_amd import ops

    npairs, T, B, n = 49950, 30, 10, 2001
    pre, w, lin, off, G = _stack_inputs(dev, npairs, T, B, n, "pws", 11)
    out = torch.empty_like(lin)
    sec = _time_calls(lambda: ops.pws(pre, w, off, lin, out=out), reps)
    res = _stage("pws", npairs, "pairs/s", sec, npairs * 4 * n * (G + 2), "hbm", HBM_PEAK_GBS,
                 "GB/s", {"pairs": npairs, "timestamps": T, "substack": B, "groups": G, "n": n})
    del pre, w, lin, out
    torch.cuda.empty_cache()
    if cpu:
        from oracle import cbind

        th = STAGE_CPU_THREADS
        m = 8 * th
        pre_c, w_c, off_c, lin_c = _stack_inputs_host(m, T, B, n, 11)
        res["cpu_baseline"] = _stage_cpu(
            lambda: cbind.cpu_pws(pre_c, w_c, off_c, lin_c, nthreads=th), m, "pairs/s",
            f"{m} pairs of {G} sub-stacks x {n} lags (oracle/cpu_stages.c cpu_pws, FP32)",
            threads=th)
    return res


def stage_tfpws(dev, reps, cpu=True):
    """2000 pairs, 30 timestamps in sub-stacks of 10, 2001 lags, all 1001 voices."""
    import math

    import torch
    from fastxc_amd import ops

    npairs, T, B, n = 2000, 30, 10, 2001
    pre, w, lin, off, G = _stack_inputs(dev, npairs, T, B, n, "tfpws", 13)
    out = torch.empty_like(lin)
    sec = _time_calls(lambda: ops.tfpws(pre, w, off, lin, out=out), reps)
    flops = npairs * ((G + 1) * (n // 2 + 1) + 1) * 5.0 * n * math.log2(n)
    res = _stage("tfpws", npairs, "pairs/s", sec, flops, "fp32-vector", FP32_VECTOR_PEAK_TFLOPS,
                 "TFLOP/s", {"pairs": npairs, "timestamps": T, "substack": B, "groups": G, "n": n,
                             "voices": n // 2 + 1})
    del pre, w, lin, out
    torch.cuda.empty_cache()
    if cpu:
        from oracle import cbind

        th = STAGE_CPU_THREADS
        pre_c, w_c, off_c, lin_c = _stack_inputs_host(th, T, B, n, 13)
        res["cpu_baseline"] = _stage_cpu(
            lambda: cbind.cpu_tfpws(pre_c, w_c, off_c, lin_c, nthreads=th), th, "pairs/s",
            f"{th} pairs of {G} sub-stacks x {n} lags, {n // 2 + 1} voices "
            "(oracle/cpu_stages.c cpu_tfpws, FP32)", threads=th)
    return res


def stage_linear(dev, reps, cpu=True):
    """49950 pairs x 30 timestamps x 2001 lags, float64 running sum."""
    import numpy as np
    import torch
    from fastxc_amd import ops

    npairs, T, n = 49950, 30, 2001
    g = torch.Generator(device=dev)
    g.manual_seed(17)
    x = torch.randn((npairs * T, n), dtype=torch.float32, device=dev, generator=g)
    off = np.arange(0, npairs * T + 1, T, dtype=np.int64)
    out = torch.empty((npairs, n), dtype=torch.float32, device=dev)
    sec = _time_calls(lambda: ops.linear_stack(x, off, out=out), reps)
    res = _stage("linear_stack", npairs, "pairs/s", sec, npairs * 4 * n * (T + 1), "hbm",
                 HBM_PEAK_GBS, "GB/s", {"pairs": npairs, "timestamps": T, "n": n,
                                        "accumulator": "f64"})
    del x, out
    torch.cuda.empty_cache()
    if cpu:
        from oracle import cbind

        th = STAGE_CPU_THREADS
        m = 64 * th
        recs = np.random.default_rng(17).standard_normal((m * T, n)).astype(np.float32)
        off_c = np.arange(0, m * T + 1, T, dtype=np.int64)
        res["cpu_baseline"] = _stage_cpu(
            lambda: cbind.cpu_linear_stack(recs, off_c, nthreads=th), m, "pairs/s",
            f"{m} pairs of {T} records x {n} lags (oracle/cpu_stages.c cpu_linear_stack, "
            "FP64 sum)", threads=th)
    return res


def stage_rotate(dev, reps, cpu=True):
    """49950 station pairs x 9 components x 2001 lags, ENZ -> RTZ."""
    import numpy as np
    import torch
    from fastxc_amd import ops

    npairs, n = 49950, 2001
    g = torch.Generator(device=dev)
    g.manual_seed(23)
    x = torch.randn((npairs, 9, n), dtype=torch.float32, device=dev, generator=g)
    ang = np.deg2rad(np.random.default_rng(23).uniform(0, 360, (npairs, 2))).astype(np.float32)
    trig = torch.from_numpy(np.stack([np.sin(ang[:, 0]), np.cos(ang[:, 0]), np.sin(ang[:, 1]),
                                      np.cos(ang[:, 1])], axis=1)).to(dev)
    out = torch.empty_like(x)
    sec = _time_calls(lambda: ops.rotate_rtz(x, trig, out=out), max(reps, 20))
    res = _stage("rotate", npairs, "pairs/s", sec, npairs * (72 * n + 16), "hbm", HBM_PEAK_GBS,
                 "GB/s", {"pairs": npairs, "components": 9, "n": n})
    del x, out
    torch.cuda.empty_cache()
    if cpu:
        from oracle import rotation as o

        enz = np.random.default_rng(23).standard_normal((64, 9, n)).astype(np.float32)
        angs = np.rad2deg(ang[:64].astype(np.float64))
        res["cpu_baseline"] = _stage_cpu(
            lambda: [o.rotate64(enz[i], angs[i, 0], angs[i, 1]) for i in range(64)], 64,
            "pairs/s", f"64 pairs of 9 components x {n} lags (oracle/rotation.py rotate64, FP64)")
    return res


_STAGES = {"sac2spec": "stage_sac2spec", "pws": "stage_pws", "tfpws": "stage_tfpws",
           "linear_stack": "stage_linear", "rotate": "stage_rotate"}


def stages(reps=3, only="", cpu=True, threads=1):
    import torch

    global STAGE_CPU_THREADS
    STAGE_CPU_THREADS = threads

    dev = torch.device("cuda:0")
    out = []
    names = [x for x in only.split(",") if x] or list(_STAGES)
    for name in names:
        fn = globals()[_STAGES[name]]
        print(f"# stage {fn.__name__}", file=sys.stderr, flush=True)
        out.append(fn(dev, reps, cpu))
    return out


# ----------------------------------------------------------------- stream --
# Streaming segment updates (fxc_stream_*): one-hour segments at dt 0.1 s for
# random subnetworks of 10-500 stations of a 2000-station array (Z
# channels), Poisson arrivals at each rate of a sweep; every segment runs
# sac2spec, xc of all its pairs (split-BF16) and the per-pair float64 linear
# stack update with a 4-record correlation history.  Latency runs from
# submit to the applied state update; pair-windows/s is the work of the
# segments of one rate over the time from its first submit to its last
# completion.  Traces come from a seeded pool (contiguous row blocks of a
# random station permutation, so submitting copies nothing).
def _stream_rate(eng, rng, rate, seconds, nsta, ncomp, pool, perm):
    """Poisson arrivals at `rate` for `seconds`; each segment a random
    subnetwork of 10-500 stations (all ncomp channels of each), channel ids
    submitted in a random order.  Offered load: pair-windows of the submitted
    segments over the arrival span; served: the same pair-windows over the
    span from the first arrival to the last completed update."""
    import numpy as np

    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=max(1, int(rate * seconds))))
    sizes = rng.integers(10, 501, size=len(arrivals))
    tickets, pairs = [], 0
    start = time.perf_counter()
    for at, n in zip(arrivals, sizes):
        delay = start + at - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        sta = rng.choice(nsta, size=int(n), replace=False)
        chan = (perm[sta][:, None] * ncomp + np.arange(ncomp)[None]).reshape(-1).astype(np.int32)
        rows = (sta[:, None] * ncomp + np.arange(ncomp)[None]).reshape(-1)
        tickets.append(eng.submit(chan, pool[rows]))
        m = int(n) * ncomp
        pairs += m * (m - 1) // 2
    lat, refused = [], 0
    for t in tickets:
        try:
            lat.append(eng.wait(t))
        except Exception:  # refused by the state footprint (FXC_ERR_ALLOC)
            refused += 1
    done = time.perf_counter() - start
    lat = np.array(lat) if lat else np.array([float("nan")])
    offered = pairs / float(arrivals[-1])
    served = pairs / done
    return {"rate_per_s": rate, "segments": len(tickets), "refused": refused,
            "seconds": round(float(arrivals[-1]), 1),
            "p50_ms": round(float(np.percentile(lat, 50)), 2),
            "p99_ms": round(float(np.percentile(lat, 99)), 2), "max_ms": round(float(lat.max()), 2),
            "offered_pair_windows_per_s": round(offered, 1),
            "served_pair_windows_per_s": round(served, 1),
            "served_over_offered": round(served / offered, 3)}


def stream_sweep(rates, seconds, seed=5, ncomp=1, dt=0.1, nsta=2000, lanes=4, history=4):
    """Streaming sweep of one engine configuration.  A rate holds when the
    engine serves >= 0.98 of the offered pair-windows and p99 < 1 s; the
    rated rate is the highest rate that holds."""
    import numpy as np
    from fastxc_amd.ops import StreamEngine

    if dt == 0.1:
        npts, nfft_f, nfft_o, half, idx = 36000, 36288, 37044, 1000, [240, 360, 7200, 9597]
    else:
        npts, nfft_f, nfft_o, half, idx = 72000, 72576, 74088, 2000, [240, 360, 7200, 9597]
    seg = npts
    resp = _butter_power(nfft_f, dt, 0.1, 2.0)[None]
    rng = np.random.default_rng(seed)
    nchan = nsta * ncomp
    print(f"# stream: trace pool ({nchan} channels, dt {dt})", file=sys.stderr, flush=True)
    pool = rng.standard_normal((nchan, npts), dtype=np.float32)
    perm = rng.permutation(nsta).astype(np.int64)
    max_pairs = nchan * (nchan - 1) // 2
    t0 = time.perf_counter()
    eng = StreamEngine(npts, seg, seg, nfft_f, nfft_o, ncomp, dt, 0.1, idx, resp, [0.1], nchan,
                       half, precision="bf16x3", history=history, max_pairs=max_pairs,
                       max_seg_chan=500 * ncomp, lanes=lanes, pws_substack=1)
    setup_s = time.perf_counter() - t0
    info0 = eng.info()
    out = []
    for rate in rates:
        print(f"# stream: rate {rate}/s for {seconds} s", file=sys.stderr, flush=True)
        out.append(_stream_rate(eng, rng, rate, seconds, nsta, ncomp, pool, perm))
        print(f"#   p99 {out[-1]['p99_ms']} ms, served/offered {out[-1]['served_over_offered']}",
              file=sys.stderr, flush=True)
    info = eng.info()
    eng.close()
    ok = [r["rate_per_s"] for r in out
          if r["p99_ms"] < 1000.0 and r["served_over_offered"] >= 0.98 and r["refused"] == 0]
    comps = "Z channels" if ncomp == 1 else "E/N/Z channels"
    return {"workload": f"one-hour segments (dt {dt} s, nfft {nfft_o}, {2 * half + 1} lags), random "
                        f"10-500 station subnetworks of a {nsta}-station array, {comps}, Poisson "
                        f"arrivals, {seconds:g} s per rate",
            "form": "bf16x3", "lanes": lanes, "history": history, "running_pws_substack": 1,
            "setup_s": round(setup_s, 2), "state_gb": round(info0["state_bytes"] / 1e9, 1),
            "lane_gb": round(info0["lane_bytes"] / 1e9, 2), "pairs_with_state": info["pairs_used"],
            "rates": out, "rated_rate_per_s": max(ok) if ok else None,
            "p99_ms_at_rated": next((r["p99_ms"] for r in out if ok and r["rate_per_s"] == max(ok)),
                                    None)}


def stream_all(rates, seconds):
    """The Z sweep at dt 0.1 s (the rated configuration), then a
    three-component sweep (three rates) and a dt 0.05 s sweep (two rates)."""
    res = {"z_dt0.1": stream_sweep(rates, seconds)}
    res["enz_dt0.1"] = stream_sweep([8, 12, 16], max(10.0, seconds / 2), seed=6, ncomp=3, nsta=700,
                                    lanes=3, history=0)
    res["z_dt0.05"] = stream_sweep([8, 12], max(10.0, seconds / 2), seed=7, dt=0.05, nsta=1000,
                                   history=0)
    return res


# ------------------------------------------------------------- sourcepack --
_XC_COLUMNS = ("timestamp worker_id anchor_block target_begin_block target_end_block block_size "
               "anchor_begin anchor_end target_begin target_end pack_path offset bytes src_id "
               "rec_id src_network src_station src_location src_component rec_network "
               "rec_station rec_location rec_component npts dt dist az baz "
               "final_pair_path").split()


def _write_xcpack(root, nsta, lags=2001, ts="20200502T04:00", limit=None):
    """The xcpack TSV xc_fast writes for one timestamp of nsta stations (Z),
    all pairs in task order, over a sparse pack file of the right size."""
    x = os.path.join(root, "ncf", "xcpack")
    os.makedirs(x, exist_ok=True)
    stem = ts + ".i000000.j000000.w000.p000"
    rb = 632 + 4 * lags
    lines, n = ["\t".join(_XC_COLUMNS)], 0
    for a in range(1, nsta + 1):
        for b in range(a + 1, nsta + 1):
            if limit is not None and n == limit:
                break
            sa, sb = "S%04d" % a, "S%04d" % b
            lines.append(f"{ts}\t0\t0\t0\t0\t{nsta}\t0\t{nsta}\t0\t{nsta}\t{stem}.xcpack\t"
                         f"{n * rb}\t{rb}\t{a}\t{b}\tXX\t{sa}\t\tBHZ\tXX\t{sb}\t\tBHZ\t{lags}\t0.1\t"
                         f"{0.1 * (a + b):.1f}\t{a % 360}\t{b % 360}\t/out/ncf/XX.{sa}/XX.{sb}/"
                         f"XX-XX.{sa}-{sb}.BHZ-BHZ.sac")
            n += 1
    with open(os.path.join(x, stem + ".tsv"), "w") as fh:
        fh.write("\n".join(lines) + "\n")
    with open(os.path.join(x, stem + ".xcpack"), "wb") as fh:
        fh.truncate(n * rb)
    with open(os.path.join(x, ts + ".done"), "w") as fh:
        fh.write("DONE\n")
    return x, n


# ------------------------------------------------------------------ batch --
# Uneven time segments in one set of launches (fxc_xc_batch_*): 64 segments,
# stations ~U[20,400], windows ~U[6,48], 10 % of station-windows gaps,
# one-hour windows at dt 0.1 s (nspec 18523, 2001 lags), gaps excluded from
# each pair's stack.  The grouping tuner picks one plan or one plan per
# operand depth; the chosen batch is captured into a HIP graph and replayed.
# pair-windows = sum over segments of pairs x windows.  Memory is estimated
# first and the segment count lowered (and said so) if it would not fit.
def batch_measure(nseg=64, seed=13, reps=3, budget_gb=235.0, max_rows=65536):
    import numpy as np
    import torch
    from fastxc_amd.ops import XcBatch

    nspec, half = 18523, 1000
    rng = np.random.default_rng(seed)
    nsta = rng.integers(20, 401, size=nseg)
    nstep = rng.integers(6, 49, size=nseg)

    def depth(w):  # operand depth of a plan (the compiled depths of the contraction)
        return next(d for d in (8, 16, 24, 32, 40, 48, 64, 96, 128) if w <= d)

    def gbytes(k):
        """Peak device bytes (GB) of grouping 0 and of grouping 1: spectra,
        outputs, the split operands of the largest group and one shared
        cross-spectrum workspace."""
        a, w = nsta[:k], nstep[:k]
        fixed = float((a * w).sum()) * nspec * 8 + float((a * (a - 1) // 2).sum()) * (2 * half + 1) * 4
        ws = max_rows * 18528 * 8.0
        padded = (a + 31) // 32 * 32
        one = float(padded.sum()) * nspec * 2 * depth(int(w.max())) * 4 + ws
        kinds = sorted({depth(int(x)) for x in w})
        per = max(float(padded[[depth(int(x)) == d for x in w]].sum()) * nspec * 2 * d * 4
                  for d in kinds) + ws
        return (fixed + one) / 1e9, (fixed + per) / 1e9

    k = nseg
    while k > 1 and min(gbytes(k)) > budget_gb:
        k -= 1
    nsta, nstep = nsta[:k], nstep[:k]
    g0, g1 = gbytes(k)
    dev = torch.device("cuda:0")
    print(f"# batch: {k} segments, ~{g0:.0f} / {g1:.0f} GB (grouping 0 / 1)", file=sys.stderr,
          flush=True)
    b = XcBatch(nspec, 0.1, half, nsta.tolist(), nstep.tolist(), gaps="exclude",
                max_rows=max_rows)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    spec = torch.randn(b.spec_elems, dtype=torch.complex64, device=dev, generator=g)
    valid = (torch.rand(b.valid_elems, device=dev, generator=g) >= 0.1).to(torch.uint8)
    out = torch.empty((b.ntasks, b.cc_size), dtype=torch.float32, device=dev)
    # the tuner times grouping x row chunk; candidates whose workspaces do not
    # fit beside the spectra, the outputs and the operands are skipped
    room = budget_gb - max(g0, g1) + max_rows * 18528 * 8.0 / 1e9
    tune = b.autotune(spec, valid, out, rows=(32768, 65536, 131072), reps=1, budget_gb=room)
    tuner = {"key": tune["key"], "grouping": tune["grouping"], "max_rows": tune["max_rows"],
             "trials": [{k: (round(v, 2) if isinstance(v, float) else v) for k, v in t.items()}
                        for t in tune["trials"]]}
    ngroups = b.info()["groups"]
    sec = _time_calls(lambda: b.correlate(spec, valid, out), reps)
    b.capture(spec, valid, out)
    torch.cuda.synchronize()
    gsec = _time_calls(lambda: b.replay(), reps)
    pw = float((nsta * (nsta - 1) // 2 * nstep).sum())
    b.correlate(spec, valid, out)
    torch.cuda.synchronize()
    # the PWS leg in FP32: every pair of the shared station array over the
    # segments containing it (sub-stacks of one record)
    from fastxc_amd.ops import batch_pws
    b_nsta, b_base = list(b.nsta), b.task_base
    b.close()
    del spec, valid
    torch.cuda.empty_cache()

    class _Rows:  # the row layout batch_pws reads
        nsta, task_base = b_nsta, b_base

    batch_pws(_Rows, out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    psrc, _, pout = batch_pws(_Rows, out)
    torch.cuda.synchronize()
    psec = time.perf_counter() - t0
    pws_leg = {"pairs": int(len(psrc)), "records": int(out.shape[0]),
               "ms": round(psec * 1e3, 2), "pairs_per_s": round(len(psrc) / psec, 1),
               "note": "gather of each pair's records + linear mean (torch) + fxc_pws, FP32"}
    del pout
    res = {"segments": int(k), "pairs": int(b.ntasks), "pair_windows": pw,
           "stations_total": int(nsta.sum()),
           "windows_min_max": [int(nstep.min()), int(nstep.max())],
           "gaps": "10 % of station-windows, excluded per pair", "nspec": nspec,
           "lags": 2 * half + 1, "max_rows": tune["max_rows"], "tuner": tuner,
           "groups": ngroups,
           "ms_eager": round(sec * 1e3, 2), "ms_graph": round(gsec * 1e3, 2),
           "pair_windows_per_s": round(pw / min(sec, gsec), 1),
           "est_gb": [round(g0, 1), round(g1, 1)], "pws_leg": pws_leg}
    del out
    torch.cuda.empty_cache()
    return res


def sourcepack_measure(nsta=1000, sample=50000):
    import shutil
    import subprocess
    import tempfile

    exe = os.path.join(ROOT, "bin", "sourcepack_index")
    tmp = tempfile.mkdtemp(prefix="fxc_sourcepack_")
    try:
        x, n = _write_xcpack(os.path.join(tmp, "full"), nsta)
        t0 = time.perf_counter()
        r = subprocess.run([exe, "-I", x, "-O", os.path.join(tmp, "full", "sp")],
                           capture_output=True, text=True)
        sec = time.perf_counter() - t0
        if r.returncode != 0:
            raise RuntimeError("sourcepack_index failed: " + r.stderr[-500:])
        xs, ns = _write_xcpack(os.path.join(tmp, "sample"), nsta, limit=sample)
        from oracle import sourcepack as osp

        t0 = time.perf_counter()
        osp.build_all(xs, os.path.join(tmp, "sample", "sp"))
        port = time.perf_counter() - t0
        return {"records": n, "value": round(n / sec, 1), "unit": "records/s",
                "sec": round(sec, 3), "threads": 1,
                "cpu_port": {"value": round(ns / port, 1), "unit": "records/s", "cores": 1,
                             "kind": "port", "sample": f"{ns} records, same TSV layout"},
                "detail": "one timestamp of 1000 stations (all pairs), file to file "
                          "(read xcpack TSV, sort, write index + _SUCCESS)"}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


# ---------------------------------------------------------------- planner --
def planner_measure(nsta=2000, sample=200, seed=9):
    """All station pairs (with autocorrelations) of nsta stations in a
    15 x 20 degree box: the native planner (threads = this process's cores,
    at most 16) against the plain Python loop of the oracle on the pairs of
    the first `sample` stations."""
    import numpy as np
    from fastxc_amd.geodesy import distkm_az_baz_rudoe
    from fastxc_amd.planner import GeoNode, PairFilterConfig, build_valid_paths
    from oracle import planner as opl

    rng = np.random.default_rng(seed)
    nodes = [GeoNode(k + 1, "1", "XX", f"S{k:04d}", "00", "x",
                     lat=float(np.float32(rng.uniform(30, 45))),
                     lon=float(np.float32(rng.uniform(-120, -100)))) for k in range(nsta)]
    threads = min(16, len(os.sched_getaffinity(0)))
    t0 = time.perf_counter()
    paths, checks = build_valid_paths(nodes, PairFilterConfig(), threads=threads)
    sec = time.perf_counter() - t0
    sub = [(n.lat, n.lon, n.group) for n in nodes[:sample]]
    cfg = dict(allow_autocorr=True, autocorr_only=False, group_mode="all", use_distance=False,
               min_km=0.0, max_km=50000.0, use_azimuth=False, ranges=((0.0, 360.0),))
    t0 = time.perf_counter()
    _, pchecks = opl.valid_paths(sub, cfg, distkm_az_baz_rudoe)
    psec = time.perf_counter() - t0
    return {"stations": nsta, "pairs": len(paths), "value": round(checks / sec, 1),
            "unit": "pairs/s", "sec": round(sec, 3), "threads": threads,
            "cpu_port": {"value": round(pchecks / psec, 1), "unit": "pairs/s", "cores": 1,
                         "kind": "port", "sample": f"{pchecks} pairs of {sample} stations"},
            "detail": "Rudoe geometry and filters of every station pair, float64 libm, "
                      "bit-identical values"}


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
    rates = [float(x) for x in args.stream_rates.split(",") if x]
    if args.stages_only:
        print(json.dumps({"stages": stages(only=args.stage, cpu=not args.no_cpu, threads=cpu_threads(args))}), flush=True)
        return
    if args.batch_only:
        print(json.dumps({"batch": batch_measure()}), flush=True)
        return
    if args.xc_file_to_file:
        print(json.dumps({"xc_file_to_file": xc_file_to_file()}), flush=True)
        return
    if args.stream_only:
        print(json.dumps({"stream": stream_all(rates, args.stream_seconds)}), flush=True)
        return
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(1234)
    spec = torch.randn((args.nstep, args.nsta, args.nspec), dtype=torch.complex64, device=dev,
                       generator=g)
    src, rec = np.triu_indices(args.nsta, k=1)
    src = src.astype(np.int32)
    rec = rec.astype(np.int32)
    out = torch.empty((len(src), 2 * args.half_cc + 1), dtype=torch.float32, device=dev)
    check_pairs, check_rows = sample_pairs(args.nsta, args.check_pairs)
    modes = [run_mode(p, spec, src, rec, args, out, check_rows) for p in args.precision.split(",")]
    # every form against an FP64 host evaluation of the sampled pairs (first,
    # last, tile-diagonal and off-diagonal ones included)
    want = fp64_rows(spec, check_pairs, args.half_cc)
    for m in modes:
        rows = m.pop("_rows")
        d = rows - want
        m["rel_l2_vs_fp64"] = float(np.linalg.norm(d) / np.linalg.norm(want))
        m["max_abs_vs_fp64"] = float(np.abs(d).max())
        m["max_row_rel_l2_vs_fp64"] = float(
            (np.linalg.norm(d, axis=1) / np.linalg.norm(want, axis=1)).max())
    fp64_check = {"pairs": len(check_pairs), "peak_abs": float(np.abs(want).max()),
                  "sample": "first and last pairs, pairs inside a diagonal and an off-diagonal "
                            "32-station tile, random pairs; FP64 numpy evaluation on the host"}
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
                  "fp16x3": "f16x3-split/f32-acc", "plain": "f32"}[head["precision"]],
        "data": "synthetic",
        "config": {"workload": f"xc_fast one timestamp: {args.nsta} stations x {args.nstep} windows, "
                               f"nspec {args.nspec} (nfft {2 * (args.nspec - 1)}), {len(src)} pairs, "
                               f"{2 * args.half_cc + 1} lags",
                   "stations": args.nsta, "windows": args.nstep, "nspec": args.nspec,
                   "pairs": len(src), "lags": 2 * args.half_cc + 1, "chunk_rows": args.chunk_rows},
        "modes": modes,
        "fp64_check": fp64_check,
        "roofline": roofline(head, args, len(src), default_workload),
    }
    for m in modes[1:]:
        res["ratio_vs_" + m["precision"]] = round(head["pair_windows_per_s"] /
                                                   m["pair_windows_per_s"], 3)
    del spec
    torch.cuda.empty_cache()
    if not args.no_band:
        print("# band spec mode", file=sys.stderr, flush=True)
        res["band_spec_mode"] = band_measure(args, src, rec, out, check_pairs, check_rows)
    del out
    torch.cuda.empty_cache()
    if not args.no_stages:
        res["stages"] = stages(only=args.stage, cpu=not args.no_cpu, threads=cpu_threads(args))
    if not args.no_stream:
        res["stream"] = stream_all(rates, args.stream_seconds)
    if not args.no_sourcepack:
        print("# sourcepack indexer", file=sys.stderr, flush=True)
        res["sourcepack"] = sourcepack_measure()
    if not args.no_batch:
        try:
            res["batch"] = batch_measure()
        except Exception as e:  # the headline line must still be printed
            res["batch"] = {"error": str(e)[:300]}
            torch.cuda.empty_cache()
    if not args.no_planner:
        print("# path planner", file=sys.stderr, flush=True)
        res["planner"] = planner_measure()
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
