"""The measured workloads run as chained GPU workloads against the FP64 oracle.

config 0  8 stations, Z, one day at dt 0.1 s (864000 samples), 24 one-hour
          windows, RUN-ABS, whitening after normalisation on 0.1-2 Hz, 100 s
          lags: sac2spec -> xc_fast (fp32 and split-bf16) -> linear stack,
          every stage file to file through the executables.
config 1  the production shape at a few stations: 3 components, 47 windows a
          day (shift 1800 s), nfft 37044, three days: sac2spec -> xc_fast ->
          linear stack and ncf_pws (B 1; power 1 as the reference, power 2 as
          the spec mode) over the days.
config 2  1000 stations x 24 windows x 18523 bins, every form of the
          contraction at full size, 300 sampled pairs (the first, the last and
          pairs of diagonal and off-diagonal tiles) against an FP64 evaluation
          of just those pairs.

Inputs are tests/synth.py's seeded wavefield.  Tolerances: relative L2 1e-5
for the FP32 chain and 1e-3 for split-bf16 / fp16 (the expected values are
printed); pair and lag order are exact because every record is matched to its
oracle row by (source, receiver, components).
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import sac2spec as s2s
from oracle import stack
from oracle import xc
from tests import fxformats as ff
from tests import synth

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "bin")


def run(exe, args, timeout=600):
    r = subprocess.run([os.path.join(BIN, exe)] + args, capture_output=True, text=True,
                       timeout=timeout)
    if r.returncode != 0:
        print(r.stderr[-3000:])
    return r


def _oracle_spectra(traces, plan, nch):
    """traces [files, npts] (channels of a station adjacent) -> [nstep, files, nspec]."""
    win = s2s.cut_windows(np.asarray(traces, np.float64), plan)
    steps, files, seg = win.shape
    out = s2s.process_batch(win.reshape(steps * files // nch, nch, seg), plan, whiten=2,
                            normalize=3)
    return out.reshape(steps, files, -1)


def _key(row):
    return (int(row["src_id"]), int(row["rec_id"]), row["src_component"], row["rec_component"])


def _oracle_day(tmp, traces, plan, nch, shape, allowed, dt):
    """FP64 spectra and correlations of one day: {(src, rec, cs, cr): ccf}."""
    nsta = traces.shape[0]
    comps = synth.COMPONENTS[nch]
    spec = _oracle_spectra(traces.reshape(nsta * nch, -1), plan, nch)
    gnsl = [s + 1 for s in range(nsta) for _ in range(nch)]
    job = xc.build_jobs(len(gnsl), len(gnsl))[0]
    tasks, _ = xc.build_tasks_for_job(gnsl, allowed, job)
    src = np.array([t.src_meta for t in tasks])
    rec = np.array([t.rec_meta for t in tasks])
    cc = xc.correlate(spec, src, rec, shape.nfft, shape.half_cc, dt)
    out = {}
    for t, row in zip(tasks, cc):
        out[(gnsl[t.src_meta], gnsl[t.rec_meta], comps[t.src_meta % nch],
             comps[t.rec_meta % nch])] = row
    return spec, out


def _sourcepack_indexes(spd):
    """The per-timestamp indexes xc_fast --sourcepack-dir wrote (timestamp
    directories in xc's normalised spelling, chronological when sorted)."""
    return [p / "sourcepack_index.tsv" for p in sorted(spd.iterdir())
            if (p / "sourcepack_index.tsv").exists()]


def _compare(name, got, want, tol):
    keys = sorted(want)
    assert sorted(got) == keys, f"{name}: pair set differs"
    g = np.stack([got[k] for k in keys]).astype(np.float64)
    w = np.stack([want[k] for k in keys])
    err = xc.rel_l2(g, w)
    row = np.linalg.norm(g - w, axis=1) / np.linalg.norm(w, axis=1)
    print(f"{name}: records={len(keys)} rel_l2={err:.3e} max_row={row.max():.3e} "
          f"max_abs={np.abs(g - w).max():.3e} (peak {np.abs(w).max():.3e})", flush=True)
    assert err < tol and row.max() < 10 * tol
    return err


def test_config0_chain(gpu, tmp_path):
    dt, npts, nsta = 0.1, 864000, 8
    traces, _, _ = synth.wavefield(nsta, npts, dt, seed=2024, ncomp=1)
    ts = "2020.001.0000"
    rows = synth.write_day(str(tmp_path), traces, dt, ts, ff.write_sac)
    index = str(tmp_path / "sac_index.tsv")
    synth.write_index(index, rows)
    filt = str(tmp_path / "filter.txt")
    s2s.write_filter_file(filt, dt, "0.1/2")
    ws = tmp_path / "ws"
    r = run("sac2spec", ["-I", index, "-O", str(ws), "-C", "1", "-B", filt, "-L",
                         "86400/3600/3600/100", "-W", "2/0", "-N", "3", "-T", "8"])
    assert r.returncode == 0
    plan = s2s.plan_lengths(86400, dt, 3600, 3600, 100, s2s.read_filters(filt))
    assert (plan.seg, plan.nstep, plan.nfft_out, plan.nspec) == (36000, 24, 37044, 18523)
    packs = sorted(p for p in os.listdir(ws / "stepack") if p.endswith(".stepack"))
    _, nodes, spec = ff.read_stepack(str(ws / "stepack" / packs[0]))
    shape = xc.finalize_shape(plan.nspec, plan.nstep, dt, 100.0)
    assert (shape.nfft, shape.half_cc) == (37044, 1000)
    allowed = {a * 10000 + b for a in range(1, nsta + 1) for b in range(a + 1, nsta + 1)}
    want_spec, want = _oracle_day(str(tmp_path), traces, plan, 1, shape, allowed, dt)
    err = xc.rel_l2(spec, want_spec)
    print(f"config0 sac2spec: rel_l2={err:.3e} max_abs={np.abs(spec - want_spec).max():.3e}")
    assert err < 1e-5
    paths = str(tmp_path / "allowed_paths.tsv")
    ff.write_allowed_paths(paths, [(a // 10000, a % 10000, 0.1, 10.0, 1.0, 2.0)
                                   for a in sorted(allowed)])
    for precision, tol in (("fp32", 1e-5), ("bf16x3", 1e-3)):
        out = tmp_path / ("ncf_" + precision)
        spd = tmp_path / ("sp_" + precision)
        r = run("xc_fast", ["-I", str(ws), "-P", paths, "-O", str(out), "-C", "100",
                            "--precision", precision, "--sourcepack-dir", str(spd)])
        assert r.returncode == 0
        recs = ff.read_xcpack_dir(str(out / "xcpack"))
        assert len(recs) == 28
        got = {_key(row): data for row, _, data in recs}
        _compare(f"config0 xc {precision}", got, want, tol)
        if precision == "fp32":
            # linear stack of the day's index: one record per pair, returned as is
            from fastxc_amd.linear_stack import linear_stack_sourcepack_indexes
            idx = _sourcepack_indexes(spd)
            assert len(idx) == 1
            res = linear_stack_sourcepack_indexes(idx, tmp_path / "stack", device=gpu)
            assert len(res) == 28
            lin = ff.read_sourcepack_index(
                str(tmp_path / "stack" / "linearstack_sourcepack" / "STACK" / "sourcepack_index.tsv"))
            lg = {_key(row): data for row, _, data in lin}
            assert all(np.array_equal(lg[k], got[k]) for k in got)
            _compare("config0 linear", lg, want, 1e-5)


def test_config1_production_shape(gpu, tmp_path):
    dt, npts, nsta, nch, days = 0.1, 864000, 4, 3, 3
    filt = str(tmp_path / "filter.txt")
    s2s.write_filter_file(filt, dt, "0.1/2")
    plan = s2s.plan_lengths(86400, dt, 3600, 1800, 100, s2s.read_filters(filt))
    assert (plan.nstep, plan.nfft_out) == (47, 37044)
    shape = xc.finalize_shape(plan.nspec, plan.nstep, dt, 100.0)
    allowed = {a * 10000 + b for a in range(1, nsta + 1) for b in range(a + 1, nsta + 1)}
    rows, want_days, stamps = [], [], []
    for d in range(days):
        ts = "2020.%03d.0000" % (10 + d)
        stamps.append(ts)
        traces, _, _ = synth.wavefield(nsta, npts, dt, seed=100 + d, ncomp=nch)
        rows += synth.write_day(str(tmp_path), traces, dt, ts, ff.write_sac)
        want_days.append(_oracle_day(str(tmp_path), traces, plan, nch, shape, allowed, dt)[1])
    index = str(tmp_path / "sac_index.tsv")
    synth.write_index(index, rows)
    ws = tmp_path / "ws"
    r = run("sac2spec", ["-I", index, "-O", str(ws), "-C", str(nch), "-B", filt, "-L",
                         "86400/3600/1800/100", "-W", "2/0", "-N", "3", "-T", "8"])
    assert r.returncode == 0
    paths = str(tmp_path / "allowed_paths.tsv")
    ff.write_allowed_paths(paths, [(a // 10000, a % 10000, 0.1, 10.0, 1.0, 2.0)
                                   for a in sorted(allowed)])
    out, spd = tmp_path / "ncf", tmp_path / "sp"
    r = run("xc_fast", ["-I", str(ws), "-P", paths, "-O", str(out), "-C", "100",
                        "--sourcepack-dir", str(spd)])
    assert r.returncode == 0
    recs = ff.read_xcpack_dir(str(out / "xcpack"))
    npair = len(allowed) * nch * nch
    assert len(recs) == days * npair
    xc_stamps = sorted({row["timestamp"] for row, _, _ in recs})  # xc's spelling, in day order
    assert len(xc_stamps) == days
    for d, ts in enumerate(xc_stamps):
        got = {_key(row): data for row, _, data in recs if row["timestamp"] == ts}
        _compare(f"config1 xc day {d}", got, want_days[d], 1e-5)
    keys = sorted(want_days[0])
    series = {k: np.stack([w[k] for w in want_days]) for k in keys}   # [days, 2001] per pair
    # linear stack over the days
    from fastxc_amd.linear_stack import linear_stack_sourcepack_indexes
    idx = _sourcepack_indexes(spd)
    assert len(idx) == days
    res = linear_stack_sourcepack_indexes(idx, tmp_path / "stack", device=gpu)
    assert len(res) == npair and all(x.segment_count == days for x in res)
    lin = ff.read_sourcepack_index(
        str(tmp_path / "stack" / "linearstack_sourcepack" / "STACK" / "sourcepack_index.tsv"))
    _compare("config1 linear", {_key(row): data for row, _, data in lin},
             {k: series[k].mean(axis=0) for k in keys}, 1e-5)
    # PWS over the days: the reference (power 1) and the spec mode (power 2)
    lst = str(tmp_path / "sourcepacks.txt")
    with open(lst, "w") as f:
        f.write("\n".join(str(p) for p in idx) + "\n")
    for power in (1.0, 2.0):
        o = str(tmp_path / ("pws_p%g" % power))
        args = ["--sourcepack-list", lst, "--output-sourcepack", o, "-B", "1"]
        if power != 1.0:
            args += ["--power", "%g" % power]
        r = run("ncf_pws", args)
        assert r.returncode == 0
        pr = ff.read_sourcepack_index(os.path.join(o, "sourcepack_index.tsv"))
        want = {}
        for k in keys:
            g, w, l = stack.pws_prepare(series[k], 1)
            want[k] = stack.pws(g, w, l, power=power)
        _compare(f"config1 pws power {power:g}", {_key(row): data for row, _, data in pr}, want,
                 1e-4)


def _sample_pairs(nsta, tile=32, count=300, seed=5):
    """Pairs (i < j) over the triangle: the first and last, every pair of one
    diagonal tile corner, pairs of off-diagonal tiles, and random ones."""
    rng = np.random.default_rng(seed)
    pick = {(0, 1), (nsta - 2, nsta - 1), (0, nsta - 1)}
    for a in range(3):
        for b in range(a + 1, 4):
            pick.add((tile + a, tile + b))          # inside diagonal tile 1
            pick.add((a, 5 * tile + b))             # off-diagonal tile (0, 5)
            pick.add((nsta - 40 + a, nsta - 4 + b - 1))
    while len(pick) < count:
        i, j = sorted(rng.choice(nsta, 2, replace=False))
        pick.add((int(i), int(j)))
    return sorted(pick)


def test_config2_full_size_every_form(gpu):
    import torch
    from fastxc_amd.ops import XcPlan

    nsta, nstep, nspec, H, dt = 1000, 24, 18523, 1000, 0.1
    nfft = 2 * (nspec - 1)
    g = torch.Generator(device=gpu)
    g.manual_seed(7)
    spec = torch.randn((nstep, nsta, nspec, 2), generator=g, device=gpu, dtype=torch.float32)
    spec = torch.view_as_complex(spec).contiguous()
    pairs = _sample_pairs(nsta)
    src_all, rec_all = np.triu_indices(nsta, k=1)
    row_of = {(int(a), int(b)): r for r, (a, b) in enumerate(zip(src_all, rec_all))}
    rows = torch.tensor([row_of[p] for p in pairs], device=gpu)
    stations = sorted({s for p in pairs for s in p})
    sub = spec[:, stations, :].cpu().numpy()
    loc = {s: i for i, s in enumerate(stations)}
    ps = np.array([loc[a] for a, _ in pairs])
    pr = np.array([loc[b] for _, b in pairs])
    want = xc.correlate(sub, ps, pr, nfft, H, dt)
    for precision, tol in (("fp32", 1e-5), ("bf16x3", 1e-3), ("fp16x3", 1e-3), ("fp16", 1e-3)):
        plan = XcPlan(nspec, nstep, dt, H, precision)
        plan.set_pairs(nsta, src_all.astype(np.int32), rec_all.astype(np.int32))
        cc = plan.correlate(spec)
        got = cc[rows].cpu().numpy().astype(np.float64)
        del cc
        plan.close()
        torch.cuda.empty_cache()
        err = xc.rel_l2(got, want)
        row = np.linalg.norm(got - want, axis=1) / np.linalg.norm(want, axis=1)
        print(f"config2 {precision}: pairs={len(pairs)} rel_l2={err:.3e} max_row={row.max():.3e} "
              f"max_abs={np.abs(got - want).max():.3e} (peak {np.abs(want).max():.3e})", flush=True)
        assert err < tol and row.max() < 10 * tol
