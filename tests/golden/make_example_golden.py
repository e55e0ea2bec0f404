"""Writes the example fixtures under tests/golden/example/.

Inputs: a subset of the reference's bundled example (example/data in the
reference tree): stations A7K2 and B4Q8, components E, N, Z, days 2022.138
and 2022.139 at 1 Hz, copied byte for byte as SAC files, and the filter file
of the example's bands (example/config.ini: 0.2/0.4 and 0.1/0.2 Hz, so the
broad band 0.1/0.4 first), written as the reference's design_filter writes it.

Expected outputs: the FP64 oracle (oracle/) run on those inputs with the
example's settings -- sac_len 86400, win_len 3600, shift AUTO (= win_len),
RUN-ABS-MF (normalize AUTO with two bands), whitening AFTER, max_lag 100 s
(201 lags at 1 Hz), autocorrelations off -- then the linear stack over the
two days and PWS / TF-PWS over the days with sub-stacks of one record
(pre_stack_size 10 would put both days in one sub-stack, where PWS equals
the linear stack).  Saved in golden.npz: per day the nine component-pair
correlations of the pair A7K2-B4Q8 in the reference's task order, and the
three stacks.

Day 2022.137 is left out: B4Q8 records zeros there for 4.8 hours (and C9M3
for 3.5), and a window that is part data, part zeros hands RUN-ABS-MF the
prefilter's decaying tail, which it scales up to O(1) from values near
FP32 rounding -- no FP32 path (the reference's included) matches FP64 there.

Run in a tree that holds the reference at /root/reference:
    python tests/golden/make_example_golden.py
"""
from __future__ import annotations

import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

REF_DATA = "/root/reference/example/data"
STATIONS = ["A7K2", "B4Q8"]
DAYS = ["138", "139"]
COMPS = ["E", "N", "Z"]
BANDS = "0.2/0.4 0.1/0.2"
OUT = os.path.join(HERE, "example")


def sac_path(root, sta, day, comp):
    return os.path.join(root, sta, "2022", f"{sta}.2022.{day}.0000.{comp}.sac")


def copy_inputs():
    for sta in STATIONS:
        for day in DAYS:
            for c in COMPS:
                dst = sac_path(os.path.join(OUT, "data"), sta, day, c)
                os.makedirs(os.path.dirname(dst), exist_ok=True)
                shutil.copyfile(sac_path(REF_DATA, sta, day, c), dst)


def read_sac_f32(path, npts):
    with open(path, "rb") as f:
        b = f.read()
    n = int(np.frombuffer(b, dtype="<i4", count=1, offset=280 + 9 * 4)[0])
    x = np.zeros(npts, np.float32)
    m = min(n, npts)
    x[:m] = np.frombuffer(b, dtype="<f4", count=m, offset=632)
    return x


def oracle_outputs(data_root, filter_path):
    """Correlations per day and the stacks, FP64 oracle."""
    from oracle import sac2spec as s2s
    from oracle import stack
    from oracle import xc

    plan = s2s.plan_lengths(86400, 1.0, 3600, 3600, 100, s2s.read_filters(filter_path))
    shape = xc.finalize_shape(plan.nspec, plan.nstep, 1.0, 100.0)
    nch = len(COMPS)
    gnsl = [s + 1 for s in range(len(STATIONS)) for _ in COMPS]
    allowed = {1 * 10000 + 2}
    job = xc.build_jobs(len(gnsl), len(gnsl))[0]
    tasks, _ = xc.build_tasks_for_job(gnsl, allowed, job)
    keys = [(gnsl[t.src_meta], gnsl[t.rec_meta], COMPS[t.src_meta % nch], COMPS[t.rec_meta % nch])
            for t in tasks]
    days = []
    for day in DAYS:
        tr = np.stack([read_sac_f32(sac_path(data_root, sta, day, c), plan.npts)
                       for sta in STATIONS for c in COMPS]).astype(np.float64)
        win = s2s.cut_windows(tr, plan)
        steps, files, seg = win.shape
        spec = s2s.process_batch(win.reshape(steps * files // nch, nch, seg), plan, whiten=2,
                                 normalize=1).reshape(steps, files, -1)
        src = np.array([t.src_meta for t in tasks])
        rec = np.array([t.rec_meta for t in tasks])
        days.append(xc.correlate(spec, src, rec, shape.nfft, shape.half_cc, 1.0))
    days = np.stack(days)                        # [day, task, lag]
    lin = days.mean(axis=0)
    pws, tfpws = [], []
    for t in range(len(tasks)):
        g, w, l = stack.pws_prepare(days[:, t], 1)
        pws.append(stack.pws(g, w, l))
        s, w2, l2 = stack.tfpws_prepare(days[:, t], 1)
        tfpws.append(stack.tfpws(s, w2, l2, delta=1.0))
    return plan, keys, days, lin, np.stack(pws), np.stack(tfpws)


def main():
    from oracle import sac2spec as s2s

    copy_inputs()
    filt = os.path.join(OUT, "filter.txt")
    s2s.write_filter_file(filt, 1.0, BANDS)
    plan, keys, days, lin, pws, tfpws = oracle_outputs(os.path.join(OUT, "data"), filt)
    np.savez_compressed(os.path.join(OUT, "golden.npz"), keys=np.array(keys, dtype=object).astype(str),
                        xc=days, linear=lin, pws=pws, tfpws=tfpws,
                        plan=np.array([plan.npts, plan.seg, plan.nfft_filter, plan.nfft_out,
                                       plan.nspec, plan.nstep]))
    print(f"wrote {OUT}: {len(keys)} records x {days.shape[0]} days x {days.shape[2]} lags; "
          f"plan npts={plan.npts} seg={plan.seg} nfft_filter={plan.nfft_filter} "
          f"nfft_out={plan.nfft_out}")


if __name__ == "__main__":
    main()
