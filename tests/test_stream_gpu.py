"""Streaming engine (fxc_stream_*) on the GPU.

Segments of one 3-minute window (dt 0.1 s, seg 1800, nfft_out 2000, 401
lags) for random subnetworks of a 12-channel array arrive one after another;
each updates every channel pair's float64 linear stack and its history of
correlations.  Checked against:
  - the FP64 oracle: sac2spec + xc of each segment, float64 mean over the
    segments a pair occurs in (relative L2 <= 1e-5 per pair, FP32 form);
  - the same engine with one lane (segments one at a time): results in the
    mix equal the results alone exactly (state updates are applied in
    admission order, so the float64 sums add in the same order);
  - PWS / TF-PWS of the history against fxc_pws / fxc_tfpws on the same
    correlations prepared on the host as the reference does (1e-5 / 1e-4).
"""
import numpy as np
import pytest

from oracle import sac2spec as s2s
from oracle import xc as oxc
from oracle.xc import rel_l2

pytestmark = pytest.mark.gpu

NCHAN = 12


def _plan(tmp_path):
    path = str(tmp_path / "filter.txt")
    s2s.write_filter_file(path, 0.1, "0.1/2")
    f = s2s.read_filters(path)
    return s2s.plan_lengths(180, 0.1, 180, 180, 20, f)


def _engine(plan, lanes, history=0, precision="fp32", max_pairs=1 << 12, max_seg_chan=NCHAN,
            pws_substack=0, H=200):
    from fastxc_amd.ops import StreamEngine

    resp = np.stack([r[: plan.nfft_filter // 2 + 1] for r in plan.responses])
    return StreamEngine(plan.npts, plan.seg, plan.shift, plan.nfft_filter, plan.nfft_out, 1,
                        plan.delta, plan.freq_low, plan.idx, resp, plan.freq_lows, NCHAN, H,
                        precision=precision, history=history, max_pairs=max_pairs,
                        max_seg_chan=max_seg_chan, lanes=lanes, pws_substack=pws_substack)


def _segments(seed, count):
    rng = np.random.default_rng(seed)
    segs = []
    for _ in range(count):
        n = int(rng.integers(3, NCHAN + 1))
        chan = rng.choice(NCHAN, size=n, replace=False).astype(np.int32)
        tr = rng.standard_normal((n, 1800)).astype(np.float32)
        tr += np.float32(0.4) * np.sin(np.arange(1800) * 0.03 + chan[:, None]).astype(np.float32)
        segs.append((chan, tr))
    return segs


def _run(eng, segs):
    tickets = [eng.submit(c, t) for c, t in segs]
    return [eng.wait(t) for t in tickets]


def _oracle_ccfs(plan, segs, H=200):
    """{(a, b): [ccf per segment, submission order]} in FP64."""
    out = {}
    for chan, tr in segs:
        win = s2s.cut_windows(tr.astype(np.float64), plan)          # [1, n, seg]
        spec = s2s.process_batch(win.reshape(len(chan), 1, plan.seg), plan, 2, 3, False)
        spec = spec.reshape(1, len(chan), plan.nspec)
        order = np.argsort(chan)
        for i in range(len(order)):
            for j in range(i + 1, len(order)):
                a, b = order[i], order[j]
                cc = oxc.correlate(spec, [a], [b], plan.nfft_out, H, plan.delta)[0]
                out.setdefault((int(chan[a]), int(chan[b])), []).append(cc)
    return out


def test_linear_stack_matches_oracle_and_mix_equals_alone(gpu, tmp_path):
    plan = _plan(tmp_path)
    segs = _segments(1, 14)
    want = _oracle_ccfs(plan, segs)
    keys = sorted(want)
    src = [k[0] for k in keys]
    rec = [k[1] for k in keys]
    mix = _engine(plan, lanes=4)
    lat = _run(mix, segs)
    assert all(x > 0 for x in lat)
    got, cnt = mix.read(src, rec, "linear")
    got = got.cpu().numpy()
    info = mix.info()
    assert info["pairs_used"] == len(keys) and info["segments_done"] == len(segs)
    for i, k in enumerate(keys):
        assert cnt[i] == len(want[k])
        ref = np.mean(want[k], axis=0)
        err = rel_l2(got[i], ref)
        assert err < 1e-5, (k, err)
    alone = _engine(plan, lanes=1)
    _run(alone, segs)
    got1, cnt1 = alone.read(src, rec, "linear")
    assert np.array_equal(cnt, cnt1)
    assert np.array_equal(got, got1.cpu().numpy())
    mix.close()
    alone.close()


@pytest.mark.parametrize("kind,B,tol", [("pws", 2, 1e-5), ("tfpws", 1, 1e-4)])
def test_history_stacks_match_batch_kernels(gpu, tmp_path, kind, B, tol):
    import torch
    from fastxc_amd import ops

    plan = _plan(tmp_path)
    segs = _segments(2, 16)
    # 4 lanes and a 4-record ring: a slow lane must not let a newer record
    # of a pair land before an older one in the same ring slot
    eng = _engine(plan, lanes=4, history=4)
    _run(eng, segs)
    # the same correlations computed segment by segment through the batch
    # entry points are the records the reference's stacks would read
    recs = {}
    xp = ops.XcPlan(plan.nspec, 1, plan.delta, 200, "fp32")
    sp = ops.S2sPlan(plan.npts, plan.seg, plan.shift, plan.nfft_filter, plan.nfft_out, 1,
                     plan.delta, plan.freq_low, plan.idx,
                     np.stack([r[: plan.nfft_filter // 2 + 1] for r in plan.responses]),
                     plan.freq_lows, 2, 3, False)
    for chan, tr in segs:
        spec = sp.process(torch.from_numpy(tr).to(gpu), [0])
        order = np.argsort(chan)
        pa = [(order[i], order[j]) for i in range(len(order)) for j in range(i + 1, len(order))]
        xp.set_pairs(len(chan), [p[0] for p in pa], [p[1] for p in pa])
        cc = xp.correlate(spec).cpu().numpy()
        for (a, b), row in zip(pa, cc):
            recs.setdefault((int(chan[a]), int(chan[b])), []).append(row)
    keys = sorted(k for k in recs if len(recs[k]) >= 2)
    got, cnt = eng.read([k[0] for k in keys], [k[1] for k in keys], kind, substack=B)
    got = got.cpu().numpy()
    # host preparation of the last 4 records (submission order), as the reference does
    pre, w, lin, off = [], [], [], [0]
    for k in keys:
        r = np.stack(recs[k][-4:]).astype(np.float32)
        T = r.shape[0]
        lsum = np.zeros(r.shape[1], np.float32)
        for a in range(0, T, B):
            g = r[a:a + B]
            s = np.zeros(r.shape[1], np.float32)
            for row in g:
                s += row
            if kind == "pws":
                s = s * (np.float32(1) / np.float32(len(g)))
                wg = np.float32(len(g)) / np.float32(T)
                lsum += s * wg
            else:
                wg = np.float32(len(g)) * (np.float32(1) / np.float32(T))
                lsum += s * (np.float32(1) / np.float32(T))
            pre.append(s)
            w.append(wg)
        lin.append(lsum)
        off.append(len(w))
    d = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a, np.float32))).to(gpu)
    fn = ops.pws if kind == "pws" else ops.tfpws
    want = fn(d(pre), d(w), np.array(off), d(np.stack(lin))).cpu().numpy()
    for i, k in enumerate(keys):
        assert cnt[i] == len(recs[k])
        err = rel_l2(got[i], want[i])
        assert err < tol, (k, err)
    eng.close()


def test_refuses_segment_beyond_pair_state(gpu, tmp_path):
    from fastxc_amd._lib import FxcError

    plan = _plan(tmp_path)
    eng = _engine(plan, lanes=2, max_pairs=10)
    rng = np.random.default_rng(3)
    big = np.arange(6, dtype=np.int32)                      # 15 pairs > 10 slots
    t = eng.submit(big, rng.standard_normal((6, 1800)).astype(np.float32))
    with pytest.raises(FxcError, match="pair state full"):
        eng.wait(t)
    small = np.array([7, 3, 9], dtype=np.int32)             # 3 pairs fit
    t = eng.submit(small, rng.standard_normal((3, 1800)).astype(np.float32))
    assert eng.wait(t) > 0
    assert eng.info()["pairs_used"] == 3
    with pytest.raises(FxcError, match="duplicate channel"):
        eng.submit(np.array([1, 1], np.int32), np.zeros((2, 1800), np.float32))
    with pytest.raises(FxcError, match="out of range"):
        eng.submit(np.array([1, NCHAN], np.int32), np.zeros((2, 1800), np.float32))
    eng.close()


@pytest.mark.parametrize("B", [1, 3])
def test_running_pws_and_history_tfpws_match_fp64_oracle(gpu, tmp_path, B):
    """The engine's running PWS over every record of a pair (sub-stacks of B in
    submission order; power 1 and the spec mode's power 2) and its TF-PWS over
    the history ring, against oracle/stack.py over the FP64 oracle's
    correlations of the same segments (channels submitted out of id order)."""
    from oracle import stack as ostack

    plan = _plan(tmp_path)
    segs = _segments(20 + B, 12)
    want = _oracle_ccfs(plan, segs)
    keys = sorted(k for k in want if len(want[k]) >= 2)
    src = [k[0] for k in keys]
    rec = [k[1] for k in keys]
    hist = 4
    eng = _engine(plan, lanes=4, history=hist, pws_substack=B)
    _run(eng, segs)
    for power in (1.0, 2.0):
        got, cnt = eng.read(src, rec, "pws", substack=B, power=power)
        got = got.cpu().numpy()
        for i, k in enumerate(keys):
            recs = np.stack(want[k])
            assert cnt[i] == len(recs)
            g, w, lin = ostack.pws_prepare(recs, B)
            err = rel_l2(got[i], ostack.pws(g, w, lin, power=power))
            assert err < 1e-4, (k, power, err)
    got, _ = eng.read(src, rec, "tfpws", substack=B)
    got = got.cpu().numpy()
    for i, k in enumerate(keys):
        recs = np.stack(want[k])[-hist:]
        s_, w, lin = ostack.tfpws_prepare(recs, B)
        err = rel_l2(got[i], ostack.tfpws(s_, w, lin, delta=plan.delta))
        assert err < 1e-3, (k, err)
    eng.close()


@pytest.mark.parametrize("precision", ["bf16x3", "fp32"])
def test_production_length_segments_match_oracle(gpu, tmp_path, precision):
    """One-hour segments at dt 0.1 s (seg 36000, output FFT 37044, 2001 lags):
    the lengths of the fused sac2spec path and of xc's one-window product
    form.  Linear and running PWS (sub-stacks of one record) read-outs
    against the FP64 oracle's correlations of the same segments."""
    from oracle import stack as ostack

    path = str(tmp_path / "filter.txt")
    s2s.write_filter_file(path, 0.1, "0.1/2")
    plan = s2s.plan_lengths(3600, 0.1, 3600, 3600, 100, s2s.read_filters(path))
    assert plan.seg == 36000 and plan.nfft_out == 37044
    H = 1000
    rng = np.random.default_rng(31)
    segs = []
    for _ in range(4):
        n = int(rng.integers(3, 6))
        chan = rng.choice(NCHAN, size=n, replace=False).astype(np.int32)
        tr = rng.standard_normal((n, 36000)).astype(np.float32)
        tr += np.float32(0.4) * np.sin(np.arange(36000) * 0.03 + chan[:, None]).astype(np.float32)
        segs.append((chan, tr))
    want = _oracle_ccfs(plan, segs, H)
    keys = sorted(want)
    src = [k[0] for k in keys]
    rec = [k[1] for k in keys]
    eng = _engine(plan, lanes=2, precision=precision, max_seg_chan=6, pws_substack=1, H=H)
    _run(eng, segs)
    got, cnt = eng.read(src, rec, "linear")
    got = got.cpu().numpy()
    for i, k in enumerate(keys):
        assert cnt[i] == len(want[k])
        err = rel_l2(got[i], np.mean(want[k], axis=0))
        assert err < 1e-5, (k, err)
    got, _ = eng.read(src, rec, "pws", substack=1)
    got = got.cpu().numpy()
    for i, k in enumerate(keys):
        g, w, lin = ostack.pws_prepare(np.stack(want[k]), 1)
        err = rel_l2(got[i], ostack.pws(g, w, lin))
        assert err < 1e-4, (k, err)
    eng.close()
