"""Batched uneven segments (fxc_xc_batch_*) on the GPU vs the FP64 oracle.

Each test builds segments of different station and window counts with 10 %
of station-windows marked as gaps.  Expected rows per segment:
  gaps "keep":    oracle/xc.py correlate on the segment with its gaps zeroed
                  (the reference divides by the window count);
  gaps "exclude": oracle/xc.py lags(cross_spectrum_valid(...)).
Tolerance: split-BF16 <= 1e-3 relative L2 against FP64 per segment (expected
~1e-5).  Rows keep the batch order (segments back to back, pairs (i, j), i < j,
row-major), so every row is compared with the oracle row of the same pair.
"""
import numpy as np
import pytest

from oracle import xc

pytestmark = pytest.mark.gpu

TOL = 1e-3


def _operand_depth(nstep):
    """The MFMA operand depth a window count compiles to: 8, multiples of 8
    up to 48, then 64, 96, 128 (fxc_xc.hip xc_operand_depth)."""
    if nstep <= 48:
        return 8 if nstep <= 8 else (nstep + 7) // 8 * 8
    return 64 if nstep <= 64 else 96 if nstep <= 96 else 128


def _spectra(rng, nstep, nsta, seg, nfft, dt, common=0.5):
    base = rng.standard_normal((nstep, 1, seg))
    x = common * base + rng.standard_normal((nstep, nsta, seg))
    return (np.fft.rfft(x, n=nfft, axis=-1) * dt).astype(np.complex64)


def _segments(rng, nseg, seg, nfft, dt, sta_range=(2, 45), steps=(3, 6, 17, 24, 40, 47, 70)):
    nsta = [int(rng.integers(*sta_range)) for _ in range(nseg)]
    nsta[1] = 1  # a segment without pairs
    nstep = [int(rng.choice(steps)) for _ in range(nseg)]
    specs = [_spectra(rng, w, n, seg, nfft, dt) for n, w in zip(nsta, nstep)]
    valid = [(rng.random((w, n)) >= 0.1).astype(np.uint8) for n, w in zip(nsta, nstep)]
    return nsta, nstep, specs, valid


def _expected(spec, valid, gaps, nfft, H, dt):
    n = spec.shape[1]
    src, rec = np.triu_indices(n, k=1)
    if gaps == "keep":
        z = spec.copy()
        z[valid == 0] = 0
        return xc.correlate(z, src, rec, nfft, H, dt)
    return xc.lags(xc.cross_spectrum_valid(spec, valid, src, rec), nfft, H, dt)


def _check_segments(batch, cc, specs, valid, gaps, nfft, H, dt):
    for s, (sp, v) in enumerate(zip(specs, valid)):
        rows = cc[batch.segment_rows(s)]
        want = _expected(sp, v, gaps, nfft, H, dt)
        assert rows.shape == want.shape, (s, rows.shape, want.shape)
        if want.size:
            err = xc.rel_l2(rows, want)
            assert err < TOL, (s, batch.nsta[s], batch.nstep[s], err)


@pytest.mark.parametrize("gaps", ["keep", "exclude"])
@pytest.mark.parametrize("grouping", [0, 1])
def test_batch_matches_oracle(gpu, gaps, grouping):
    import torch
    from fastxc_amd.ops import XcBatch

    rng = np.random.default_rng(51 + grouping)
    seg, nfft, H, dt = 600, 1764, 200, 0.1
    nsta, nstep, specs, valid = _segments(rng, 9, seg, nfft, dt)
    b = XcBatch(nfft // 2 + 1, dt, H, nsta, nstep, gaps=gaps)
    b.set_grouping(grouping)
    spec = XcBatch.pack([torch.from_numpy(x) for x in specs], gpu)
    vmask = XcBatch.pack([torch.from_numpy(v) for v in valid], gpu)
    cc = b.correlate(spec, vmask).cpu().numpy()
    info = b.info()
    assert info["grouping"] == grouping
    assert info["groups"] == (1 if grouping == 0 else len({_operand_depth(w) for w in nstep}))
    _check_segments(b, cc, specs, valid, gaps, nfft, H, dt)
    b.close()


def test_batch_segment_equals_plan_alone(gpu):
    """Each segment's rows equal fxc_xc_correlate of that segment alone
    (gaps zero-filled, gap mode keep) to float rounding of the row scale."""
    import torch
    from fastxc_amd.ops import XcBatch, XcPlan

    rng = np.random.default_rng(57)
    seg, nfft, H, dt = 600, 1764, 200, 0.1
    nsta, nstep, specs, valid = _segments(rng, 6, seg, nfft, dt)
    zs = []
    for sp, v in zip(specs, valid):
        z = sp.copy()
        z[v == 0] = 0
        zs.append(z)
    b = XcBatch(nfft // 2 + 1, dt, H, nsta, nstep, gaps="keep")
    cc = b.correlate(XcBatch.pack([torch.from_numpy(z) for z in zs], gpu)).cpu().numpy()
    for s, z in enumerate(zs):
        if nsta[s] < 2:
            continue
        p = XcPlan(nfft // 2 + 1, nstep[s], dt, H, "bf16x3")
        src, rec = np.triu_indices(nsta[s], k=1)
        p.set_pairs(nsta[s], src, rec)
        alone = p.correlate(torch.from_numpy(z).to(gpu)).cpu().numpy()
        p.close()
        err = xc.rel_l2(cc[b.segment_rows(s)], alone)
        assert err < 1e-6, (s, nsta[s], nstep[s], err)
    b.close()


def test_batch_tune_keeps_results(gpu):
    import torch
    from fastxc_amd.ops import XcBatch

    rng = np.random.default_rng(61)
    seg, nfft, H, dt = 600, 1764, 200, 0.1
    nsta, nstep, specs, valid = _segments(rng, 8, seg, nfft, dt)
    b = XcBatch(nfft // 2 + 1, dt, H, nsta, nstep, gaps="exclude")
    spec = XcBatch.pack([torch.from_numpy(x) for x in specs], gpu)
    vmask = XcBatch.pack([torch.from_numpy(v) for v in valid], gpu)
    t = b.tune(spec, vmask, reps=1)
    assert t["grouping"] in (0, 1) and all(ms > 0 for ms in t["ms"])
    assert t["grouping"] == (1 if t["ms"][1] < t["ms"][0] else 0)
    cc = b.correlate(spec, vmask).cpu().numpy()
    _check_segments(b, cc, specs, valid, "exclude", nfft, H, dt)
    b.close()


def test_batch_graph_replays_match_eager(gpu):
    """A captured batch replayed on new data in the same buffers: every replay
    equals the eager run on that data, with eager work and synchronisation
    between the replays."""
    import torch
    from fastxc_amd.ops import XcBatch

    rng = np.random.default_rng(67)
    seg, nfft, H, dt = 600, 1764, 200, 0.1
    nsta, nstep, specs, valid = _segments(rng, 7, seg, nfft, dt)
    b = XcBatch(nfft // 2 + 1, dt, H, nsta, nstep, gaps="exclude")
    spec = XcBatch.pack([torch.from_numpy(x) for x in specs], gpu)
    vmask = XcBatch.pack([torch.from_numpy(v) for v in valid], gpu)
    out = torch.empty((b.ntasks, b.cc_size), dtype=torch.float32, device=gpu)
    b.capture(spec, vmask, out)
    torch.cuda.synchronize()
    first = out.cpu().numpy()
    _check_segments(b, first, specs, valid, "exclude", nfft, H, dt)
    eager_out = torch.empty_like(out)
    for r in range(3):
        new = [_spectra(rng, w, n, seg, nfft, dt) for n, w in zip(nsta, nstep)]
        newv = [(rng.random((w, n)) >= 0.1).astype(np.uint8) for n, w in zip(nsta, nstep)]
        spec.copy_(XcBatch.pack([torch.from_numpy(x) for x in new], gpu))
        vmask.copy_(XcBatch.pack([torch.from_numpy(v) for v in newv], gpu))
        b.replay()
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        b.correlate(spec, vmask, out=eager_out)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(got, eager_out.cpu().numpy(), err_msg=f"replay {r}")
        _check_segments(b, got, new, newv, "exclude", nfft, H, dt)
    b.close()


def test_batch_production_length(gpu):
    """nfft 37044 (the compile-time lag plan), three segments with 24, 6 and 47
    windows; a sample of rows of each against the oracle."""
    import torch
    from fastxc_amd.ops import XcBatch

    rng = np.random.default_rng(71)
    seg, nfft, H, dt = 36000, 37044, 1000, 0.1
    nsta, nstep = [20, 7, 12], [24, 6, 47]
    specs = [_spectra(rng, w, n, seg, nfft, dt) for n, w in zip(nsta, nstep)]
    valid = [(rng.random((w, n)) >= 0.1).astype(np.uint8) for n, w in zip(nsta, nstep)]
    b = XcBatch(nfft // 2 + 1, dt, H, nsta, nstep, gaps="exclude")
    spec = XcBatch.pack([torch.from_numpy(x) for x in specs], gpu)
    vmask = XcBatch.pack([torch.from_numpy(v) for v in valid], gpu)
    cc = b.correlate(spec, vmask).cpu().numpy()
    for s, (sp, v) in enumerate(zip(specs, valid)):
        src, rec = np.triu_indices(nsta[s], k=1)
        pick = np.linspace(0, len(src) - 1, min(len(src), 12)).round().astype(int)
        want = xc.lags(xc.cross_spectrum_valid(sp, v, src[pick], rec[pick]), nfft, H, dt)
        err = xc.rel_l2(cc[b.segment_rows(s)][pick], want)
        assert err < TOL, (s, err)
    b.close()


@pytest.mark.parametrize("power", [1.0, 2.0])
def test_batch_pws_leg_matches_oracle(gpu, power):
    """The batch's PWS leg (ops.batch_pws, FP32): every pair's records over
    the segments containing it, stacked as the reference's PWS with
    sub-stacks of one record (and the spec mode's power 2), against
    oracle/stack.py over the FP64 oracle's correlations of each segment."""
    import torch
    from fastxc_amd.ops import XcBatch, batch_pws
    from oracle import stack

    rng = np.random.default_rng(61)
    seg, nfft, H, dt = 600, 1764, 200, 0.1
    nsta = [12, 9, 15, 6, 12, 11]
    nstep = [6, 17, 24, 6, 40, 9]
    specs = [_spectra(rng, w, n, seg, nfft, dt) for n, w in zip(nsta, nstep)]
    b = XcBatch(nfft // 2 + 1, dt, H, nsta, nstep, gaps="keep")
    spec = XcBatch.pack([torch.from_numpy(x) for x in specs], gpu)
    cc = b.correlate(spec)
    src, rec, out = batch_pws(b, cc, power=power)
    out = out.cpu().numpy()
    want_seg = [_expected(sp, np.ones(sp.shape[:2], np.uint8), "keep", nfft, H, dt) for sp in specs]
    assert len(src) == max(nsta) * (max(nsta) - 1) // 2
    worst = 0.0
    for p, (i, j) in enumerate(zip(src, rec)):
        recs = []
        for s, n in enumerate(nsta):
            if j < n:
                r = i * n - i * (i + 1) // 2 + (j - i - 1)
                recs.append(want_seg[s][r])
        g, w, lin = stack.pws_prepare(np.stack(recs), 1)
        err = xc.rel_l2(out[p], stack.pws(g, w, lin, power=power))
        worst = max(worst, err)
        assert err < 1e-4, (i, j, len(recs), err)
    print(f"batch pws power={power} pairs={len(src)} worst rel_l2={worst:.3e}")
    b.close()


def test_batch_autotune_searches_and_caches(gpu, tmp_path):
    """The tuner times grouping x row chunk on the data, keeps results
    identical to the untuned batch, and a second batch of the same class
    takes the cached choice without timing."""
    import torch
    from fastxc_amd.ops import XcBatch

    rng = np.random.default_rng(71)
    seg, nfft, H, dt = 600, 1764, 200, 0.1
    nsta, nstep, specs, valid = _segments(rng, 6, seg, nfft, dt)
    b = XcBatch(nfft // 2 + 1, dt, H, nsta, nstep, gaps="exclude")
    spec = XcBatch.pack([torch.from_numpy(x) for x in specs], gpu)
    vmask = XcBatch.pack([torch.from_numpy(v) for v in valid], gpu)
    ref = b.correlate(spec, vmask).clone()
    cache = str(tmp_path / "tune.json")
    res = b.autotune(spec, vmask, rows=(64, 4096), cache=cache)
    assert not res["cached"] and len(res["trials"]) == 4
    assert (res["grouping"], res["max_rows"]) == (b.info()["grouping"], b.max_rows)
    cc = b.correlate(spec, vmask)
    assert xc.rel_l2(cc.cpu().numpy(), ref.cpu().numpy()) < 1e-6
    _check_segments(b, cc.cpu().numpy(), specs, valid, "exclude", nfft, H, dt)
    b2 = XcBatch(nfft // 2 + 1, dt, H, nsta, nstep, gaps="exclude")
    res2 = b2.autotune(spec, vmask, rows=(64, 4096), cache=cache)
    assert res2["cached"] and (res2["grouping"], res2["max_rows"]) == (res["grouping"], res["max_rows"])
    b.close()
    b2.close()
