"""sac2spec stage on the GPU vs the FP64 oracle (oracle/sac2spec.py).

Small configuration with the reference's own length rules: dt 0.1 s,
24-minute records, 3-minute windows (seg 1800), max_lag 20 s
(nfft_out 2000), butter(2) bands; 2 stations x 3 channels so frames share
weights across channels.  Tolerance: relative L2 <= 1e-5 (FP32 vs FP64) on
the output spectra, per row as well as overall.
"""
import numpy as np
import pytest

from oracle import sac2spec as s2s
from oracle.xc import rel_l2

pytestmark = pytest.mark.gpu


def _setup(tmp_path, bands):
    path = str(tmp_path / "filter.txt")
    s2s.write_filter_file(path, 0.1, bands)
    f = s2s.read_filters(path)
    plan = s2s.plan_lengths(1440, 0.1, 180, 180, 20, f)
    return plan


def _gpu_plan(plan, nch, whiten, normalize, phase_only):
    from fastxc_amd.ops import S2sPlan

    resp = np.stack([r[: plan.nfft_filter // 2 + 1] for r in plan.responses])
    return S2sPlan(plan.npts, plan.seg, plan.shift, plan.nfft_filter, plan.nfft_out, nch,
                   plan.delta, plan.freq_low, plan.idx, resp, plan.freq_lows, whiten, normalize,
                   phase_only)


def _traces(rng, nfiles, npts):
    t = rng.standard_normal((nfiles, npts)).astype(np.float32)
    t += np.float32(0.3) * np.sin(np.arange(npts) * 0.05).astype(np.float32)  # shared component
    return t


@pytest.mark.parametrize("whiten,normalize,phase_only,bands", [
    (2, 3, 0, "0.1/2"),          # defaults: RUN-ABS, whiten after
    (0, 0, 0, "0.1/2"),          # prefilter only
    (1, 2, 0, "0.1/2"),          # whiten before, one-bit
    (3, 1, 1, "0.1/0.5 0.5/2"),  # RUN-ABS-MF, whiten both, phase-only output
    (2, 1, 0, "0.1/0.5 0.5/2"),  # RUN-ABS-MF, whiten after
])
def test_process_matches_oracle(gpu, tmp_path, whiten, normalize, phase_only, bands):
    import torch

    plan = _setup(tmp_path, bands)
    nch, nsta = 3, 2
    rng = np.random.default_rng(whiten * 10 + normalize)
    tr = _traces(rng, nch * nsta, plan.npts)
    steps = [0, 2, 3, 7]
    gp = _gpu_plan(plan, nch, whiten, normalize, phase_only)
    info = gp.info()
    assert info["nspec"] == plan.nspec
    out = gp.process(torch.from_numpy(tr).to(gpu), steps).cpu().numpy()
    win = s2s.cut_windows(tr.astype(np.float64), plan, steps)  # [nsteps, files, seg]
    frames = win.reshape(len(steps) * nsta, nch, plan.seg)
    want = s2s.process_batch(frames, plan, whiten, normalize, bool(phase_only))
    want = want.reshape(len(steps), nch * nsta, plan.nspec)
    if phase_only:
        # Phase-only output maps every bin to unit magnitude, so bins at the
        # FP32 noise floor (out of band) carry O(1) phase differences against
        # FP64.  Compare bins whose FP64 amplitude is above 1e-3 of the row
        # maximum; every other bin must still be unit or zero magnitude.
        amp = np.abs(s2s.process_batch(frames, plan, whiten, normalize, False)).reshape(want.shape)
        keep = amp > 1e-3 * amp.max(axis=-1, keepdims=True)
        mag = np.abs(out[~keep])
        assert np.all((np.abs(mag - 1) < 1e-3) | (mag == 0))
        out, want = out[keep], want[keep]
        err = rel_l2(out, want)
        print(f"phase-only compared bins: {keep.mean():.3f} rel_l2={err:.3e}")
        assert err < 1e-4
        return
    err = rel_l2(out, want)
    print(f"whiten={whiten} normalize={normalize} phase_only={phase_only} rel_l2={err:.3e}")
    assert err < 1e-5
    rows = np.linalg.norm((out - want).reshape(-1, plan.nspec), axis=1) / \
        np.linalg.norm(want.reshape(-1, plan.nspec), axis=1)
    assert np.all(rows < 1e-4)


def test_production_lengths_one_station(gpu, tmp_path):
    """seg 36000, nfft_filter 36288, nfft_out 37044 at dt = 0.1 (3 windows)."""
    import torch

    path = str(tmp_path / "filter.txt")
    s2s.write_filter_file(path, 0.1, "0.1/2")
    plan = s2s.plan_lengths(3 * 3600, 0.1, 3600, 3600, 100, s2s.read_filters(path))
    assert (plan.seg, plan.nfft_filter, plan.nfft_out) == (36000, 36288, 37044)
    rng = np.random.default_rng(9)
    tr = _traces(rng, 1, plan.npts)
    gp = _gpu_plan(plan, 1, 2, 3, 0)
    out = gp.process(torch.from_numpy(tr).to(gpu), [0, 1, 2]).cpu().numpy()
    win = s2s.cut_windows(tr.astype(np.float64), plan, [0, 1, 2])
    want = s2s.process_batch(win.reshape(3, 1, plan.seg), plan, 2, 3, False).reshape(3, 1, -1)
    assert rel_l2(out, want) < 1e-5


@pytest.mark.parametrize("whiten,normalize,bands,nch", [(2, 3, "0.1/2", 1), (3, 1, "0.1/0.5 0.5/2", 3),
                                                        (1, 2, "0.1/2", 1), (0, 0, "0.1/2", 1)])
def test_long_windows_dt005(gpu, tmp_path, whiten, normalize, bands, nch):
    """dt 0.05 s, one-hour windows: seg 72000, nfft_filter 72576, nfft_out
    74088, whose transforms exceed one workgroup's LDS and run as four-step
    transforms over HBM, against the FP64 oracle."""
    import torch

    path = str(tmp_path / "filter.txt")
    s2s.write_filter_file(path, 0.05, bands)
    plan = s2s.plan_lengths(2 * 3600, 0.05, 3600, 3600, 100, s2s.read_filters(path))
    # 72576 with the 0.1-2 Hz band; the 0.1-0.5 Hz band pads the filter FFT further
    assert (plan.seg, plan.nfft_out) == (72000, 74088) and plan.nfft_filter >= 72576
    rng = np.random.default_rng(10 + whiten + normalize)
    nsta = 2
    tr = _traces(rng, nsta * nch, plan.npts)
    gp = _gpu_plan(plan, nch, whiten, normalize, 0)
    out = gp.process(torch.from_numpy(tr).to(gpu), [0, 1]).cpu().numpy()
    win = s2s.cut_windows(tr.astype(np.float64), plan, [0, 1])
    want = s2s.process_batch(win.reshape(2 * nsta, nch, plan.seg), plan, whiten, normalize,
                             False).reshape(2, nsta * nch, -1)
    err = rel_l2(out, want)
    print(f"dt 0.05 whiten={whiten} normalize={normalize} rel_l2={err:.3e}")
    assert err < 1e-5


@pytest.mark.parametrize("whiten,normalize,phase_only", [
    (2, 3, 0), (0, 3, 0), (0, 0, 1), (1, 2, 0), (3, 3, 0), (2, 2, 1), (1, 3, 0), (3, 0, 0)])
def test_fused_production_path(gpu, tmp_path, monkeypatch, whiten, normalize, phase_only):
    """The fused one-channel path (front/back kernels on compile-time plans,
    taken at seg 36000 / 36288 / 37044 with one channel) for every whitening
    mode and the RUN-ABS, one-bit and no normalisation: against the FP64
    oracle and against the five-kernel path of the same plan (FXC_S2S_FUSED=0).
    Three full windows of two stations (a window that runs past the record
    end is zero padded, and RUN-ABS or one-bit then scale the prefilter's
    leakage in the padding to O(1) values that FP32 and FP64 round apart)."""
    import torch

    path = str(tmp_path / "filter.txt")
    s2s.write_filter_file(path, 0.1, "0.1/2")
    plan = s2s.plan_lengths(3 * 3600, 0.1, 3600, 3600, 100, s2s.read_filters(path))
    assert (plan.seg, plan.nfft_filter, plan.nfft_out) == (36000, 36288, 37044)
    rng = np.random.default_rng(40 + 4 * whiten + normalize)
    tr = _traces(rng, 2, plan.npts)
    steps = [0, 1, 2]
    gp = _gpu_plan(plan, 1, whiten, normalize, phase_only)
    out = gp.process(torch.from_numpy(tr).to(gpu), steps).cpu().numpy()
    monkeypatch.setenv("FXC_S2S_FUSED", "0")
    ref = _gpu_plan(plan, 1, whiten, normalize, phase_only)
    out5 = ref.process(torch.from_numpy(tr).to(gpu), steps).cpu().numpy()
    win = s2s.cut_windows(tr.astype(np.float64), plan, steps)
    want = s2s.process_batch(win.reshape(-1, 1, plan.seg), plan, whiten, normalize,
                             bool(phase_only)).reshape(len(steps), 2, -1)
    if phase_only:
        amp = np.abs(s2s.process_batch(win.reshape(-1, 1, plan.seg), plan, whiten, normalize,
                                       False)).reshape(want.shape)
        keep = amp > 1e-3 * amp.max(axis=-1, keepdims=True)
        err = rel_l2(out[keep], want[keep])
        print(f"fused phase-only whiten={whiten} normalize={normalize} rel_l2={err:.3e}")
        assert err < 1e-4
        assert rel_l2(out[keep], out5[keep]) < 1e-4
        return
    err = rel_l2(out, want)
    print(f"fused whiten={whiten} normalize={normalize} rel_l2={err:.3e} "
          f"vs five-kernel {rel_l2(out, out5):.3e}")
    assert err < 1e-5
    assert rel_l2(out5, want) < 1e-5


@pytest.mark.parametrize("whiten,normalize", [(2, 3), (3, 3), (1, 3)])
def test_compensated_fp32_running_means(gpu, tmp_path, whiten, normalize):
    """Spec mode fxc_s2s_set_sums(1): the running means (RUN-ABS weights,
    whitening weights) in float with Neumaier compensation instead of FP64
    sums agree with the FP64-sum output to 1e-6 relative L2 and with the FP64
    oracle to 1e-5; the five-kernel path rejects the mode."""
    import torch

    path = str(tmp_path / "filter.txt")
    s2s.write_filter_file(path, 0.1, "0.1/2")
    plan = s2s.plan_lengths(3 * 3600, 0.1, 3600, 3600, 100, s2s.read_filters(path))
    rng = np.random.default_rng(77 + whiten)
    tr = _traces(rng, 2, plan.npts)
    steps = [0, 1, 2]
    gp = _gpu_plan(plan, 1, whiten, normalize, 0)
    ref = gp.process(torch.from_numpy(tr).to(gpu), steps).cpu().numpy()
    gp.set_sums(True)
    comp = gp.process(torch.from_numpy(tr).to(gpu), steps).cpu().numpy()
    win = s2s.cut_windows(tr.astype(np.float64), plan, steps)
    want = s2s.process_batch(win.reshape(-1, 1, plan.seg), plan, whiten, normalize,
                             False).reshape(len(steps), 2, -1)
    d = rel_l2(comp, ref)
    print(f"compensated vs FP64 sums whiten={whiten}: {d:.3e}, vs oracle {rel_l2(comp, want):.3e}")
    assert d < 1e-6
    assert rel_l2(comp, want) < 1e-5
    gp.set_sums(False)
    assert np.array_equal(gp.process(torch.from_numpy(tr).to(gpu), steps).cpu().numpy(), ref)
    multi = _gpu_plan(plan, 3, whiten, normalize, 0)  # three channels: five-kernel path
    with pytest.raises(RuntimeError):
        multi.set_sums(True)
