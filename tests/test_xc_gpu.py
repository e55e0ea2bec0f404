"""xc_fast hot path on the GPU vs the FP64 oracle (oracle/xc.py).

Tolerances (relative L2 against FP64): fp32 <= 1e-5; bf16x3 <= 1e-3 (expected
~1e-5); fp16 <= 1e-3.  Row order (task order) and lag order are exact: every
row is compared with the oracle row of the same task.
"""
import numpy as np
import pytest

from oracle import xc

pytestmark = pytest.mark.gpu

TOL = {"fp32": 1e-5, "bf16x3": 1e-3, "fp16": 1e-3, "fp16x3": 1e-3, "plain": 1e-5}


def _spectra(rng, nstep, nsta, seg, nfft, dt, common=0.5):
    # band-limited noise with a shared component so cross-spectra are correlated
    base = rng.standard_normal((nstep, 1, seg))
    x = common * base + rng.standard_normal((nstep, nsta, seg))
    return (np.fft.rfft(x, n=nfft, axis=-1) * dt).astype(np.complex64)


def _all_pairs(nsta):
    src, rec = np.triu_indices(nsta, k=1)
    return src.astype(np.int32), rec.astype(np.int32)


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "fp16", "plain"])
@pytest.mark.parametrize("nsta,nstep,seg,nfft,H", [(40, 5, 800, 882 * 2, 100), (9, 24, 1500, 1764, 300),
                                                    (33, 47, 600, 1764, 200), (17, 100, 600, 1764, 200)])
def test_correlate_matches_oracle(gpu, precision, nsta, nstep, seg, nfft, H):
    import torch
    from fastxc_amd.ops import XcPlan

    rng = np.random.default_rng(nsta * 100 + nstep)
    dt = 0.1
    spec = _spectra(rng, nstep, nsta, seg, nfft, dt)
    nspec = nfft // 2 + 1
    src, rec = _all_pairs(nsta)
    plan = XcPlan(nspec, nstep, dt, H, precision)
    plan.set_pairs(nsta, src, rec)
    sd = torch.from_numpy(spec).to(gpu)
    xs = plan.cross_spectra(sd).cpu().numpy()
    want_xs = xc.cross_spectrum(spec, src, rec)
    assert xc.rel_l2(xs, want_xs) < TOL[precision]
    cc = plan.correlate(sd).cpu().numpy()
    want = xc.lags(want_xs, nfft, H, dt)
    err = xc.rel_l2(cc, want)
    print(f"precision={precision} nsta={nsta} rel_l2={err:.3e} maxabs={np.abs(cc - want).max():.3e}")
    assert err < TOL[precision]
    # per-row check: rows are in task order
    row_err = np.linalg.norm(cc - want, axis=1) / np.linalg.norm(want, axis=1)
    assert np.all(row_err < 10 * TOL[precision])


@pytest.mark.parametrize("precision", ["bf16x3", "fp16", "fp16x3"])
@pytest.mark.parametrize("nstep", [1, 12, 20, 28, 36, 44, 60, 90, 128, 200])
def test_every_operand_depth(gpu, precision, nstep):
    """One window count per compiled MFMA operand depth (kh 8, 16, 24, 32, 40,
    48, 64, 96, 128) and one beyond them (200 windows: seven chunks of 32
    accumulated in registers), two station tiles, against FP64."""
    import torch
    from fastxc_amd.ops import XcPlan

    rng = np.random.default_rng(500 + nstep)
    nsta, seg, nfft, H, dt = 37, 600, 1764, 200, 0.1
    spec = _spectra(rng, nstep, nsta, seg, nfft, dt)
    src, rec = _all_pairs(nsta)
    plan = XcPlan(nfft // 2 + 1, nstep, dt, H, precision)
    plan.set_pairs(nsta, src, rec)
    sd = torch.from_numpy(spec).to(gpu)
    xs = plan.cross_spectra(sd).cpu().numpy()
    want_xs = xc.cross_spectrum(spec, src, rec)
    err_xs = xc.rel_l2(xs, want_xs)
    cc = plan.correlate(sd).cpu().numpy()
    want = xc.lags(want_xs, nfft, H, dt)
    err = xc.rel_l2(cc, want)
    print(f"depth test {precision} nstep={nstep} xspec rel_l2={err_xs:.3e} cc rel_l2={err:.3e}")
    assert err_xs < TOL[precision] and err < TOL[precision]
    row_err = np.linalg.norm(cc - want, axis=1) / np.linalg.norm(want, axis=1)
    assert np.all(row_err < 10 * TOL[precision])


def test_production_length_fp32(gpu):
    """nfft = 37044 (2^2 3^3 7^3), H = 1000 at dt = 0.1, a handful of pairs."""
    import torch
    from fastxc_amd.ops import XcPlan

    rng = np.random.default_rng(7)
    nsta, nstep, seg, nfft, H, dt = 5, 3, 36000, 37044, 1000, 0.1
    spec = _spectra(rng, nstep, nsta, seg, nfft, dt)
    src = np.array([0, 0, 1, 3, 2, 4], np.int32)
    rec = np.array([1, 4, 3, 4, 2, 0], np.int32)
    plan = XcPlan(nfft // 2 + 1, nstep, dt, H, "fp32")
    plan.set_pairs(nsta, src, rec)
    cc = plan.correlate(torch.from_numpy(spec).to(gpu)).cpu().numpy()
    want = xc.correlate(spec, src, rec, nfft, H, dt)
    assert xc.rel_l2(cc, want) < 1e-5


def test_delayed_copy_lag_position_exact(gpu):
    import torch
    from fastxc_amd.ops import XcPlan

    rng = np.random.default_rng(3)
    seg, d, H, dt, nfft = 1500, 37, 300, 0.1, 1764
    base = rng.standard_normal(seg + d)
    x = np.stack([base[d:d + seg], base[:seg]])[None]
    spec = (np.fft.rfft(x, n=nfft, axis=-1) * dt).astype(np.complex64)
    plan = XcPlan(nfft // 2 + 1, 1, dt, H, "fp32")
    plan.set_pairs(2, np.array([0], np.int32), np.array([1], np.int32))
    cc = plan.correlate(torch.from_numpy(spec).to(gpu)).cpu().numpy()[0]
    want = xc.correlate(spec, [0], [1], nfft, H, dt)[0]
    # receiver delayed by d (rec(t) = src(t - d)): conj on the source side puts
    # the peak at lag +d, i.e. output index H + d
    assert int(np.argmax(cc)) == int(np.argmax(want)) == H + d


def test_chunked_matches_unchunked(gpu):
    import torch
    from fastxc_amd.ops import XcPlan

    rng = np.random.default_rng(11)
    nsta, nstep, seg, nfft, H, dt = 70, 4, 800, 1764, 200, 0.1
    spec = torch.from_numpy(_spectra(rng, nstep, nsta, seg, nfft, dt)).to(gpu)
    src, rec = _all_pairs(nsta)
    a = XcPlan(nfft // 2 + 1, nstep, dt, H, "fp32")
    a.set_pairs(nsta, src, rec)
    b = XcPlan(nfft // 2 + 1, nstep, dt, H, "fp32", chunk_rows=700)
    b.set_pairs(nsta, src, rec)
    assert b.pairs_info()["nchunks"] > 1
    ca = a.correlate(spec).cpu().numpy()
    cb = b.correlate(spec).cpu().numpy()
    assert np.array_equal(ca, cb)


@pytest.mark.parametrize("precision", ["bf16x3", "fp16x3", "fp16"])
def test_production_length_mfma_forms(gpu, precision):
    """The MFMA forms at nfft 37044: split contraction + the fixed-plan lag
    kernel, against FP64."""
    import torch
    from fastxc_amd.ops import XcPlan

    rng = np.random.default_rng(8)
    nsta, nstep, seg, nfft, H, dt = 40, 6, 36000, 37044, 1000, 0.1
    spec = _spectra(rng, nstep, nsta, seg, nfft, dt)
    src, rec = _all_pairs(nsta)
    plan = XcPlan(nfft // 2 + 1, nstep, dt, H, precision)
    plan.set_pairs(nsta, src, rec)
    cc = plan.correlate(torch.from_numpy(spec).to(gpu)).cpu().numpy()
    want = xc.correlate(spec, src, rec, nfft, H, dt)
    err = xc.rel_l2(cc, want)
    print(f"{precision} nfft 37044 rel_l2={err:.3e}")
    assert err < TOL[precision]


@pytest.mark.parametrize("precision", ["bf16x3", "fp16", "fp16x3", "fp32"])
def test_one_window_product_form(gpu, precision):
    """One window at nfft 37044: the stack of a pair is conj(S[src]) S[rec],
    which the lag kernel forms in FP32 from the two spectra as it reads them
    (no contraction).  Every form then meets the FP32 tolerance against FP64
    and agrees with the PLAIN form (the reference's step order through the
    cross-spectrum); pairs in both orientations, a delayed copy lands at
    exactly H + d."""
    import torch
    from fastxc_amd.ops import XcPlan

    rng = np.random.default_rng(21)
    nsta, seg, nfft, H, dt, d = 40, 36000, 37044, 1000, 0.1, 123
    spec = _spectra(rng, 1, nsta, seg, nfft, dt)
    base = rng.standard_normal(seg + d)
    spec[0, :2] = (np.fft.rfft(np.stack([base[d:d + seg], base[:seg]]), n=nfft, axis=-1)
                   * dt).astype(np.complex64)
    src, rec = _all_pairs(nsta)
    src = np.concatenate([src, rec[:50]]).astype(np.int32)  # 50 pairs reversed
    rec = np.concatenate([rec, src[:50]]).astype(np.int32)
    sd = torch.from_numpy(spec).to(gpu)
    plan = XcPlan(nfft // 2 + 1, 1, dt, H, precision)
    plan.set_pairs(nsta, src, rec)
    cc = plan.correlate(sd).cpu().numpy()
    want = xc.correlate(spec, src, rec, nfft, H, dt)
    err = xc.rel_l2(cc, want)
    print(f"{precision} one window rel_l2={err:.3e} maxabs={np.abs(cc - want).max():.3e}")
    assert err < 1e-5
    ref = XcPlan(nfft // 2 + 1, 1, dt, H, "plain")
    ref.set_pairs(nsta, src, rec)
    assert xc.rel_l2(cc, ref.correlate(sd).cpu().numpy()) < 1e-6
    assert int(np.argmax(cc[0])) == H + d


@pytest.mark.parametrize("precision", ["bf16x3", "fp32"])
def test_one_window_product_form_dt005(gpu, precision):
    """One window at nfft 74088 (dt 0.05 s): the four-step lag synthesis folds
    the product conj(S[src]) S[rec] straight from the spectra; against FP64
    and the PLAIN form, pairs in both orientations."""
    import torch
    from fastxc_amd.ops import XcPlan

    rng = np.random.default_rng(75)
    nsta, seg, nfft, H, dt = 9, 72000, 74088, 2000, 0.05
    spec = _spectra(rng, 1, nsta, seg, nfft, dt)
    src, rec = _all_pairs(nsta)
    src, rec = np.concatenate([src, rec[:5]]), np.concatenate([rec, src[:5]])
    src, rec = src.astype(np.int32), rec.astype(np.int32)
    sd = torch.from_numpy(spec).to(gpu)
    plan = XcPlan(nfft // 2 + 1, 1, dt, H, precision)
    plan.set_pairs(nsta, src, rec)
    cc = plan.correlate(sd).cpu().numpy()
    want = xc.correlate(spec, src, rec, nfft, H, dt)
    err = xc.rel_l2(cc, want)
    print(f"nfft 74088 one window {precision} rel_l2={err:.3e}")
    assert cc.shape == (len(src), 4001) and err < 1e-5
    ref = XcPlan(nfft // 2 + 1, 1, dt, H, "plain")
    ref.set_pairs(nsta, src, rec)
    assert xc.rel_l2(cc, ref.correlate(sd).cpu().numpy()) < 1e-6


@pytest.mark.parametrize("precision", ["bf16x3", "fp32"])
def test_uneven_segments_with_gaps(gpu, precision):
    """Segments of different station and window counts one after another, 10 %
    of station-windows zero-filled (gaps).  As the reference does, a gap
    still counts in the 1/nstep of its segment.  One plan per window count,
    its pair list re-set for every segment."""
    import torch
    from fastxc_amd.ops import XcPlan

    rng = np.random.default_rng(33)
    seg, nfft, H, dt = 600, 1764, 200, 0.1
    plans = {}
    for s in range(8):
        nsta = int(rng.integers(20, 61))
        nstep = int(rng.choice([6, 9, 12]))
        spec = _spectra(rng, nstep, nsta, seg, nfft, dt)
        gaps = rng.random((nstep, nsta)) < 0.1
        spec[gaps] = 0
        src, rec = _all_pairs(nsta)
        plan = plans.setdefault(nstep, XcPlan(nfft // 2 + 1, nstep, dt, H, precision))
        plan.set_pairs(nsta, src, rec)
        cc = plan.correlate(torch.from_numpy(spec).to(gpu)).cpu().numpy()
        want = xc.correlate(spec, src, rec, nfft, H, dt)
        err = xc.rel_l2(cc, want)
        assert err < TOL[precision], (s, nsta, nstep, err)


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_long_nfft_dt005(gpu, precision):
    """nfft 74088 (dt 0.05 s, one-hour windows), 4001 lags: the lag synthesis
    runs the four-step inverse transform; against FP64, lag order exact."""
    import torch
    from fastxc_amd.ops import XcPlan

    rng = np.random.default_rng(74)
    nsta, nstep, seg, nfft, H, dt = 6, 3, 72000, 74088, 2000, 0.05
    spec = _spectra(rng, nstep, nsta, seg, nfft, dt)
    src, rec = _all_pairs(nsta)
    plan = XcPlan(nfft // 2 + 1, nstep, dt, H, precision)
    plan.set_pairs(nsta, src, rec)
    cc = plan.correlate(torch.from_numpy(spec).to(gpu)).cpu().numpy()
    want = xc.correlate(spec, src, rec, nfft, H, dt)
    err = xc.rel_l2(cc, want)
    print(f"nfft 74088 {precision} rel_l2={err:.3e}")
    assert cc.shape == (len(src), 4001) and err < TOL[precision]
    row_err = np.linalg.norm(cc - want, axis=1) / np.linalg.norm(want, axis=1)
    assert np.all(row_err < 10 * TOL[precision])


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "fp16x3"])
def test_band_spec_mode(gpu, precision):
    """Spec mode fxc_xc_set_band: the stack is formed for the 16-bin chunks
    holding bins k_lo..k_hi and every other bin is 0.  Spectra that are 0
    outside the band give the parity result exactly; general spectra give the
    FP64 lags of the spectra with the bins outside those chunks zeroed; a
    parity call after a band call (and a band call after it) on the same plan
    and workspace are unaffected by the other mode."""
    import torch
    from fastxc_amd.ops import XcPlan

    rng = np.random.default_rng(31)
    nsta, nstep, seg, nfft, H, dt = 20, 6, 1500, 1764, 300, 0.1
    nspec = nfft // 2 + 1
    spec = _spectra(rng, nstep, nsta, seg, nfft, dt)
    src, rec = _all_pairs(nsta)
    k_lo, k_hi = 70, 611
    c0, c1 = k_lo // 16 * 16, -(-(k_hi + 1) // 16) * 16  # chunk-rounded band [c0, c1)
    inband = spec.copy()
    inband[..., :c0] = 0
    inband[..., c1:] = 0
    plan = XcPlan(nspec, nstep, dt, H, precision)
    plan.set_pairs(nsta, src, rec)
    parity_in = plan.correlate(torch.from_numpy(inband).to(gpu)).cpu().numpy()
    parity_gen = plan.correlate(torch.from_numpy(spec).to(gpu)).cpu().numpy()
    plan.set_band(k_lo, k_hi)
    band_in = plan.correlate(torch.from_numpy(inband).to(gpu)).cpu().numpy()
    assert np.array_equal(band_in, parity_in)
    band_gen = plan.correlate(torch.from_numpy(spec).to(gpu)).cpu().numpy()
    want = xc.lags(xc.cross_spectrum(inband, src, rec), nfft, H, dt)
    err = xc.rel_l2(band_gen, want)
    print(f"precision={precision} band rel_l2={err:.3e}")
    assert err < TOL[precision]
    plan.set_band()
    again = plan.correlate(torch.from_numpy(spec).to(gpu)).cpu().numpy()
    assert np.array_equal(again, parity_gen)
    plan.set_band(k_lo, k_hi)
    assert np.array_equal(plan.correlate(torch.from_numpy(spec).to(gpu)).cpu().numpy(), band_gen)
    plan.close()


def test_band_rejects_bad_ranges(gpu):
    from fastxc_amd.ops import XcPlan

    plan = XcPlan(883, 4, 0.1, 100, "bf16x3")
    with pytest.raises(RuntimeError):
        plan.set_band(10, 5)
    with pytest.raises(RuntimeError):
        plan.set_band(0, 883)
    plan.close()
    plain = XcPlan(883, 4, 0.1, 100, "plain")
    with pytest.raises(RuntimeError):
        plain.set_band(10, 500)
    plain.set_band()  # parity is always accepted
    plain.close()
