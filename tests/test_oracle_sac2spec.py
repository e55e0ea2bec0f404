"""Pins the sac2spec oracle (CPU only): closed forms, brute force, scipy and
the reference's own C helpers compiled into oracle/_ref."""
import os

import numpy as np
import pytest
from scipy.signal import freqz

from oracle import cbind
from oracle import sac2spec as s2s


@pytest.fixture(scope="module")
def filt_cfg2(tmp_path_factory):
    path = str(tmp_path_factory.mktemp("f") / "filter.txt")
    s2s.write_filter_file(path, 0.1, "0.1/2")
    return path


def test_config2_plan_closed_forms(filt_cfg2):
    f = s2s.read_filters(filt_cfg2)
    assert len(f) == 2  # broad band + the configured band
    p = s2s.plan_lengths(86400, 0.1, 3600, 3600, 100, f)
    assert (p.npts, p.seg, p.nstep, p.nfft_out, p.nspec) == (864000, 36000, 24, 37044, 18523)
    assert p.nfft_filter == 36288
    assert p.idx == (240, 360, 7200, 9597)
    assert s2s.runabs_window(float(np.float32(np.float32(0.1) * np.float32(0.667))), 0.1) == 299
    assert s2s.whiten_window(36000, 0.1) == 72
    # 50 % overlap gives 47 windows a day
    assert s2s.plan_lengths(86400, 0.1, 3600, 1800, 100, f).nstep == 47
    # dt = 0.05: nfft_out 74088, nspec 37045
    assert s2s.plan_lengths(86400, 0.05, 3600, 3600, 100, f).nfft_out == 74088


@pytest.mark.skipif(not cbind.have_ref(), reason="reference helpers not built (make -C oracle ref)")
def test_lengths_and_filters_match_reference_c(filt_cfg2):
    ref = cbind.ref_s2s()
    for n in list(range(1, 400)) + [36000, 36001, 37000, 72001, 99991]:
        assert ref.next_even_smooth_2357_length(n) == s2s.next_even_smooth_2357(n), n
    for secs, dt, mx in [(3600, 0.1, 864000), (3600, 0.05, 1728000), (7200, 1.0, 3000),
                         (123.4, 0.1, 5000), (9000, 1.0, 3000)]:
        assert ref.segment_length_from_seconds(secs, mx, dt) == \
            s2s.segment_length_from_seconds(secs, mx, dt)
    f = s2s.read_filters(filt_cfg2)
    cnt, pad, resp = cbind.ref_filters(filt_cfg2, 36288, 36000)
    assert cnt == len(f)
    assert pad == s2s.filter_padding(f, 36000)
    mine = np.stack([s2s.power_response(x, 36288)[: 36288 // 2 + 1] for x in f])
    assert np.max(np.abs(mine - resp) / np.maximum(np.abs(resp), 1e-30)) < 1e-6


def test_power_response_matches_freqz(filt_cfg2):
    f = s2s.read_filters(filt_cfg2)[0]
    n = 4000
    w = np.pi * np.arange(n // 2 + 1) / (n // 2)
    _, h = freqz(f.b, f.a, worN=w)
    got = s2s.power_response(f, n)[: n // 2 + 1]
    assert np.allclose(got, np.abs(h) ** 2, rtol=1e-6, atol=1e-12)


def test_preprocess_removes_ramp():
    W = 500
    x = (3.0 + 0.01 * np.arange(W))[None, None, :]
    y = s2s.preprocess(x, 0.5, 0.1)
    assert np.max(np.abs(y)) < 1e-9
    # taper closed form: T = (size_t)(2*(1/0.5)/0.1) = 40
    t = s2s.time_taper(W, 40)
    assert t[0] == 0.0 and t[20] == 0.5 * (1 - np.cos(np.pi * 20 / 40)) and t[100] == 1.0
    assert t[W - 20] == 0.5 * (1 - np.cos(np.pi * 20 / 40))


@pytest.mark.parametrize("W,win", [(10, 3), (10, 4), (17, 1), (9, 15), (50, 8)])
def test_smooth_rows_bruteforce(W, win):
    rng = np.random.default_rng(W + win)
    x = rng.standard_normal((2, W))
    got = s2s.smooth_rows(x, win)
    want = np.zeros_like(x)
    for r in range(2):
        for c in range(W):
            start = c - win // 2
            want[r, c] = sum(x[r, i] for i in range(start, start + win) if 0 <= i < W) / win
    assert np.allclose(got, want)


def test_runabs_constant_and_onebit():
    x = np.full((1, 1, 400), 2.5)
    y = s2s.runabs(x, 0.25, 0.1)  # W = 2*int(1/0.025)+1 = 81
    assert np.allclose(y[..., 40:-40], 1.0)
    # two identical channels share the weight: each normalises to 1/2
    y2 = s2s.runabs(np.concatenate([x, x], axis=1), 0.25, 0.1)
    assert np.allclose(y2[..., 40:-40], 0.5)
    assert np.array_equal(s2s.onebit(np.array([-2.0, 0.0, 3.0])), [-1.0, 0.0, 1.0])


def test_whiten_flat_spectrum_gives_taper():
    class P:
        seg = 1000
        delta = 0.1
        idx = (20, 30, 200, 260)
    X = np.full((1, 1, 501), 0.1 + 0j)
    out = s2s.whiten_spectrum(X, P)
    win = s2s.whiten_window(1000, 0.1)  # int(0.02*1000*0.1) = 2
    assert win == 2
    taper = s2s.spectral_taper(501, P.idx)
    # interior bins: smoothed |X| equals |X| so X/w = 1
    assert np.allclose(out[0, 0, 1:-1].real, taper[1:-1])


def test_output_spectrum_is_scaled_rfft():
    class P:
        nfft_out = 64
        nspec = 33
        delta = 0.25
    x = np.random.default_rng(0).standard_normal((1, 1, 50))
    assert np.allclose(s2s.output_spectrum(x, P), np.fft.rfft(x, n=64) * 0.25)


def test_cpu_xc_restatement_matches_oracle():
    from oracle import xc

    if not cbind.have_cpu_xc():
        pytest.skip("oracle/libcpu_xc.so not built")
    rng = np.random.default_rng(4)
    nstep, nsta, nfft, H, dt = 3, 5, 1764, 150, 0.1
    x = rng.standard_normal((nstep, nsta, 1500))
    spec = (np.fft.rfft(x, n=nfft, axis=-1) * dt).astype(np.complex64)
    src, rec = np.triu_indices(nsta, k=1)
    got = cbind.cpu_xc_correlate(spec, src, rec, H, dt, nthreads=2)
    want = xc.correlate(spec, src, rec, nfft, H, dt)
    assert xc.rel_l2(got, want) < 1e-5


def _unit_plan(seg=600, nfft_filter=1200, nbands=2, delta=0.1):
    """A plan whose every power response is 1 (identity filters)."""
    return s2s.Plan(npts=seg, delta=delta, seg=seg, shift=seg, nstep=1, nfft_out=seg,
                    nspec=seg // 2 + 1, nfft_filter=nfft_filter, freq_low=0.2, freq_high=2.0,
                    idx=(0, 1, 2, 3), responses=[np.ones(nfft_filter)] * nbands,
                    freq_lows=[0.2] + [0.25] * (nbands - 1))


def test_prefilter_with_unit_response_is_identity():
    # zero padding to nfft_filter, x dt, |H|^2 = 1, inverse / (nfft dt), keep
    # the first seg samples: the identity on the window
    p = _unit_plan()
    x = np.random.default_rng(9).standard_normal((2, 1, p.seg))
    assert np.allclose(s2s.prefilter(x, p), x, atol=1e-12)


def test_prefilter_is_circular_convolution_with_response_kernel():
    # |H|^2[k] = 1 + cos(2 pi k d / n) is the spectrum of the kernel
    # delta[0] + (delta[d] + delta[-d]) / 2, so on the zero-padded window the
    # prefilter is y[t] = x[t] + (xp[t-d] + xp[t+d]) / 2 (indices mod n)
    p = _unit_plan()
    n, d = p.nfft_filter, 37
    k = np.arange(n)
    resp = 1.0 + np.cos(2 * np.pi * k * d / n)
    p.responses = [resp, resp]
    x = np.random.default_rng(11).standard_normal(p.seg)
    xp = np.zeros(n)
    xp[: p.seg] = x
    t = np.arange(p.seg)
    want = x + 0.5 * (xp[(t - d) % n] + xp[(t + d) % n])
    got = s2s.prefilter(x[None, None], p)[0, 0]
    assert np.allclose(got, want, atol=1e-12)


def test_runabs_mf_single_unit_band_equals_runabs():
    # one extra band with |H|^2 = 1: the band signal is the window itself, so
    # RUN-ABS-MF reduces to RUN-ABS at that band's low corner (no 0.667 factor)
    p = _unit_plan(nbands=2)
    x = np.random.default_rng(10).standard_normal((3, 2, p.seg))
    assert np.allclose(s2s.runabs_mf(x, p), s2s.runabs(x, 0.25, p.delta), atol=1e-12)
    # two identical bands add up to twice that
    p3 = _unit_plan(nbands=3)
    assert np.allclose(s2s.runabs_mf(x, p3), 2 * s2s.runabs(x, 0.25, p.delta), atol=1e-12)
