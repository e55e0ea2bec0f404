"""Pins the xc oracle against brute force and invariants (CPU only)."""
import itertools

import numpy as np

from oracle import xc


def _spectra_from_windows(x, nfft, dt):
    # sac2spec output step: zero-pad to nfft, R2C, x dt (native/sac2spec/worker/compute.cu:226-252)
    return np.fft.rfft(x, n=nfft, axis=-1) * dt


def test_lags_match_time_domain_bruteforce():
    rng = np.random.default_rng(1)
    nstep, seg, H, dt = 3, 40, 7, 0.25
    nfft = 50  # >= seg + H so lags |tau| <= H do not wrap
    x = rng.standard_normal((nstep, 2, seg))
    spec = _spectra_from_windows(x, nfft, dt)
    got = xc.correlate(spec, [0], [1], nfft, H, dt)[0]
    want = np.zeros(2 * H + 1)
    for w in range(nstep):
        for tau in range(-H, H + 1):
            s = 0.0
            for t in range(seg):
                if 0 <= t + tau < seg:
                    s += x[w, 0, t] * x[w, 1, t + tau]
            want[tau + H] += dt * s
    want /= nstep
    # oracle applies the float32 weight 1/((float)nfft*dt): ~1e-7 relative
    assert np.allclose(got, want, rtol=1e-6, atol=1e-9)


def test_delayed_copy_peaks_at_positive_lag():
    rng = np.random.default_rng(2)
    seg, d, H, dt = 200, 13, 30, 0.1
    base = rng.standard_normal(seg + d)
    src = base[d:d + seg]
    rec = base[:seg]  # rec(t) = src(t - d): receiver sees the wave later
    x = np.stack([src, rec])[None]
    nfft = 256
    cc = xc.correlate(_spectra_from_windows(x, nfft, dt), [0], [1], nfft, H, dt)[0]
    # c(tau) = sum_t src(t) rec(t + tau) peaks where rec(t + tau) = src(t): tau = +d
    assert int(np.argmax(cc)) - H == d
    want = np.array([sum(src[t] * rec[t + tau] for t in range(seg) if 0 <= t + tau < seg)
                     for tau in range(-H, H + 1)]) * dt
    assert np.allclose(cc, want, rtol=1e-6, atol=1e-9)


def test_finalize_shape_closed_forms():
    s = xc.finalize_shape(18523, 24, 0.1, 100.0)
    assert (s.nfft, s.half_cc, s.cc_size) == (37044, 1000, 2001)
    s = xc.finalize_shape(37045, 24, 0.05, 100.0)
    assert (s.nfft, s.half_cc, s.cc_size) == (74088, 2000, 4001)
    # clamping H to nspec-1 gives cc_size = nfft+1, which the reference rejects
    import pytest
    with pytest.raises(ValueError):
        xc.finalize_shape(11, 1, 1.0, 1000.0)


def _allowed_all(ids):
    u = sorted(set(int(i) for i in ids))
    return {xc.make_allowed_path_id(a, b) for a, b in itertools.combinations_with_replacement(u, 2)}


def test_task_enumeration_covers_each_pair_once():
    # 7 stations x 3 channels, block size 4 rows -> several diagonal/off-diagonal jobs
    ids = np.repeat(np.arange(1, 8), 3)
    allowed = _allowed_all(ids)
    allowed.discard(xc.make_allowed_path_id(2, 5))  # one forbidden path
    jobs = xc.build_jobs(len(ids), 4)
    seen = []
    for job in jobs:
        tasks, loaded = xc.build_tasks_for_job(ids, allowed, job)
        for t in tasks:
            assert loaded[t.src_local] == t.src_meta and loaded[t.rec_local] == t.rec_meta
            assert ids[t.src_meta] <= ids[t.rec_meta]
            seen.append((t.src_meta, t.rec_meta))
    # brute force of the same rule: every ordered NSLC row pair (a,b) with
    # gnsl(a) < gnsl(b), plus every (a,b) of one station, if the path is allowed
    want = set()
    for a in range(len(ids)):
        for b in range(len(ids)):
            ga, gb = ids[a], ids[b]
            if xc.make_allowed_path_id(min(ga, gb), max(ga, gb)) not in allowed:
                continue
            if ga < gb or ga == gb:
                want.add((a, b))
    assert len(seen) == len(set(seen))
    assert set(seen) == want


def test_task_order_within_diagonal_block():
    ids = np.arange(1, 5)
    allowed = _allowed_all(ids)
    job = xc.build_jobs(4, 4)[0]
    tasks, _ = xc.build_tasks_for_job(ids, allowed, job)
    got = [(t.src_meta, t.rec_meta) for t in tasks]
    assert got == [(0, 0), (0, 1), (0, 2), (0, 3), (1, 1), (1, 2), (1, 3), (2, 2), (2, 3), (3, 3)]


def test_valid_stack_all_valid_is_the_reference_stack():
    rng = np.random.default_rng(41)
    spec = rng.standard_normal((7, 5, 33)) + 1j * rng.standard_normal((7, 5, 33))
    src, rec = np.triu_indices(5, k=1)
    got = xc.cross_spectrum_valid(spec, np.ones((7, 5), np.uint8), src, rec)
    np.testing.assert_allclose(got, xc.cross_spectrum(spec, src, rec), rtol=1e-13, atol=0)


def test_valid_stack_divides_by_shared_windows():
    # station 0 valid in windows {0, 1}, station 1 in {1, 2}, station 2 in {3}:
    # pair (0, 1) shares window 1 only, pairs (0, 2) and (1, 2) share none
    rng = np.random.default_rng(42)
    spec = rng.standard_normal((4, 3, 9)) + 1j * rng.standard_normal((4, 3, 9))
    valid = np.zeros((4, 3), np.uint8)
    valid[[0, 1], 0] = 1
    valid[[1, 2], 1] = 1
    valid[3, 2] = 1
    got = xc.cross_spectrum_valid(spec, valid, [0, 0, 1], [1, 2, 2])
    np.testing.assert_allclose(got[0], np.conj(spec[1, 0]) * spec[1, 1], rtol=1e-14)
    assert not got[1:].any()
    # two shared windows: their mean, not the mean over all four windows
    valid[2, 0] = 1
    got = xc.cross_spectrum_valid(spec, valid, [0], [1])
    want = (np.conj(spec[1, 0]) * spec[1, 1] + np.conj(spec[2, 0]) * spec[2, 1]) / 2
    np.testing.assert_allclose(got[0], want, rtol=1e-14)
