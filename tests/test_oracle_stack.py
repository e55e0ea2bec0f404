"""Pins the stack oracle (CPU only) with closed forms and brute force."""
import numpy as np

from oracle import stack


def test_linear_stack_is_float64_mean():
    """Against the correctly rounded mean: float32 inputs summed in float64
    lose nothing for short stacks, so one rounding of fsum/T must agree."""
    import math
    r = np.random.default_rng(0).standard_normal((7, 33)).astype(np.float32)
    want = np.array([math.fsum(c) / 7.0 for c in r.T.astype(np.float64)]).astype(np.float32)
    assert np.array_equal(stack.linear_stack(r), want)
    # a large offset that cancels exactly (the sum stays exact in float64)
    big = np.vstack([r, np.full((1, 33), 2.0 ** 20, np.float32), np.full((1, 33), -2.0 ** 20,
                                                                          np.float32)])
    want9 = np.array([math.fsum(c) / 9.0 for c in big.T.astype(np.float64)]).astype(np.float32)
    assert np.array_equal(stack.linear_stack(big), want9)
    # row order matters only through rounding: a single record is returned as is
    assert np.array_equal(stack.linear_stack(r[:1]), r[0])


def test_analytic_signal_of_cosine():
    N, k = 101, 7
    t = np.arange(N)
    x = np.cos(2 * np.pi * k * t / N)
    a = np.fft.ifft(stack.analytic_spectrum(x))
    assert np.allclose(a, np.exp(2j * np.pi * k * t / N), atol=1e-12)


def test_pws_identical_traces_equal_linear():
    rng = np.random.default_rng(1)
    tr = rng.standard_normal(61)
    recs = np.tile(tr, (9, 1))
    g, w, lin = stack.pws_prepare(recs, 4)
    assert np.allclose(w, [4 / 9, 4 / 9, 1 / 9])
    assert np.allclose(stack.pws(g, w, lin), lin)


def test_pws_group_means_and_weights():
    rng = np.random.default_rng(2)
    recs = rng.standard_normal((5, 20))
    g, w, lin = stack.pws_prepare(recs, 2)
    assert np.allclose(g[0], recs[:2].mean(0)) and np.allclose(g[2], recs[4])
    assert np.allclose(lin, recs.mean(0))
    s, w2, lin2 = stack.tfpws_prepare(recs, 2)
    assert np.allclose(s[1], recs[2:4].sum(0)) and np.allclose(lin2, recs.mean(0))
    assert np.allclose(w, w2)


def test_s_transform_matches_bruteforce():
    rng = np.random.default_rng(3)
    N = 15
    A = stack.analytic_spectrum(rng.standard_normal(N))
    for f in (0, 1, 4, 7):
        got = stack.s_transform_voice(A[None], f)[0]
        want = np.zeros(N, complex)
        for tau in range(N):
            for c in range(N):
                cm = c if c < (N + 1) // 2 else N - c
                if f == 0:
                    gc = 1.0
                elif c == 0:
                    gc = 0.0
                else:
                    gc = np.exp(-2 * np.pi ** 2 * cm * cm * 0.1 / f ** 2)
                want[tau] += A[(f + c) % N] * gc * np.exp(2j * np.pi * c * tau / N)
        assert np.allclose(got, want, atol=1e-10)


def test_tfpws_identical_groups_vanish():
    # identical groups give |W| = 1/N at every tau; with G(0, f != 0) = 0 each
    # voice's time sum is N * X_f[0] = 0, and the DC voice carries A[0] = 0.
    rng = np.random.default_rng(4)
    tr = rng.standard_normal(41)
    s, w, lin = stack.tfpws_prepare(np.tile(tr, (6, 1)), 2)
    out = stack.tfpws(s / 2, w, lin)  # scaled sums equal the linear stack
    assert np.max(np.abs(out)) < 1e-9 * np.max(np.abs(lin))
    # the band-limited blend keeps the analytic linear stack outside the band
    outb = stack.tfpws(s / 2, w, lin, band=(0.2, 0.3, 0.01), delta=1.0)
    assert np.max(np.abs(outb)) > 0.1 * np.max(np.abs(lin))


def test_tfpws_band_blend_limits():
    # a band that covers every voice with no taper blends with alpha = 1
    # (plain TF-PWS); a band above Nyquist keeps the analytic linear stack,
    # whose inverse is the linear stack's analytic signal real part: lin
    # minus its mean (the DC bin is dropped by the analytic mask)
    rng = np.random.default_rng(12)
    recs = rng.standard_normal((8, 45))
    s, w, lin = stack.tfpws_prepare(recs, 2)
    plain = stack.tfpws(s, w, lin)
    full = stack.tfpws(s, w, lin, band=(0.0, 0.5, 0.0), delta=1.0)
    assert np.allclose(full, plain, atol=1e-12)
    none = stack.tfpws(s, w, lin, band=(10.0, 20.0, 0.1), delta=1.0)
    N = lin.shape[-1]
    A = stack.analytic_spectrum(lin[None])[0]
    want = (np.fft.ifft(A) * N).real / N
    assert np.allclose(none, want, atol=1e-12)
    assert np.allclose(want, lin - lin.mean(), atol=1e-9) or N % 2 == 0
