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


def _hilbert_pair(rng, N, theta):
    """x1 with no DC (and no Nyquist bin) and x2 whose positive-frequency bins
    are exp(-i theta) times x1's: x2 = cos(theta) x1 + sin(theta) H[x1], so the
    analytic signals satisfy a2 = exp(-i theta) a1 exactly."""
    half = (N + 1) // 2
    X1 = np.zeros(N, complex)
    X1[1:half] = rng.standard_normal(half - 1) + 1j * rng.standard_normal(half - 1)
    X1[N - half + 1:] = np.conj(X1[1:half][::-1])
    X2 = X1.copy()
    X2[1:half] *= np.exp(-1j * theta)
    X2[N - half + 1:] = np.conj(X2[1:half][::-1])
    return np.fft.ifft(X1).real, np.fft.ifft(X2).real


def test_pws_hilbert_rotation_closed_form():
    """Two equal-weight groups, the second the first Hilbert-rotated by theta:
    W = (phi + exp(-i theta) phi)/2, |W| = |cos(theta/2)|, so the reference's
    PWS (power 1, cuda.pws.cu:79-95) is lin |cos(theta/2)|; power 2 would give
    cos^2.  A wrong exponent, a missing weight or a dropped normalisation of
    phi fails this."""
    rng = np.random.default_rng(31)
    for N in (64, 81, 201):
        for theta in (2 * np.pi / 3, 1.1, np.pi / 2):
            x1, x2 = _hilbert_pair(rng, N, theta)
            a1 = np.fft.ifft(stack.analytic_spectrum(x1))
            assert np.allclose(np.fft.ifft(stack.analytic_spectrum(x2)), np.exp(-1j * theta) * a1,
                               atol=1e-12)
            assert np.min(np.abs(a1)) > 1e-6  # the 1e-7 floor never applies here
            pre = np.stack([x1, x2])
            w = np.array([0.5, 0.5])
            lin = 0.5 * (x1 + x2)
            c = abs(np.cos(theta / 2))
            assert np.allclose(stack.pws(pre, w, lin), lin * c, atol=1e-12)
            assert np.allclose(stack.pws(pre, w, lin, power=2.0), lin * c * c, atol=1e-12)


def test_stack_weight_homogeneity():
    """Scaling every group weight by s scales |W| by s, so PWS and the TF-PWS
    voice sums scale by s**power (power 1 in the reference)."""
    rng = np.random.default_rng(32)
    recs = rng.standard_normal((7, 45))
    g, w, lin = stack.pws_prepare(recs, 3)
    s = 0.37
    for p in (1.0, 2.0):
        assert np.allclose(stack.pws(g, s * w, lin, power=p), s ** p * stack.pws(g, w, lin, power=p),
                           atol=1e-12)
    sums, w2, lin2 = stack.tfpws_prepare(recs, 3)
    for p in (1.0, 2.0):
        a = stack.tfpws_spectrum(sums, s * w2, lin2, power=p)
        b = stack.tfpws_spectrum(sums, w2, lin2, power=p)
        assert np.allclose(a, s ** p * b, atol=1e-12 * np.abs(b).max())
        assert np.abs(b).max() > 0


def test_tfpws_spec_voices():
    """Spec-mode voices: every voice is the reference; at a voice bin the
    voice sum equals the reference's; between voices it is the sum with the
    weights blended linearly (checked at the midpoint of two voices against
    the two voices' weights computed directly)."""
    rng = np.random.default_rng(33)
    N = 61
    recs = rng.standard_normal((6, N))
    sums, w, lin = stack.tfpws_prepare(recs, 2)
    nfreq = N // 2 + 1
    ref = stack.tfpws_spectrum(sums, w, lin)
    assert np.allclose(stack.tfpws_spectrum(sums, w, lin, nvoices=nfreq), ref, atol=0)
    vb = stack.voice_bins(nfreq, 9)
    assert vb[0] == 0 and vb[-1] == nfreq - 1 and len(set(vb.tolist())) == 9
    sp = stack.tfpws_spectrum(sums, w, lin, nvoices=9)
    assert np.allclose(sp[vb], ref[vb], atol=1e-13)
    j = 3
    f = (vb[j] + vb[j + 1]) // 2
    t = (f - vb[j]) / (vb[j + 1] - vb[j])
    Ag = stack.analytic_spectrum(sums)
    Al = stack.analytic_spectrum(lin[None])[0]
    c = (1 - t) * stack.tfpws_weight(Ag, w, int(vb[j])) + t * stack.tfpws_weight(Ag, w, int(vb[j + 1]))
    want = (stack.s_transform_voice(Al[None], f)[0] * c).sum()
    assert abs(sp[f] - want) < 1e-12 * abs(want) + 1e-15
    # the full TF-PWS with every voice equals the parity form
    assert np.allclose(stack.tfpws(sums, w, lin, nvoices=0), stack.tfpws(sums, w, lin), atol=0)


def test_s_transform_even_length_bruteforce():
    """Even N: the Gaussian of gaussianModulateSub (cuda.stransform.cu:96-141)
    written as its branches -- c < (N+1)/2 direct, else the mirror N-c only
    when it is below (N+1)/2 and nonzero -- so the shift c = N/2 gets 0."""
    rng = np.random.default_rng(34)
    N = 16
    A = stack.analytic_spectrum(rng.standard_normal(N))
    half = (N + 1) // 2
    for f in (0, 1, 3, 8):
        got = stack.s_transform_voice(A[None], f)[0]
        want = np.zeros(N, complex)
        for tau in range(N):
            for c in range(N):
                if f == 0:
                    gc = 1.0
                elif c == 0:
                    gc = 0.0
                elif c < half:
                    gc = np.exp(-2 * np.pi ** 2 * c * c * 0.1 / f ** 2)
                else:
                    m = N - c
                    gc = np.exp(-2 * np.pi ** 2 * m * m * 0.1 / f ** 2) if (m < half and m != 0) else 0.0
                want[tau] += A[(f + c) % N] * gc * np.exp(2j * np.pi * c * tau / N)
        assert np.allclose(got, want, atol=1e-10)
    assert stack.gaussian(N, 3)[N // 2] == 0.0 and stack.gaussian(N + 1, 3)[N // 2] > 0.0
