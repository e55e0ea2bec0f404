"""FP64 oracle of the stacks -- test infrastructure only.

  linear_stack      fastxc/operators/stacking/linear_stack.py:228-247 (float64 running
                    sum in record order / count, rounded once to float32)
  pws_prepare       prepare_sourcepack_host_item, native/pws/io/sourcepack_io.cu:385-478
  analytic          R2C + hilbertTransformKernel, native/pws/worker/gpu.cu:195-205,
                    native/pws/kernels/cuda.pws.cu:3-26 (and the tfpws copy,
                    native/tfpws/kernels/cuda.stransform.cu:10-39)
  pws               compute_pws_host_batch, native/pws/worker/gpu.cu:195-262,
                    kernels cuda.pws.cu:28-95 (power 1: lin * |sum w_g phi_g|;
                    other powers: spec mode)
  tfpws_prepare     prepare_tfpws_sourcepack_item, native/tfpws/io/sourcepack_io.cu:403-490
                    (group SUMS, w_g = n_g/T, lin = (1/T) sum)
  tfpws             compute_tfpws_from_prestack, native/tfpws/core/tfpws_compute.cu:19-366,
                    kernels cuda.stransform.cu:59-332 (scale 0.1, power 1,
                    G(c=0, f!=0) = 0, G(n/2, f!=0) = 0 for even n, weight / N,
                    every voice); spec mode: other powers, fewer voices with
                    the weights interpolated linearly between them

Pinning (tests/test_oracle_stack.py): identical traces give |W| = 1 so
pws = lin; a group and its Hilbert rotation by theta (analytic signals
a and exp(-i theta) a) give |W| = |cos(theta/2)|, so pws = lin |cos(theta/2)|
only at power 1; scaling every weight by s scales pws and the voice sums of
tfpws by s**power; spec-mode voice sums equal the reference's at the voice
bins and tfpws with every voice equals the reference; identical groups give tfpws = 0 (the reference's zero Gaussian at
c = 0 removes every voice's time sum); a cosine's analytic signal is
exp(i w t); the FFT S-transform equals a brute-force DFT evaluation on small
N; the band blend with a band covering every voice equals plain tfpws and
with a band above Nyquist returns the analytic linear stack's real part;
group means/sums and weights against hand-computed values; linear_stack
equals the correctly rounded mean (math.fsum) on float32 inputs and is
invariant to a constant offset that exactly cancels.
"""
from __future__ import annotations

import numpy as np


def linear_stack(records: np.ndarray) -> np.ndarray:
    """float64 running sum in record (timestamp) order, float64 division by
    the count, one rounding to float32 (_linear_stack_group,
    fastxc/operators/stacking/linear_stack.py:228-247)."""
    r = np.asarray(records, dtype=np.float32)
    acc = r[0].astype(np.float64)
    for row in r[1:]:
        acc += row.astype(np.float64)
    return (acc / float(r.shape[0])).astype(np.float32)


def _groups(T: int, B: int):
    g = 1 if B < 2 else B
    return [(i, min(i + g, T)) for i in range(0, T, g)]


def pws_prepare(records: np.ndarray, B: int):
    """records [T, N] (timestamp order) -> (group means [G, N], w [G], lin [N]), FP64."""
    r = np.asarray(records, dtype=np.float64)
    T = r.shape[0]
    means, w = [], []
    for a, b in _groups(T, B):
        means.append(r[a:b].mean(axis=0))
        w.append((b - a) / T)
    means = np.stack(means)
    w = np.array(w)
    return means, w, (w[:, None] * means).sum(axis=0)


def tfpws_prepare(records: np.ndarray, B: int):
    """records [T, N] -> (group sums [G, N], w [G], lin [N]), FP64."""
    r = np.asarray(records, dtype=np.float64)
    T = r.shape[0]
    sums, w = [], []
    for a, b in _groups(T, B):
        sums.append(r[a:b].sum(axis=0))
        w.append((b - a) / T)
    return np.stack(sums), np.array(w), r.sum(axis=0) / T


def analytic_spectrum(x: np.ndarray) -> np.ndarray:
    """Full-length spectrum with bin 0 and bins >= (N+1)/2 zeroed, 1..(N+1)/2-1 doubled."""
    N = x.shape[-1]
    X = np.fft.fft(np.asarray(x, dtype=np.float64), axis=-1)
    half = (N + 1) // 2
    A = np.zeros_like(X)
    A[..., 1:half] = 2.0 * X[..., 1:half]
    return A


def pws(prestack: np.ndarray, w: np.ndarray, lin: np.ndarray, power: float = 1.0) -> np.ndarray:
    """lin * |sum_g w_g phi_g|**power; the reference is power 1 (cuda.pws.cu:79-95),
    other powers are the spec mode."""
    N = prestack.shape[-1]
    a = np.fft.ifft(analytic_spectrum(prestack), axis=-1) * N / N  # unnormalised inverse, then /N
    mod = np.maximum(np.abs(a), 1e-7)
    phi = a / mod
    W = (np.asarray(w)[:, None] * phi).sum(axis=0)
    return np.asarray(lin, dtype=np.float64) * np.abs(W) ** power


def voice_bins(nfreq: int, nvoices: int) -> np.ndarray:
    """Spec-mode voices: nvoices bins spread evenly over [0, nfreq-1],
    v_j = floor(j (nfreq-1) / (nvoices-1)); nvoices <= 0 or >= nfreq means
    every bin (the reference, tfpws_compute.cu:224)."""
    if nvoices <= 0 or nvoices >= nfreq:
        return np.arange(nfreq)
    if nvoices == 1:
        return np.array([0])
    j = np.arange(nvoices, dtype=np.int64)
    return (j * (nfreq - 1)) // (nvoices - 1)


def gaussian(N: int, f: int, scale: float = 0.1) -> np.ndarray:
    """G(c, f) of gaussianModulateSub (cuda.stransform.cu:59-153)."""
    if f == 0:
        return np.ones(N)
    c = np.arange(N)
    half = (N + 1) // 2
    cm = np.where(c < half, c, N - c).astype(np.float64)
    g = np.exp(-2.0 * np.pi ** 2 * cm * cm * scale / (f * f))
    g[0] = 0.0
    # the mirror branch requires N - c < (N+1)/2: for even N the shift c = N/2
    # meets neither branch and gets 0 (cuda.stransform.cu:119-141)
    g[cm >= half] = 0.0
    return g


def s_transform_voice(A: np.ndarray, f: int, scale: float = 0.1) -> np.ndarray:
    """ST_f[tau] = sum_c A[(f+c) mod N] G(c, f) exp(+2 pi i c tau / N) (unnormalised)."""
    N = A.shape[-1]
    X = np.roll(A, -f, axis=-1) * gaussian(N, f, scale)
    return np.fft.ifft(X, axis=-1) * N


def tfpws_weight(Ag: np.ndarray, w: np.ndarray, f: int, scale: float = 0.1,
                 power: float = 1.0) -> np.ndarray:
    """|W_f(tau)|**power, W_f = sum_g w_g ST_g / (|ST_g|' N) (cuda.stransform.cu:176-198,
    |ST| <= 1e-7 replaced by 1)."""
    N = Ag.shape[-1]
    STg = s_transform_voice(Ag, f, scale)
    fa = np.abs(STg)
    fa = np.where(fa <= 1e-7, 1.0, fa)
    W = (np.asarray(w, dtype=np.float64)[:, None] * STg / (fa * N)).sum(axis=0)
    return np.abs(W) ** power


def tfpws_spectrum(prestack: np.ndarray, w: np.ndarray, lin: np.ndarray, scale: float = 0.1,
                   power: float = 1.0, nvoices: int = 0) -> np.ndarray:
    """Voice sums v[f] = sum_tau ST_lin,f(tau) c_f(tau), f < N/2+1.  c_f is
    |W_f|**power at the voices (every bin in the reference); between two
    voices v_j < f < v_j+1 (spec mode) it is interpolated linearly in f,
    c_f = (1-t) c_vj + t c_vj+1 with t = (f - v_j) / (v_j+1 - v_j)."""
    prestack = np.asarray(prestack, dtype=np.float64)
    N = prestack.shape[-1]
    nfreq = N // 2 + 1
    Ag = analytic_spectrum(prestack)
    Al = analytic_spectrum(np.asarray(lin, dtype=np.float64)[None])[0]
    vb = voice_bins(nfreq, nvoices)
    cw = {int(v): tfpws_weight(Ag, w, int(v), scale, power) for v in vb}
    out = np.zeros(nfreq, dtype=np.complex128)
    for f in range(nfreq):
        j = int(np.searchsorted(vb, f, side="right")) - 1
        if vb[j] == f or j + 1 >= len(vb):
            c = cw[int(vb[j])]
        else:
            t = (f - vb[j]) / float(vb[j + 1] - vb[j])
            c = (1.0 - t) * cw[int(vb[j])] + t * cw[int(vb[j + 1])]
        out[f] = (s_transform_voice(Al[None], f, scale)[0] * c).sum()
    return out


def tfpws(prestack: np.ndarray, w: np.ndarray, lin: np.ndarray, band=None,
          delta: float = 1.0, scale: float = 0.1, power: float = 1.0,
          nvoices: int = 0) -> np.ndarray:
    """band: None or (fmin, fmax, taper_hz) for the band-limited blend.
    power 1 and every voice (nvoices 0) is the reference; the rest is the
    spec mode (tfpws_spectrum)."""
    prestack = np.asarray(prestack, dtype=np.float64)
    N = prestack.shape[-1]
    nfreq = N // 2 + 1
    Al = analytic_spectrum(np.asarray(lin, dtype=np.float64)[None])[0]
    vs = tfpws_spectrum(prestack, w, lin, scale, power, nvoices)
    out = np.zeros(N, dtype=np.complex128)
    if band is not None:
        out[:] = Al
    df = 1.0 / (N * delta)
    for f in range(nfreq):
        v = vs[f]
        if band is None:
            out[f] = v
        else:
            fmin, fmax, taper = band
            freq = f * df
            alpha = 0.0
            if fmin <= freq <= fmax:
                t = min(taper, 0.5 * (fmax - fmin))
                if t <= 0:
                    alpha = 1.0
                elif freq < fmin + t:
                    alpha = 0.5 - 0.5 * np.cos(np.pi * (freq - fmin) / t)
                elif freq > fmax - t:
                    alpha = 0.5 - 0.5 * np.cos(np.pi * (fmax - freq) / t)
                else:
                    alpha = 1.0
            out[f] = Al[f] * (1 - alpha) + v * alpha
    return (np.fft.ifft(out) * N).real / N


def tfpws_chunk_overlaps_band(f_start: int, count: int, df: float, fmin: float, fmax: float) -> bool:
    """A voice chunk is processed when any of its frequencies lies in [fmin, fmax]."""
    lo = f_start * df
    hi = (f_start + count - 1) * df
    return not (hi < fmin or lo > fmax)
