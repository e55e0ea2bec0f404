"""FP64 oracle of the stacks -- test infrastructure only.

  linear_stack      fastxc/operators/stacking/linear_stack.py:228-247 (float64 running
                    sum in record order / count, rounded once to float32)
  pws_prepare       prepare_sourcepack_host_item, native/pws/io/sourcepack_io.cu:385-478
  analytic          R2C + hilbertTransformKernel, native/pws/worker/gpu.cu:195-205,
                    native/pws/kernels/cuda.pws.cu:3-26 (and the tfpws copy,
                    native/tfpws/kernels/cuda.stransform.cu:10-39)
  pws               compute_pws_host_batch, native/pws/worker/gpu.cu:195-262,
                    kernels cuda.pws.cu:28-95 (power 1: lin * |sum w_g phi_g|)
  tfpws_prepare     prepare_tfpws_sourcepack_item, native/tfpws/io/sourcepack_io.cu:403-490
                    (group SUMS, w_g = n_g/T, lin = (1/T) sum)
  tfpws             compute_tfpws_from_prestack, native/tfpws/core/tfpws_compute.cu:19-366,
                    kernels cuda.stransform.cu:59-332 (scale 0.1, power 1,
                    G(c=0, f!=0) = 0, weight / N)

Pinning (tests/test_oracle_stack.py): identical traces give |W| = 1 so
pws = lin; identical groups give tfpws = 0 (the reference's zero Gaussian at
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


def pws(prestack: np.ndarray, w: np.ndarray, lin: np.ndarray) -> np.ndarray:
    N = prestack.shape[-1]
    a = np.fft.ifft(analytic_spectrum(prestack), axis=-1) * N / N  # unnormalised inverse, then /N
    mod = np.maximum(np.abs(a), 1e-7)
    phi = a / mod
    W = (np.asarray(w)[:, None] * phi).sum(axis=0)
    return np.asarray(lin, dtype=np.float64) * np.abs(W)


def gaussian(N: int, f: int, scale: float = 0.1) -> np.ndarray:
    """G(c, f) of gaussianModulateSub (cuda.stransform.cu:59-153)."""
    if f == 0:
        return np.ones(N)
    c = np.arange(N)
    half = (N + 1) // 2
    cm = np.where(c < half, c, N - c).astype(np.float64)
    g = np.exp(-2.0 * np.pi ** 2 * cm * cm * scale / (f * f))
    g[0] = 0.0
    return g


def s_transform_voice(A: np.ndarray, f: int, scale: float = 0.1) -> np.ndarray:
    """ST_f[tau] = sum_c A[(f+c) mod N] G(c, f) exp(+2 pi i c tau / N) (unnormalised)."""
    N = A.shape[-1]
    X = np.roll(A, -f, axis=-1) * gaussian(N, f, scale)
    return np.fft.ifft(X, axis=-1) * N


def tfpws(prestack: np.ndarray, w: np.ndarray, lin: np.ndarray, band=None,
          delta: float = 1.0, scale: float = 0.1) -> np.ndarray:
    """band: None or (fmin, fmax, taper_hz) for the band-limited blend."""
    prestack = np.asarray(prestack, dtype=np.float64)
    N = prestack.shape[-1]
    nfreq = N // 2 + 1
    Ag = analytic_spectrum(prestack)
    Al = analytic_spectrum(np.asarray(lin, dtype=np.float64)[None])[0]
    w = np.asarray(w, dtype=np.float64)
    out = np.zeros(N, dtype=np.complex128)
    if band is not None:
        out[:] = Al
    df = 1.0 / (N * delta)
    for f in range(nfreq):
        STg = s_transform_voice(Ag, f, scale)
        fa = np.abs(STg)
        fa = np.where(fa <= 1e-7, 1.0, fa)
        W = (w[:, None] * STg / (fa * N)).sum(axis=0)
        STl = s_transform_voice(Al[None], f, scale)[0] * np.abs(W)
        v = STl.sum()
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
