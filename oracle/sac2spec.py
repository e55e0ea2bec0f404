"""FP64 oracle of sac2spec (stage 1) -- test infrastructure only.

Follows the reference step order for one batch of windows; arrays are
[frames, nch, samples] (a frame is one station-window; its channels share the
normalisation and whitening weights).

  plan_lengths           native/sac2spec/cuda.main.cu:190-296, fft_length.c:20-109,
                         core/filter_response.c:10-105 (padding), :336-420 (|H|^2)
  read_filters           core/filter_response.c:134-205 (filter.txt)
  preprocess             core/preprocess.cu:23-105, kernels/rdcrtr.cu:3-119,
                         kernels/taper.cu:71-99, kernels/real_matrix.cu:39-52
  prefilter              worker/compute.cu:59-86
  smooth_rows            kernels/real_matrix.cu:155-191 (always divides by W)
  runabs                 core/normalization.cu:91-213
  runabs_mf              core/normalization.cu:232-393
  onebit                 core/normalization.cu:77-89, kernels/real_matrix.cu:127-140
  freq_whiten            core/spectrum.cu:178-305, kernels/complex_matrix.cu:5-90,
                         kernels/taper.cu:8-69
  output_spectrum        worker/compute.cu:226-264, kernels/complex_matrix.cu:47-72
  process_batch          worker/compute.cu:8-272

Integer decisions the reference takes in float32 (taper length, window sizes,
whitening indices) are reproduced in float32; everything else is FP64.

Pinning (tests/test_oracle_sac2spec.py):
  plan_lengths       against the reference's own C (fft_length.c,
                     filter_response.c) compiled into oracle/_ref when it is
                     built, and against the closed-form config-0/2 numbers;
  |H|^2              against scipy.signal.freqz of the same coefficients;
  preprocess         a ramp plus offset detrends to zero; taper closed form;
  prefilter          unit response is the identity on the window; a
                     1 + cos(2 pi k d / n) response equals the brute-force
                     circular convolution with delta[0] + (delta[+-d])/2;
  smooth_rows        brute force on small rows (edges, even/odd windows);
  runabs / onebit    constant input normalises to 1 (x/(W*|x|/W)) and sign;
  runabs_mf          one unit-response band reduces to runabs at that band's
                     corner; two identical bands give twice that;
  freq_whiten        flat |X| whitens to the taper; brute-force smooth;
  output_spectrum    numpy.fft.rfft * dt.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

MAXVAL = 20.0
MINVAL = 1e-30
F32_MINVAL = float(np.float32(1e-30))


def _f32(x) -> np.float32:
    return np.float32(x)


def _llround(x: float) -> int:
    return int(math.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)


def is_smooth_2357(v: int) -> bool:
    if v < 1:
        return False
    for p in (2, 3, 5, 7):
        while v > 1 and v % p == 0:
            v //= p
    return v == 1


def next_smooth_2357(n: int) -> int:
    n = max(n, 1)
    while not is_smooth_2357(n):
        n += 1
    return n


def next_even_smooth_2357(n: int) -> int:
    v = next_smooth_2357(n)
    while v % 2:
        v = next_smooth_2357(v + 1)
    return v


def previous_smooth_2357(n: int) -> int:
    v = n - 1
    while v > 0:
        if is_smooth_2357(v):
            return v
        v -= 1
    return 0


def segment_length_from_seconds(seconds: float, max_samples: int, delta: float) -> int:
    if seconds <= 0 or delta <= 0:
        req = 0
    else:
        req = _llround(float(np.float64(_f32(seconds)) / np.float64(_f32(delta))))
    seg = next_smooth_2357(req)
    if seg > max_samples:
        seg = previous_smooth_2357(max_samples)
    return seg


@dataclass
class Butterworth:
    freq_low: float
    freq_high: float
    b: np.ndarray
    a: np.ndarray


def read_filters(path: str) -> list[Butterworth]:
    """'# lo/hi' line then b (5) and a (5) coefficient lines per band."""
    out = []
    state = 0
    lo = hi = 0.0
    b = None
    for line in open(path):
        if line.startswith("#"):
            try:
                body = line[1:].strip()
                lo_s, hi_s = body.split("/")[:2]
                lo, hi = float(_f32(float(lo_s))), float(_f32(float(hi_s.split()[0])))
            except Exception:
                continue
            state = 1
            continue
        vals = line.split()
        if state == 1:
            if len(vals) < 5:
                continue
            b = np.array([float(v) for v in vals[:5]])
            state = 2
        elif state == 2:
            if len(vals) < 5:
                state = 0
                continue
            a = np.array([float(v) for v in vals[:5]])
            out.append(Butterworth(lo, hi, b, a))
            state = 0
    return out


def write_filter_file(path: str, delta: float, bands, order: int = 2) -> None:
    """filter.txt as design_filter writes it (fastxc/operators/filter_design.py:14-67):
    the broad band (min low / max high) first, then each band; per band a
    '# lo/hi' line and tab-separated %.18e b and a rows of butter(order, .., 'bandpass')."""
    from scipy.signal import butter

    if isinstance(bands, str):
        bands = bands.split()
    pairs = [tuple(map(float, b.split("/"))) for b in bands]
    fny = 0.5 / delta
    lo_all = min(p[0] for p in pairs)
    hi_all = max(p[1] for p in pairs)
    rows = [(f"{lo_all}/{hi_all}",) + tuple(butter(order, [lo_all / fny, hi_all / fny], "bandpass"))]
    for lo, hi in pairs:
        rows.append((f"{lo}/{hi}",) + tuple(butter(order, [lo / fny, hi / fny], "bandpass")))
    with open(path, "w", encoding="ascii") as fh:
        for tag, b, a in rows:
            fh.write(f"# {tag}\n")
            fh.write("\t".join(f"{v:.18e}" for v in b) + "\n")
            fh.write("\t".join(f"{v:.18e}" for v in a) + "\n")


def impulse_length(f: Butterworth, max_samples: int) -> int:
    """Length holding all but 1e-6 of the impulse-response energy
    (estimateSinglePassImpulseLength, core/filter_response.c:10-75)."""
    a0 = f.a[0]
    if abs(a0) < 1e-30:
        return max_samples
    x = np.zeros(5)
    y = np.zeros(5)
    resp = np.zeros(max_samples)
    for n in range(max_samples):
        x[1:] = x[:-1].copy()
        y[1:] = y[:-1].copy()
        x[0] = 1.0 if n == 0 else 0.0
        v = float(np.dot(f.b, x)) - float(np.dot(f.a[1:], y[1:]))
        v /= a0
        y[0] = v
        resp[n] = v
    total = float(np.sum(resp * resp))
    if total <= 0:
        return 1
    tail = 0.0
    for n in range(max_samples - 1, -1, -1):
        tail += resp[n] * resp[n]
        if tail > total * 1e-6:
            return n + 1
    return 1


def filter_padding(filters, max_padding: int) -> int:
    m = 1
    for f in filters:
        m = max(m, impulse_length(f, max_padding))
    return min(max(2 * m, 1), max_padding)


def power_response(f: Butterworth, nfft: int) -> np.ndarray:
    """|B(e^-i pi k/half)|^2/|A|^2 for k = 0..half, stored as float32 in a
    length-nfft array (calFilterPowerResp, core/filter_response.c:336-388)."""
    half = nfft // 2
    resp = np.zeros(nfft, dtype=np.float32)
    k = np.arange(half + 1)
    w = np.pi * k / half
    j = np.arange(5)
    ph = np.exp(-1j * np.outer(w, j))
    num = np.abs(ph @ f.b) ** 2
    den = np.abs(ph @ f.a) ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        p = np.where(den <= np.finfo(float).tiny, 0.0, num / den)
    p = np.where(np.isfinite(p), p, 0.0)
    p = np.minimum(p, np.finfo(np.float32).max)
    resp[: half + 1] = p.astype(np.float32)
    return resp


@dataclass
class Plan:
    npts: int
    delta: float
    seg: int
    shift: int
    nstep: int
    nfft_out: int
    nspec: int
    nfft_filter: int
    freq_low: float
    freq_high: float
    idx: tuple
    responses: list = field(default_factory=list)
    freq_lows: list = field(default_factory=list)


def plan_lengths(sac_len: float, delta: float, win_len: float, shift_len: float,
                 max_lag: float, filters: list[Butterworth]) -> Plan:
    delta32 = _f32(delta)
    npts = _llround(float(np.float64(_f32(sac_len)) / np.float64(delta32)))
    seg = segment_length_from_seconds(win_len, npts, float(delta32))
    shift = _llround(float(np.float64(_f32(shift_len)) / np.float64(delta32)))
    if shift < 1:
        shift = seg
    nstep = max((npts - seg) // shift + 1, 1) if seg <= npts else 1
    keep = _llround(float(np.float64(_f32(max_lag)) / np.float64(delta32)))
    keep = min(keep, seg)
    n_out = min(next_even_smooth_2357(seg + keep), 2 * seg)
    seg_df = _f32(_f32(1.0) / _f32(_f32(seg) * delta32))
    flo, fhi = _f32(filters[0].freq_low), _f32(filters[0].freq_high)
    idx = (int(_f32(_f32(flo * _f32(0.667)) / seg_df)), int(_f32(flo / seg_df)),
           int(_f32(fhi / seg_df)), int(_f32(_f32(fhi * _f32(1.333)) / seg_df)))
    pad = filter_padding(filters, seg)
    n_f = min(next_even_smooth_2357(seg + pad), 2 * seg)
    resp = [power_response(f, n_f) for f in filters]
    return Plan(npts, float(delta32), seg, shift, nstep, n_out, n_out // 2 + 1, n_f, float(flo),
                float(fhi), idx, resp, [float(_f32(f.freq_low)) for f in filters])


def _sanitize_c(X):
    bad = ~(np.isfinite(X.real) & np.isfinite(X.imag))
    return np.where(bad, 0, X)


def _inv_weight(n: int, delta: float) -> float:
    # 1.0 / (width * dt) with width converted to float
    return 1.0 / float(_f32(_f32(n) * _f32(delta)))


def preprocess(x: np.ndarray, freq_low: float, delta: float) -> np.ndarray:
    x = np.where(np.isfinite(x), x, 0.0).astype(np.float64)
    W = x.shape[-1]
    s = x.sum(axis=-1, keepdims=True)
    x = x - s / W
    s = x.sum(axis=-1, keepdims=True)
    i = np.arange(W, dtype=np.float64)
    isum = (x * i).sum(axis=-1, keepdims=True)
    a12 = 0.5 * W * (W - 1)
    a11 = a12 * (2 * W - 1) / 3.0
    a22 = float(W)
    b = a11 * a22 - a12 * a12
    aa = (a22 * isum - a12 * s) / b
    bb = (a11 * s - a12 * isum) / b
    x = x - (aa * i + bb)
    T = int(_f32(_f32(_f32(2.0) * _f32(_f32(1.0) / _f32(freq_low))) / _f32(delta)))
    return x * time_taper(W, T)


def time_taper(W: int, T: int) -> np.ndarray:
    col = np.arange(W, dtype=np.float64)
    f = np.ones(W)
    if T > 0:
        lo = col < T
        f[lo] = 0.5 * (1.0 - np.cos(np.pi * col[lo] / T))
        hi = (~lo) & (col >= W - T)
        f[hi] = 0.5 * (1.0 - np.cos(np.pi * (W - col[hi]) / T))
    return f


def prefilter(x: np.ndarray, plan: Plan, band: int = 0) -> np.ndarray:
    X = np.fft.rfft(x, n=plan.nfft_filter, axis=-1) * plan.delta
    X = _sanitize_c(X) * plan.responses[band][: plan.nfft_filter // 2 + 1]
    y = np.fft.irfft(X, n=plan.nfft_filter, axis=-1) * plan.nfft_filter
    return y[..., : plan.seg] * _inv_weight(plan.nfft_filter, plan.delta)


def smooth_rows(x: np.ndarray, win: int) -> np.ndarray:
    """Centred running mean, start = -(win//2), edges truncated, / win."""
    W = x.shape[-1]
    c = np.concatenate([np.zeros(x.shape[:-1] + (1,)), np.cumsum(x, axis=-1)], axis=-1)
    start = np.arange(W) - win // 2
    end = start + win - 1
    lo = np.clip(start, 0, W)
    hi = np.clip(end + 1, 0, W)
    return (c[..., hi] - c[..., lo]) / win


def runabs_window(freq: float, delta: float) -> int:
    fd = _f32(_f32(freq) * _f32(delta))
    return max(2 * int(1.0 / float(fd)) + 1, 1)


def runabs(x: np.ndarray, freq: float, delta: float) -> np.ndarray:
    """x [frames, nch, W]; the weight is shared by the channels of a frame."""
    win = runabs_window(freq, delta)
    w = smooth_rows(np.abs(x), win).sum(axis=1, keepdims=True)
    w = np.where(~np.isfinite(w) | (w < F32_MINVAL), F32_MINVAL, w)
    with np.errstate(invalid="ignore", divide="ignore"):
        y = np.where(~np.isfinite(w) | (w < MINVAL), 0.0, x / w)
    y = np.where(np.isfinite(y), y, 0.0)
    return np.clip(y, -MAXVAL, MAXVAL)


def onebit(x: np.ndarray) -> np.ndarray:
    return np.sign(x)


def runabs_mf(x: np.ndarray, plan: Plan) -> np.ndarray:
    base = _sanitize_c(np.fft.rfft(x, n=plan.nfft_filter, axis=-1) * plan.delta)
    total = np.zeros_like(x)
    for i in range(1, len(plan.responses)):
        Xi = base * plan.responses[i][: plan.nfft_filter // 2 + 1]
        xi = np.fft.irfft(Xi, n=plan.nfft_filter, axis=-1) * plan.nfft_filter
        xi = xi[..., : plan.seg] * _inv_weight(plan.nfft_filter, plan.delta)
        total = total + runabs(xi, plan.freq_lows[i], plan.delta)
    return total


def spectral_taper(nbin: int, idx) -> np.ndarray:
    """Piecewise taper over bins 0..nbin-1 in the reference's if/else order
    (specTaper2DKernel, kernels/taper.cu:8-69, np = 1): 0 below idx1, rising
    raised cosine on [idx1, idx2), flat, falling raised cosine on
    [idx3, idx4), 0 from idx4 (idx4 clamped to nbin)."""
    i1, i2, i3, i4 = idx
    i4 = min(i4, nbin)
    f = np.ones(nbin)
    for col in range(nbin):
        if col < i1:
            f[col] = 0.0
        elif i1 <= col < i2:
            f[col] = (1.0 - math.cos(math.pi / (i2 - i1) * (col - i1))) / 2.0
        elif i3 <= col < i4:
            f[col] = (1.0 + math.cos(math.pi / (i4 - i3) * (col - i3))) / 2.0
        elif col >= i4:
            f[col] = 0.0
    return f


def whiten_window(seg: int, delta: float) -> int:
    return max(int(0.02 * seg * float(_f32(delta))), 1)


def whiten_spectrum(X: np.ndarray, plan: Plan) -> np.ndarray:
    """Divide by the frame's summed smoothed amplitude, taper, sanitise
    (core/spectrum.cu:205-300); X [frames, nch, seg/2+1]."""
    win = whiten_window(plan.seg, plan.delta)
    wgt = smooth_rows(np.abs(X), win).sum(axis=1, keepdims=True)
    wgt = np.where(~np.isfinite(wgt) | (wgt < F32_MINVAL), F32_MINVAL, wgt)
    X = np.where(wgt < MINVAL, 0, X / wgt)
    return _sanitize_c(X * spectral_taper(X.shape[-1], plan.idx))


def freq_whiten(x: np.ndarray, plan: Plan) -> np.ndarray:
    seg = plan.seg
    X = _sanitize_c(np.fft.rfft(x, n=seg, axis=-1) * plan.delta)
    X = whiten_spectrum(X, plan)
    y = np.fft.irfft(X, n=seg, axis=-1) * seg
    return y * _inv_weight(seg, plan.delta)


def output_spectrum(x: np.ndarray, plan: Plan, phase_only: bool = False) -> np.ndarray:
    S = _sanitize_c(np.fft.rfft(x, n=plan.nfft_out, axis=-1)[..., : plan.nspec] * plan.delta)
    if phase_only:
        a = np.abs(S)
        with np.errstate(invalid="ignore", divide="ignore"):
            S = np.where(~np.isfinite(a) | (a <= F32_MINVAL), 0, S / a)
        S = _sanitize_c(S)
    return S


def process_batch(x: np.ndarray, plan: Plan, whiten: int = 2, normalize: int = 3,
                  phase_only: bool = False) -> np.ndarray:
    """x: windows [frames, nch, seg] -> spectra [frames, nch, nspec].

    whiten: 0 none, 1 before, 2 after, 3 both; normalize: 0 none,
    1 RUN-ABS-MF, 2 one-bit, 3 RUN-ABS (worker/compute.cu:52-217)."""
    x = preprocess(x, plan.freq_low, plan.delta)
    x = prefilter(x, plan, 0)
    if whiten in (1, 3):
        x = freq_whiten(x, plan)
    if normalize == 1:
        x = runabs_mf(x, plan)
    elif normalize == 2:
        x = onebit(x)
    elif normalize == 3:
        x = runabs(x, float(_f32(_f32(plan.freq_low) * _f32(0.667))), plan.delta)
    if whiten in (2, 3):
        x = freq_whiten(x, plan)
    return output_spectrum(x, plan, phase_only)


def cut_windows(traces: np.ndarray, plan: Plan, valid_steps=None) -> np.ndarray:
    """traces [files, npts] -> windows [nvalid, files, seg] (worker/compute.cu:34-45)."""
    steps = list(range(plan.nstep)) if valid_steps is None else list(valid_steps)
    out = np.zeros((len(steps), traces.shape[0], plan.seg))
    for o, s in enumerate(steps):
        a = s * plan.shift
        seg = traces[:, a:a + plan.seg]
        out[o, :, : seg.shape[1]] = seg
    return out
