"""PWS and TF-PWS on the GPU vs the FP64 oracle (oracle/stack.py).

Lengths cover the direct mixed-radix transform (87 = 3*29, 345, 2001 =
3*23*29) and Bluestein's convolution for lengths with other prime factors
(201 = 3*67, 401 prime, 1001 = 7*11*13), as the correlation length 2H+1 is
arbitrary; at 4999 (prime, convolution length 10000) one TF-PWS workgroup's
LDS holds a single row, so groups and the linear stack take turns in it.  Inputs follow the reference host preparation in float32
(native/pws/io/sourcepack_io.cu:385-478, native/tfpws/io/sourcepack_io.cu:403-490);
the oracle prepares the same records in FP64.  Tolerance: relative L2
1e-5 for PWS, 1e-4 for TF-PWS (a sum over n/2+1 voices of FP32 inverse
transforms).
"""
import numpy as np
import pytest

from oracle import stack
from oracle.xc import rel_l2

pytestmark = pytest.mark.gpu


def _records(rng, T, N):
    t = np.arange(N)
    sig = np.sin(2 * np.pi * 0.05 * (t - N // 2)) * np.exp(-((t - N // 2) / (0.1 * N)) ** 2)
    return (sig[None] + 0.7 * rng.standard_normal((T, N))).astype(np.float32)


def _prep32(recs, B, kind):
    T, N = recs.shape
    g = 1 if B < 2 else B
    pre, w = [], []
    lin = np.zeros(N, np.float32)
    inv_t = np.float32(1.0) / np.float32(T)
    for a in range(0, T, g):
        b = min(a + g, T)
        s = np.zeros(N, np.float32)
        for i in range(a, b):
            s += recs[i]
        if kind == "pws":
            s = s * (np.float32(1.0) / np.float32(b - a))
            wg = np.float32(b - a) / np.float32(T)
            lin += s * wg
        else:
            wg = np.float32(b - a) * inv_t
            lin += s * inv_t
        pre.append(s)
        w.append(wg)
    return np.stack(pre).astype(np.float32), np.array(w, np.float32), lin


def _batch(rng, npairs, T, N, B, kind):
    recs = [_records(rng, T + p, N) for p in range(npairs)]  # ragged group counts
    pre, w, lin, off = [], [], [], [0]
    for r in recs:
        a, b, c = _prep32(r, B, kind)
        pre.append(a)
        w.append(b)
        lin.append(c)
        off.append(off[-1] + len(b))
    return recs, np.concatenate(pre), np.concatenate(w), np.stack(lin), np.array(off, np.int64)


@pytest.mark.parametrize("N", [87, 345, 2001, 201, 401, 1001])
def test_pws_matches_oracle(gpu, N):
    import torch
    from fastxc_amd import ops

    rng = np.random.default_rng(N)
    recs, pre, w, lin, off = _batch(rng, 3, 10, N, 4, "pws")
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(gpu)
    out = ops.pws(d(pre), d(w), off, d(lin)).cpu().numpy()
    for p, r in enumerate(recs):
        g, wg, l = stack.pws_prepare(r, 4)
        want = stack.pws(g, wg, l)
        err = rel_l2(out[p], want)
        print(f"pws N={N} pair={p} rel_l2={err:.3e}")
        assert err < 1e-5


def test_pws_rows_of_unlike_size_and_nonfinite_rows(gpu):
    """Two group rows share one complex transform in the analytic step: rows
    of very different size must keep their own accuracy (pairs scaled by
    1e4 and 1e-4 next to each other), an odd row count leaves one transform
    with a single row, and a non-finite value stays in its own pair."""
    import torch
    from fastxc_amd import ops

    N = 2001
    rng = np.random.default_rng(11)
    recs = [_records(rng, 10 + p, N) * np.float32(s) for p, s in enumerate((1e4, 1e-4, 1.0))]
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(gpu)

    def run(rs):
        pre, w, lin, off = [], [], [], [0]
        for r in rs:
            a, b, c = _prep32(r, 4, "pws")
            pre.append(a)
            w.append(b)
            lin.append(c)
            off.append(off[-1] + len(b))
        assert off[-1] % 2 == 1  # 3 + 3 + 3 groups: an odd row count
        return ops.pws(d(np.concatenate(pre)), d(np.concatenate(w)), np.array(off, np.int64),
                       d(np.stack(lin))).cpu().numpy()

    def check(out, pairs):
        for p in pairs:
            g, wg, l = stack.pws_prepare(recs[p].astype(np.float64), 4)
            err = rel_l2(out[p], stack.pws(g, wg, l))
            print(f"pair {p} rel_l2={err:.3e}")
            assert err < 1e-5

    check(run(recs), range(3))
    bad = [r.copy() for r in recs]
    bad[0][2, 100] = np.nan
    out = run(bad)
    assert not np.isfinite(out[0]).all()
    check(out, (1, 2))


@pytest.mark.parametrize("N,band", [(87, None), (345, None), (2001, None), (345, (0.1, 0.3, 0.02)),
                                    (201, None), (1001, None), (401, (0.1, 0.3, 0.02)),
                                    (4999, None)])
def test_tfpws_matches_oracle(gpu, N, band):
    import torch
    from fastxc_amd import ops

    rng = np.random.default_rng(N + 1)
    recs, pre, w, lin, off = _batch(rng, 2, 8, N, 3, "tfpws")
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(gpu)
    out = ops.tfpws(d(pre), d(w), off, d(lin), band=band, delta=1.0).cpu().numpy()
    for p, r in enumerate(recs):
        s, wg, l = stack.tfpws_prepare(r, 3)
        want = stack.tfpws(s, wg, l, band=band, delta=1.0)
        err = rel_l2(out[p], want)
        print(f"tfpws N={N} band={band} pair={p} rel_l2={err:.3e}")
        assert err < 1e-4


def test_tfpws_identical_groups_vanish(gpu):
    import torch
    from fastxc_amd import ops

    rng = np.random.default_rng(5)
    tr = rng.standard_normal(345).astype(np.float32)
    recs = np.tile(tr, (6, 1))
    pre, w, lin = _prep32(recs, 2, "tfpws")
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(gpu)
    out = ops.tfpws(d(pre / 2), d(w), np.array([0, 3]), d(lin[None])).cpu().numpy()
    assert np.max(np.abs(out)) < 1e-4 * np.max(np.abs(lin))


@pytest.mark.parametrize("kind", ["pws", "tfpws"])
def test_stack_rejects_length_beyond_lds(gpu, kind):
    """A length whose transforms do not fit one workgroup's LDS is an argument
    error before any launch (no truncated or faulting run)."""
    import torch
    from fastxc_amd import ops
    from fastxc_amd._lib import FxcError

    n = 9001 if kind == "tfpws" else 15013
    z = torch.zeros((2, n), dtype=torch.float32, device=gpu)
    w = torch.full((2,), 0.5, dtype=torch.float32, device=gpu)
    with pytest.raises(FxcError, match="too long"):
        getattr(ops, kind)(z, w, np.array([0, 2]), z[:1].contiguous())


@pytest.mark.parametrize("N", [2001, 4001])
def test_pws_hilbert_rotation_closed_form(gpu, N):
    """Two equal-weight groups, the second the first Hilbert-rotated by theta:
    the reference's PWS (power 1) is lin |cos(theta/2)| exactly; the spec
    mode's power 2 gives lin cos^2(theta/2)."""
    import torch
    from fastxc_amd import ops

    rng = np.random.default_rng(40 + N)
    theta = 2 * np.pi / 3
    half = (N + 1) // 2
    X1 = np.zeros(N, complex)
    X1[1:half] = rng.standard_normal(half - 1) + 1j * rng.standard_normal(half - 1)
    X1[N - half + 1:] = np.conj(X1[1:half][::-1])
    X2 = X1.copy()
    X2[1:half] *= np.exp(-1j * theta)
    X2[N - half + 1:] = np.conj(X2[1:half][::-1])
    x1, x2 = np.fft.ifft(X1).real, np.fft.ifft(X2).real
    pre = np.stack([x1, x2]).astype(np.float32)
    w = np.array([0.5, 0.5], np.float32)
    lin = (0.5 * (pre[0] + pre[1]))[None].astype(np.float32)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(gpu)
    c = abs(np.cos(theta / 2))
    for power in (1.0, 2.0):
        out = ops.pws(d(pre), d(w), np.array([0, 2]), d(lin), power=power).cpu().numpy()[0]
        want = lin[0].astype(np.float64) * c ** power
        err = rel_l2(out, want)
        print(f"pws hilbert N={N} power={power} rel_l2={err:.3e}")
        assert err < 1e-5
        assert rel_l2(out, stack.pws(pre.astype(np.float64), w.astype(np.float64), lin[0],
                                     power=power)) < 1e-5


@pytest.mark.parametrize("N,power,nvoices,band", [(2001, 2.0, 0, None), (2001, 2.0, 256, None),
                                                  (2001, 1.0, 256, (0.1, 0.3, 0.02)),
                                                  (401, 1.5, 64, None), (4001, 2.0, 256, None)])
def test_stack_spec_mode_matches_oracle(gpu, N, power, nvoices, band):
    """Spec mode (power and voice count) of both stacks against the FP64 oracle."""
    import torch
    from fastxc_amd import ops

    rng = np.random.default_rng(N + 7)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(gpu)
    recs, pre, w, lin, off = _batch(rng, 2, 6, N, 2, "pws")
    out = ops.pws(d(pre), d(w), off, d(lin), power=power).cpu().numpy()
    for p, r in enumerate(recs):
        g, wg, l = stack.pws_prepare(r, 2)
        err = rel_l2(out[p], stack.pws(g, wg, l, power=power))
        print(f"pws spec N={N} power={power} pair={p} rel_l2={err:.3e}")
        assert err < 1e-5
    recs, pre, w, lin, off = _batch(rng, 2, 6, N, 2, "tfpws")
    out = ops.tfpws(d(pre), d(w), off, d(lin), band=band, delta=1.0, power=power,
                    nvoices=nvoices).cpu().numpy()
    for p, r in enumerate(recs):
        s, wg, l = stack.tfpws_prepare(r, 2)
        want = stack.tfpws(s, wg, l, band=band, delta=1.0, power=power, nvoices=nvoices)
        err = rel_l2(out[p], want)
        print(f"tfpws spec N={N} power={power} nvoices={nvoices} pair={p} rel_l2={err:.3e}")
        assert err < 1e-4


@pytest.mark.parametrize("N", [200, 2000])
def test_tfpws_even_length_matches_oracle(gpu, N):
    """Even n: the Gaussian is 0 at the shift c = n/2 (cuda.stransform.cu:119-141)."""
    import torch
    from fastxc_amd import ops

    rng = np.random.default_rng(N + 3)
    recs, pre, w, lin, off = _batch(rng, 2, 6, N, 3, "tfpws")
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(gpu)
    out = ops.tfpws(d(pre), d(w), off, d(lin), delta=1.0).cpu().numpy()
    for p, r in enumerate(recs):
        s, wg, l = stack.tfpws_prepare(r, 3)
        err = rel_l2(out[p], stack.tfpws(s, wg, l, delta=1.0))
        print(f"tfpws even N={N} pair={p} rel_l2={err:.3e}")
        assert err < 1e-4
