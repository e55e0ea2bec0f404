"""fxc_pws_noalloc (workspace in a fixed device array) against fxc_pws, through
the C ABI: two calls in a row, the second smaller, each equal to fxc_pws on
the same inputs and within 1e-5 relative L2 of the FP64 oracle."""
import ctypes

import numpy as np
import pytest

from oracle import stack
from oracle.xc import rel_l2

pytestmark = pytest.mark.gpu


def _inputs(rng, npairs, T, N, B):
    t = np.arange(N)
    sig = np.sin(2 * np.pi * 0.05 * (t - N // 2)) * np.exp(-((t - N // 2) / (0.1 * N)) ** 2)
    recs, pre, w, lin, off = [], [], [], [], [0]
    for p in range(npairs):
        r = (sig[None] + 0.7 * rng.standard_normal((T + p, N))).astype(np.float32)
        recs.append(r)
        g, wg, l = stack.pws_prepare(r, B)
        pre.append(g.astype(np.float32))
        w.append(wg.astype(np.float32))
        lin.append(l.astype(np.float32))
        off.append(off[-1] + len(wg))
    return recs, np.concatenate(pre), np.concatenate(w), np.stack(lin), np.array(off, np.int64)


def test_pws_noalloc_matches_pws(gpu):
    import torch
    from fastxc_amd._lib import lib

    L = lib()
    N, B = 2001, 3
    assert L.fxc_stack_prepare(N) == 0, L.fxc_last_error()
    rng = np.random.default_rng(77)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(gpu)
    for npairs in (6, 2):  # the second call is smaller
        recs, pre, w, lin, off = _inputs(rng, npairs, 8, N, B)
        dp, dw, dl = d(pre), d(w), d(lin)
        out_a = torch.empty_like(dl)
        out_b = torch.full_like(dl, float("nan"))
        torch.cuda.synchronize()
        assert L.fxc_pws(dp.data_ptr(), dw.data_ptr(), off.ctypes.data, dl.data_ptr(), npairs, N,
                         out_a.data_ptr(), None) == 0, L.fxc_last_error()
        assert L.fxc_pws_noalloc(dp.data_ptr(), dw.data_ptr(), off.ctypes.data, dl.data_ptr(),
                                 npairs, N, out_b.data_ptr(), None) == 0, L.fxc_last_error()
        a, b = out_a.cpu().numpy(), out_b.cpu().numpy()
        print(f"noalloc npairs={npairs} max|a-b|={np.abs(a - b).max():.3e}")
        assert np.array_equal(a, b)
        for p, r in enumerate(recs):
            g, wg, l = stack.pws_prepare(r, B)
            assert rel_l2(b[p], stack.pws(g, wg, l)) < 1e-5


def test_pws_noalloc_rejects_what_does_not_fit(gpu):
    import torch
    from fastxc_amd._lib import lib

    L = lib()
    N = 2001
    assert L.fxc_stack_prepare(N) == 0, L.fxc_last_error()
    # 5000 pairs x 2 groups of 2001 values: 15000 rows x 16 KB > 64 MiB
    npairs = 5000
    off = np.arange(0, 2 * npairs + 1, 2, dtype=np.int64)
    z = torch.zeros((2 * npairs, N), dtype=torch.float32, device=gpu)
    w = torch.full((2 * npairs,), 0.5, dtype=torch.float32, device=gpu)
    lin = torch.zeros((npairs, N), dtype=torch.float32, device=gpu)
    out = torch.zeros_like(lin)
    rc = L.fxc_pws_noalloc(z.data_ptr(), w.data_ptr(), off.ctypes.data, lin.data_ptr(), npairs, N,
                           out.data_ptr(), None)
    assert rc == 1 and b"fixed scratch array" in L.fxc_last_error()
    # a length whose plans were never prepared is refused, not built
    rc = L.fxc_pws_noalloc(z.data_ptr(), w.data_ptr(), off[:2].ctypes.data, lin.data_ptr(), 1, 1999,
                           out.data_ptr(), None)
    assert rc == 1 and b"not prepared" in L.fxc_last_error()
