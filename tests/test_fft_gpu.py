"""Workgroup LDS FFT vs numpy (FP64), on the GPU; lengths whose transform
exceeds one workgroup's LDS (complex 36000 and 37044, real 72000, 72576 and
74088: the dt 0.05 s lengths) run the four-step transform over HBM."""
import numpy as np
import pytest

from oracle.xc import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [12, 98, 2001, 18000, 18144, 18522, 36000, 37044])
def test_c2c_forward_inverse(gpu, n):
    import torch
    from fastxc_amd import ops

    rng = np.random.default_rng(n)
    x = (rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))).astype(np.complex64)
    xd = torch.from_numpy(x).to(gpu)
    f = ops.fft_c2c(xd).cpu().numpy()
    assert rel_l2(f, np.fft.fft(x.astype(np.complex128))) < 1e-6
    b = ops.fft_c2c(xd, inverse=True).cpu().numpy()
    assert rel_l2(b, np.fft.ifft(x.astype(np.complex128)) * n) < 1e-6


@pytest.mark.parametrize("n,in_len", [(36000, 36000), (36288, 36000), (37044, 36000), (24, 20),
                                      (72000, 72000), (72576, 72000), (74088, 72000)])
def test_r2c_c2r(gpu, n, in_len):
    import torch
    from fastxc_amd import ops

    rng = np.random.default_rng(n)
    x = rng.standard_normal((4, in_len)).astype(np.float32)
    X = ops.fft_r2c(torch.from_numpy(x).to(gpu), n=n).cpu().numpy()
    want = np.fft.rfft(x.astype(np.float64), n=n)
    assert rel_l2(X, want) < 1e-6
    assert np.all(X[:, 0].imag == 0) and np.all(X[:, -1].imag == 0)
    y = ops.fft_c2r(torch.from_numpy(want.astype(np.complex64)).to(gpu), n).cpu().numpy()
    assert rel_l2(y, np.fft.irfft(want, n=n) * n) < 1e-6
