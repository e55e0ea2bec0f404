"""The multi-threaded C stage baselines (oracle/cpu_stages.c) against the FP64
oracle (CPU only): they are the CPU legs the stage benchmarks time, so they
must compute the same thing."""
import numpy as np
import pytest

from oracle import cbind
from oracle import sac2spec as s2s
from oracle import stack as ost
from oracle import xc as oxc

pytestmark = pytest.mark.skipif(not cbind.have_cpu_stages(),
                                reason="oracle/libcpu_stages.so not built (make -C oracle)")


@pytest.fixture(scope="module")
def plan_small(tmp_path_factory):
    path = str(tmp_path_factory.mktemp("f") / "filter.txt")
    s2s.write_filter_file(path, 0.1, "0.1/2")
    # 2 h at 0.1 s, 600 s windows: seg 6000, three windows
    return s2s.plan_lengths(7200, 0.1, 600, 600, 100, s2s.read_filters(path))


def _traces(nfiles, npts, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((nfiles, npts)).astype(np.float32)
    x += np.linspace(-3, 5, npts, dtype=np.float32)  # a trend the preprocess removes
    return x


@pytest.mark.parametrize("whiten,normalize,nch", [(2, 3, 1), (3, 3, 3), (1, 2, 1), (0, 0, 3),
                                                   (2, 2, 3)])
def test_s2s_matches_oracle(plan_small, whiten, normalize, nch):
    p = plan_small
    assert p.seg == 6000
    nfiles = 6
    tr = _traces(nfiles, p.npts + 500, 3 + whiten + 7 * normalize)  # tail beyond npts unused
    steps = [0, 2, 5, 11]  # window 11 runs past the end of the record (zero padded)
    got = cbind.cpu_s2s_process(tr, steps, p, whiten, normalize, nch, nthreads=3)
    win = s2s.cut_windows(tr[:, : p.npts + 500], p, steps)
    fr = win.reshape(len(steps), nfiles // nch, nch, p.seg).reshape(-1, nch, p.seg)
    want = s2s.process_batch(fr, p, whiten, normalize).reshape(len(steps), nfiles, p.nspec)
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert err < 2e-5, err


def test_s2s_rejects_runabs_mf(plan_small):
    with pytest.raises(RuntimeError):
        cbind.cpu_s2s_process(_traces(1, plan_small.npts, 0), [0], plan_small, 2, 1, 1)


def _groups(rng, npairs, T, B, n):
    recs = rng.standard_normal((npairs, T, n)).astype(np.float32)
    recs += np.sin(np.linspace(0, 20, n, dtype=np.float32))  # a common arrival
    return recs


@pytest.mark.parametrize("n,B", [(207, 1), (2001, 3)])
def test_pws_matches_oracle(n, B):
    rng = np.random.default_rng(n + B)
    npairs, T = 4, 7
    recs = _groups(rng, npairs, T, B, n)
    pre, w, lin, off = [], [], [], [0]
    want = []
    for p in range(npairs):
        m, wg, ln = ost.pws_prepare(recs[p], B)
        pre.append(m)
        w.append(wg)
        lin.append(ln)
        off.append(off[-1] + len(wg))
        want.append(ost.pws(m, wg, ln))
    got = cbind.cpu_pws(np.concatenate(pre), np.concatenate(w), off, np.stack(lin), nthreads=2)
    assert oxc.rel_l2(got, np.stack(want)) < 1e-5


@pytest.mark.parametrize("n,B", [(207, 1), (200, 2), (69, 3)])
def test_tfpws_matches_oracle(n, B):
    rng = np.random.default_rng(n * 3 + B)
    npairs, T = 3, 5
    recs = _groups(rng, npairs, T, B, n)
    pre, w, lin, off = [], [], [], [0]
    want = []
    for p in range(npairs):
        s, wg, ln = ost.tfpws_prepare(recs[p], B)
        pre.append(s)
        w.append(wg)
        lin.append(ln)
        off.append(off[-1] + len(wg))
        want.append(ost.tfpws(s, wg, ln))
    got = cbind.cpu_tfpws(np.concatenate(pre), np.concatenate(w), off, np.stack(lin), nthreads=2)
    assert oxc.rel_l2(got, np.stack(want)) < 1e-4


def test_linear_stack_matches_oracle():
    rng = np.random.default_rng(5)
    counts = [1, 4, 9]
    rec = rng.standard_normal((sum(counts), 301)).astype(np.float32)
    off = np.concatenate([[0], np.cumsum(counts)])
    got = cbind.cpu_linear_stack(rec, off, nthreads=2)
    for p in range(len(counts)):
        want = ost.linear_stack(rec[off[p]:off[p + 1]])
        assert np.array_equal(got[p], want)
