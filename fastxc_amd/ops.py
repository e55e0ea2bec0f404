"""Torch-facing wrappers over the C ABI (device tensors in, device tensors out).

Each wrapper validates shapes on the host before any launch, then enqueues the
HIP kernels on torch's current stream.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from ._lib import PRECISIONS, S2sConfig, StackSpec, check, lib


def _stream_ptr(device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _require(t: torch.Tensor, dtype, name: str) -> None:
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor")
    if t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


def fft_c2c(x: torch.Tensor, inverse: bool = False, scale: float = 1.0) -> torch.Tensor:
    """Batched complex FFT along the last axis of a complex64 tensor."""
    _require(x, torch.complex64, "x")
    n = x.shape[-1]
    batch = x.numel() // n if n else 0
    out = torch.empty_like(x)
    check(lib().fxc_fft_c2c(x.data_ptr(), n, out.data_ptr(), n, n, batch, int(inverse),
                            float(scale), _stream_ptr(x.device)), "fxc_fft_c2c")
    return out


def fft_r2c(x: torch.Tensor, n: int | None = None, scale: float = 1.0) -> torch.Tensor:
    """Batched real FFT (rows zero-padded to n) -> bins 0..n/2 as complex64."""
    _require(x, torch.float32, "x")
    in_len = x.shape[-1]
    n = in_len if n is None else n
    batch = x.numel() // in_len if in_len else 0
    out = torch.empty(x.shape[:-1] + (n // 2 + 1,), dtype=torch.complex64, device=x.device)
    check(lib().fxc_fft_r2c(x.data_ptr(), in_len, in_len, out.data_ptr(), n // 2 + 1, n, batch,
                            float(scale), _stream_ptr(x.device)), "fxc_fft_r2c")
    return out


def fft_c2r(X: torch.Tensor, n: int, out_len: int | None = None, scale: float = 1.0) -> torch.Tensor:
    """Batched real inverse FFT of bins 0..n/2 (unnormalised); first out_len samples."""
    _require(X, torch.complex64, "X")
    if X.shape[-1] != n // 2 + 1:
        raise ValueError("X must hold n/2+1 bins")
    out_len = n if out_len is None else out_len
    batch = X.numel() // X.shape[-1]
    out = torch.empty(X.shape[:-1] + (out_len,), dtype=torch.float32, device=X.device)
    check(lib().fxc_fft_c2r(X.data_ptr(), n // 2 + 1, out.data_ptr(), out_len, out_len, n, batch,
                            float(scale), _stream_ptr(X.device)), "fxc_fft_c2r")
    return out


class XcPlan:
    """One timestamp's windowed cross-correlation (xc_fast hot path).

    spectra: complex64 [nstep, nsta, nspec] (StepPack payload order).
    """

    def __init__(self, nspec: int, nstep: int, dt: float, half_cc: int,
                 precision: str = "fp32", chunk_rows: int = 0):
        self._h = ctypes.c_void_p()
        check(lib().fxc_xc_plan_create(ctypes.byref(self._h), nspec, nstep, float(dt), half_cc,
                                       PRECISIONS[precision]), "fxc_xc_plan_create")
        self.nspec, self.nstep, self.dt, self.half_cc = nspec, nstep, dt, half_cc
        self.cc_size = 2 * half_cc + 1
        self.nfft = 2 * (nspec - 1)
        self.precision = precision
        self.npairs = 0
        if chunk_rows:
            check(lib().fxc_xc_set_chunk_rows(self._h, int(chunk_rows)), "fxc_xc_set_chunk_rows")

    def close(self):
        if self._h:
            lib().fxc_xc_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_pairs(self, nsta: int, src, rec) -> None:
        src = np.ascontiguousarray(src, dtype=np.int32)
        rec = np.ascontiguousarray(rec, dtype=np.int32)
        if src.shape != rec.shape or src.ndim != 1:
            raise ValueError("src and rec must be equal-length 1-D arrays")
        if src.size and (src.min() < 0 or rec.min() < 0 or src.max() >= nsta or rec.max() >= nsta):
            raise ValueError("station index out of range")
        check(lib().fxc_xc_set_pairs(self._h, int(nsta), src.ctypes.data, rec.ctypes.data,
                                     int(src.size)), "fxc_xc_set_pairs")
        self.nsta = nsta
        self.npairs = int(src.size)

    def pairs_info(self):
        a, b, c = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
        check(lib().fxc_xc_pairs_info(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)),
              "fxc_xc_pairs_info")
        return {"ntiles": a.value, "nchunks": b.value, "max_chunk_rows": c.value}

    def _check_spec(self, spec: torch.Tensor) -> None:
        _require(spec, torch.complex64, "spec")
        if tuple(spec.shape) != (self.nstep, self.nsta, self.nspec):
            raise ValueError(f"spec shape {tuple(spec.shape)} != "
                             f"{(self.nstep, self.nsta, self.nspec)}")

    def cross_spectra(self, spec: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        self._check_spec(spec)
        if out is None:
            out = torch.empty((self.npairs, self.nspec), dtype=torch.complex64, device=spec.device)
        _require(out, torch.complex64, "out")
        check(lib().fxc_xc_cross_spectra(self._h, spec.data_ptr(), out.data_ptr(),
                                         _stream_ptr(spec.device)), "fxc_xc_cross_spectra")
        return out

    def lags(self, xspec: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        _require(xspec, torch.complex64, "xspec")
        if xspec.shape[-1] != self.nspec:
            raise ValueError("xspec rows must hold nspec bins")
        rows = xspec.numel() // self.nspec
        if out is None:
            out = torch.empty((rows, self.cc_size), dtype=torch.float32, device=xspec.device)
        _require(out, torch.float32, "out")
        check(lib().fxc_xc_lags(self._h, xspec.data_ptr(), rows, out.data_ptr(),
                                _stream_ptr(xspec.device)), "fxc_xc_lags")
        return out

    def set_band(self, k_lo: int | None = None, k_hi: int | None = None) -> None:
        """Spec mode: form the windowed stack only for bins k_lo..k_hi (rounded
        out to 16-bin chunks), every other cross-spectrum bin 0.  No
        arguments: the parity mode (every bin)."""
        lo = 0 if k_lo is None else int(k_lo)
        hi = self.nspec - 1 if k_hi is None else int(k_hi)
        check(lib().fxc_xc_set_band(self._h, lo, hi), "fxc_xc_set_band")

    def set_timing(self, enable: bool) -> None:
        check(lib().fxc_xc_set_timing(self._h, int(enable)), "fxc_xc_set_timing")

    def get_timing(self):
        ms = (ctypes.c_double * 3)()
        n = (ctypes.c_longlong * 3)()
        check(lib().fxc_xc_get_timing(self._h, ms, n), "fxc_xc_get_timing")
        return {"split_ms": ms[0], "contract_ms": ms[1], "lag_ms": ms[2],
                "split_launches": n[0], "contract_launches": n[1], "lag_launches": n[2]}

    def correlate(self, spec: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        self._check_spec(spec)
        if out is None:
            out = torch.empty((self.npairs, self.cc_size), dtype=torch.float32, device=spec.device)
        _require(out, torch.float32, "out")
        if tuple(out.shape) != (self.npairs, self.cc_size):
            raise ValueError("out must be [npairs, 2*half_cc+1]")
        check(lib().fxc_xc_correlate(self._h, spec.data_ptr(), out.data_ptr(),
                                     _stream_ptr(spec.device)), "fxc_xc_correlate")
        return out


class XcBatch:
    """Uneven time segments correlated in one set of launches (split-BF16).

    Segment s has nsta[s] stations and nstep[s] windows; its pairs are all
    (i, j), i < j, in row-major order, and output rows are the segments'
    pairs one segment after another.  spec: complex64, the segments'
    [nstep, nsta, nspec] spectra flattened back to back (see pack); valid:
    None or uint8, one byte per station-window ([nstep, nsta] per segment,
    back to back), 0 = a gap, which enters the stack as zeros.
    gaps="keep" divides by the segment's window count (the reference's rule
    for zero-filled windows); gaps="exclude" divides each pair by the windows
    in which both stations are valid.
    """

    GAPS = {"keep": 0, "exclude": 1}

    def __init__(self, nspec: int, dt: float, half_cc: int, nsta, nstep, gaps: str = "keep",
                 precision: str = "bf16x3", max_rows: int = 0):
        self.nsta = [int(n) for n in nsta]
        self.nstep = [int(n) for n in nstep]
        if len(self.nsta) != len(self.nstep) or not self.nsta:
            raise ValueError("nsta and nstep must be equal-length, non-empty")
        if gaps not in self.GAPS:
            raise ValueError(f"gaps must be one of {sorted(self.GAPS)}")
        self._h = ctypes.c_void_p()
        a = np.ascontiguousarray(self.nsta, dtype=np.int32)
        b = np.ascontiguousarray(self.nstep, dtype=np.int32)
        check(lib().fxc_xc_batch_create(ctypes.byref(self._h), int(nspec), float(dt),
                                        int(half_cc), PRECISIONS[precision], len(a),
                                        a.ctypes.data, b.ctypes.data, self.GAPS[gaps],
                                        int(max_rows)), "fxc_xc_batch_create")
        self.nspec, self.dt, self.half_cc, self.gaps = nspec, dt, half_cc, gaps
        self.cc_size = 2 * half_cc + 1
        info = self.info()
        self.ntasks, self.spec_elems, self.valid_elems = (info["ntasks"], info["spec_elems"],
                                                          info["valid_elems"])
        self.task_base = np.concatenate([[0], np.cumsum([n * (n - 1) // 2 for n in self.nsta])])
        self.max_rows = int(max_rows)
        self._stream = None

    def close(self):
        if self._h:
            lib().fxc_xc_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        t, s, v = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
        g, m = ctypes.c_int(), ctypes.c_int()
        check(lib().fxc_xc_batch_info(self._h, ctypes.byref(t), ctypes.byref(s), ctypes.byref(v),
                                      ctypes.byref(g), ctypes.byref(m)), "fxc_xc_batch_info")
        return {"ntasks": t.value, "spec_elems": s.value, "valid_elems": v.value,
                "groups": g.value, "grouping": m.value}

    def set_grouping(self, grouping: int) -> None:
        check(lib().fxc_xc_batch_set_grouping(self._h, int(grouping)),
              "fxc_xc_batch_set_grouping")

    @staticmethod
    def pack(parts, device) -> torch.Tensor:
        """Concatenates per-segment arrays (host or device) into one flat device tensor."""
        return torch.cat([torch.as_tensor(p).reshape(-1) for p in parts]).to(device).contiguous()

    def segment_rows(self, s: int) -> slice:
        return slice(int(self.task_base[s]), int(self.task_base[s + 1]))

    def _args(self, spec, valid, out):
        _require(spec, torch.complex64, "spec")
        if spec.numel() != self.spec_elems:
            raise ValueError(f"spec holds {spec.numel()} values, the segments {self.spec_elems}")
        vp = None
        if valid is not None:
            _require(valid, torch.uint8, "valid")
            if valid.numel() != self.valid_elems or valid.device != spec.device:
                raise ValueError(f"valid must hold {self.valid_elems} bytes on spec's device")
            vp = valid.data_ptr()
        if out is None:
            out = torch.empty((self.ntasks, self.cc_size), dtype=torch.float32, device=spec.device)
        _require(out, torch.float32, "out")
        if tuple(out.shape) != (self.ntasks, self.cc_size):
            raise ValueError("out must be [ntasks, 2*half_cc+1]")
        return vp, out

    def correlate(self, spec: torch.Tensor, valid: torch.Tensor | None = None,
                  out: torch.Tensor | None = None) -> torch.Tensor:
        vp, out = self._args(spec, valid, out)
        check(lib().fxc_xc_batch_correlate(self._h, spec.data_ptr(), vp, out.data_ptr(),
                                           _stream_ptr(spec.device)), "fxc_xc_batch_correlate")
        return out

    def tune(self, spec: torch.Tensor, valid: torch.Tensor | None = None,
             out: torch.Tensor | None = None, reps: int = 2):
        """Times grouping 0 and 1 on this data and keeps the faster."""
        vp, out = self._args(spec, valid, out)
        ms = (ctypes.c_double * 2)()
        check(lib().fxc_xc_batch_tune(self._h, spec.data_ptr(), vp, out.data_ptr(), int(reps),
                                      _stream_ptr(spec.device), ms), "fxc_xc_batch_tune")
        return {"ms": [ms[0], ms[1]], "grouping": self.info()["grouping"]}

    def set_rows(self, max_rows: int) -> None:
        check(lib().fxc_xc_batch_set_rows(self._h, int(max_rows)), "fxc_xc_batch_set_rows")
        self.max_rows = int(max_rows)

    def tune_key(self) -> str:
        """Class of this batch for the tuning cache: segment count, median and
        largest station count (to 50), largest window count (to 8)."""
        a = np.asarray(self.nsta)
        return (f"seg{len(a)}_sta{int(np.median(a)) // 50 * 50}-{int(a.max()) // 50 * 50}"
                f"_win{(max(self.nstep) + 7) // 8 * 8}_nspec{self.nspec}")

    def autotune(self, spec: torch.Tensor, valid: torch.Tensor | None = None,
                 out: torch.Tensor | None = None, rows=(32768, 65536, 131072), reps: int = 1,
                 cache: str | None = None, budget_gb: float | None = None):
        """Chooses the plan grouping (0: one plan; 1: one per operand depth)
        and the row chunk of every plan's persistent launches by timing each
        candidate on this data (HIP events on the current stream), skipping
        candidates whose cross-spectrum workspaces exceed budget_gb; the
        choice is kept in the JSON file `cache` under tune_key(), and a batch
        of a cached class takes it without timing."""
        import json
        import os

        key = self.tune_key()
        table = {}
        if cache and os.path.exists(cache):
            with open(cache) as f:
                table = json.load(f)
        if key in table:
            g, r = table[key]["grouping"], table[key]["max_rows"]
            self.set_grouping(g)
            self.set_rows(r)
            return {"key": key, "grouping": g, "max_rows": r, "cached": True, "ms": None}
        xpitch = (self.nspec + 15) // 16 * 16
        trials = []
        cur = torch.cuda.current_stream(spec.device)
        for g in (0, 1):
            for r in rows:
                # the groups share one cross-spectrum workspace of the largest chunk
                ws = min(r, self.ntasks) * xpitch * 8 / 1e9
                if budget_gb is not None and ws > budget_gb:
                    trials.append({"grouping": g, "max_rows": r, "ms": None, "skipped_gb": ws})
                    continue
                self.set_grouping(g)
                self.set_rows(r)
                self.correlate(spec, valid, out)  # creates the workspaces
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(cur)
                for _ in range(reps):
                    self.correlate(spec, valid, out)
                e1.record(cur)
                e1.synchronize()
                trials.append({"grouping": g, "max_rows": r, "ms": e0.elapsed_time(e1) / reps})
                torch.cuda.empty_cache()
        best = min((t for t in trials if t["ms"] is not None), key=lambda t: t["ms"])
        self.set_grouping(best["grouping"])
        self.set_rows(best["max_rows"])
        if cache:
            table[key] = {"grouping": best["grouping"], "max_rows": best["max_rows"],
                          "ms": round(best["ms"], 3)}
            os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
            with open(cache, "w") as f:
                json.dump(table, f, indent=1, sort_keys=True)
        return {"key": key, "grouping": best["grouping"], "max_rows": best["max_rows"],
                "cached": False, "trials": trials}

    def capture(self, spec: torch.Tensor, valid: torch.Tensor | None = None,
                out: torch.Tensor | None = None) -> torch.Tensor:
        """Records one run into a HIP graph bound to these tensors (on a side
        stream ordered after the current one); replay() re-runs it."""
        vp, out = self._args(spec, valid, out)
        cur = torch.cuda.current_stream(spec.device)
        self._stream = torch.cuda.Stream(device=spec.device)
        self._stream.wait_stream(cur)
        check(lib().fxc_xc_batch_capture(self._h, spec.data_ptr(), vp, out.data_ptr(),
                                         ctypes.c_void_p(self._stream.cuda_stream)),
              "fxc_xc_batch_capture")
        cur.wait_stream(self._stream)
        self._graph_io = (spec, valid, out)
        return out

    def replay(self) -> torch.Tensor:
        if self._stream is None:
            raise RuntimeError("capture() first")
        spec = self._graph_io[0]
        cur = torch.cuda.current_stream(spec.device)
        self._stream.wait_stream(cur)
        check(lib().fxc_xc_batch_replay(self._h, ctypes.c_void_p(self._stream.cuda_stream)),
              "fxc_xc_batch_replay")
        cur.wait_stream(self._stream)
        return self._graph_io[2]


def batch_pws(batch, cc, power=1.0):
    """PWS of every station pair over the batch's segments, in FP32.

    Segment s's stations are taken as stations 0..nsta[s]-1 of one array, so
    the pair (i, j) recurs in every segment with nsta > j.  Its records are
    those segments' rows of `cc` (the batch output, segment order = time
    order); they form sub-stacks of one record each, and the stack is the
    reference's PWS over them (fxc_pws; native/pws/io/sourcepack_io.cu:385-478
    with B = 1: weights 1/T, lin the mean).  Returns (src, rec, pws) with pws
    [npairs, 2H+1] on cc's device, pairs ordered (i, j) row-major."""
    nsta = np.asarray(batch.nsta, dtype=np.int64)
    M = int(nsta.max())
    gids, rows = [], []
    for s, n in enumerate(nsta):
        i, j = np.triu_indices(int(n), k=1)
        gids.append(i * M - i * (i + 1) // 2 + (j - i - 1))
        rows.append(batch.task_base[s] + np.arange(len(i)))
    gid = np.concatenate(gids)
    row = np.concatenate(rows)
    order = np.argsort(gid, kind="stable")       # segment order within a pair
    gid, row = gid[order], row[order]
    npairs_all = M * (M - 1) // 2
    cnt = np.bincount(gid, minlength=npairs_all)
    present = np.flatnonzero(cnt)
    off = np.zeros(len(present) + 1, dtype=np.int64)
    off[1:] = np.cumsum(cnt[present])
    dev = cc.device
    rows_d = torch.from_numpy(row).to(dev)
    pre = cc.index_select(0, rows_d)               # [groups, n] FP32 records
    pair_of_group = torch.from_numpy(np.repeat(np.arange(len(present)), cnt[present])).to(dev)
    counts = torch.from_numpy(cnt[present].astype(np.float32)).to(dev)
    lin = torch.zeros((len(present), cc.shape[1]), dtype=torch.float32, device=dev)
    lin.index_add_(0, pair_of_group, pre)
    lin /= counts[:, None]
    w = (1.0 / counts).index_select(0, pair_of_group).contiguous()
    out = torch.empty_like(lin)
    pws(pre, w, off, lin, out=out, power=power)
    ii, jj = np.triu_indices(M, k=1)
    return ii[present], jj[present], out


class S2sPlan:
    """sac2spec per-window pipeline (preprocess, prefilter, normalisation,
    whitening, output spectrum).

    responses: float32 [filter_count, nfft_filter//2 + 1] power responses
    (band 0 = broad band); freq_lows: [filter_count]; idx: whitening taper
    bins on the window grid.
    """

    def __init__(self, npts, seg, shift, nfft_filter, nfft_out, num_ch, delta, freq_low, idx,
                 responses, freq_lows, whiten=2, normalize=3, phase_only=False):
        resp = np.ascontiguousarray(responses, dtype=np.float32)
        if resp.ndim != 2 or resp.shape[1] != nfft_filter // 2 + 1:
            raise ValueError("responses must be [filter_count, nfft_filter//2+1]")
        lows = np.ascontiguousarray(freq_lows, dtype=np.float32)
        cfg = S2sConfig(npts, seg, shift, nfft_filter, nfft_out, num_ch, float(delta),
                        float(freq_low), (ctypes.c_int * 4)(*[int(v) for v in idx]), int(whiten),
                        int(normalize), int(bool(phase_only)), resp.shape[0], resp.ctypes.data,
                        lows.ctypes.data)
        self._h = ctypes.c_void_p()
        check(lib().fxc_s2s_plan_create(ctypes.byref(self._h), ctypes.byref(cfg)),
              "fxc_s2s_plan_create")
        self.npts, self.seg, self.num_ch = npts, seg, num_ch
        self.nspec = nfft_out // 2 + 1

    def info(self):
        a, b, c, d = (ctypes.c_int() for _ in range(4))
        check(lib().fxc_s2s_plan_info(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                      ctypes.byref(d)), "fxc_s2s_plan_info")
        return {"nspec": a.value, "taper_size": b.value, "runabs_win": c.value,
                "whiten_win": d.value}

    def close(self):
        if self._h:
            lib().fxc_s2s_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_sums(self, compensated: bool) -> None:
        """Spec mode: compensated-FP32 running means instead of FP64 window
        sums (fused one-channel path only)."""
        check(lib().fxc_s2s_set_sums(self._h, int(bool(compensated))), "fxc_s2s_set_sums")

    def process(self, traces: torch.Tensor, steps, out: torch.Tensor | None = None) -> torch.Tensor:
        """traces: float32 [nfiles, npts] (device); steps: window indices.
        Returns complex64 [len(steps), nfiles, nspec]."""
        _require(traces, torch.float32, "traces")
        if traces.ndim != 2 or traces.shape[1] != self.npts:
            raise ValueError("traces must be [nfiles, npts]")
        nfiles = traces.shape[0]
        if nfiles % self.num_ch:
            raise ValueError("nfiles must be a multiple of num_ch")
        st = np.ascontiguousarray(steps, dtype=np.int32)
        if out is None:
            out = torch.empty((len(st), nfiles, self.nspec), dtype=torch.complex64,
                              device=traces.device)
        _require(out, torch.complex64, "out")
        check(lib().fxc_s2s_process(self._h, traces.data_ptr(), nfiles, st.ctypes.data, len(st),
                                    out.data_ptr(), _stream_ptr(traces.device)),
              "fxc_s2s_process")
        return out


def _stack_args(prestack, weights, offsets, linear, out):
    _require(prestack, torch.float32, "prestack")
    _require(weights, torch.float32, "weights")
    _require(linear, torch.float32, "linear")
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    npairs, n = linear.shape
    if prestack.ndim != 2 or prestack.shape[1] != n or prestack.shape[0] != off[-1]:
        raise ValueError("prestack must be [offsets[-1], n]")
    if weights.shape != (prestack.shape[0],) or off.shape != (npairs + 1,):
        raise ValueError("weights/offsets shape mismatch")
    if out is None:
        out = torch.empty_like(linear)
    _require(out, torch.float32, "out")
    return off, npairs, n, out


def pws(prestack, weights, offsets, linear, out=None, power=1.0):
    """Phase-weighted stack of every pair; prestack rows are group means.
    power 1 is the reference (fxc_pws); other powers go through fxc_pws_spec."""
    off, npairs, n, out = _stack_args(prestack, weights, offsets, linear, out)
    if power == 1.0:
        check(lib().fxc_pws(prestack.data_ptr(), weights.data_ptr(), off.ctypes.data,
                            linear.data_ptr(), npairs, n, out.data_ptr(),
                            _stream_ptr(linear.device)), "fxc_pws")
        return out
    spec = StackSpec(float(power), 0)
    check(lib().fxc_pws_spec(prestack.data_ptr(), weights.data_ptr(), off.ctypes.data,
                             linear.data_ptr(), npairs, n, ctypes.byref(spec), out.data_ptr(),
                             _stream_ptr(linear.device)), "fxc_pws_spec")
    return out


def tfpws(prestack, weights, offsets, linear, band=None, delta=1.0, out=None, power=1.0,
          nvoices=0):
    """Time-frequency PWS (S-transform, scale 0.1); prestack rows are group sums.
    band: None or (fmin, fmax, taper_hz).  power 1 over every voice (nvoices 0)
    is the reference (fxc_tfpws); otherwise fxc_tfpws_spec."""
    off, npairs, n, out = _stack_args(prestack, weights, offsets, linear, out)
    bl = band is not None
    fmin, fmax, taper = band if bl else (0.0, 0.0, 0.0)
    args = (prestack.data_ptr(), weights.data_ptr(), off.ctypes.data, linear.data_ptr(), npairs,
            n, int(bl), float(delta), float(fmin), float(fmax), float(taper))
    if power == 1.0 and nvoices == 0:
        check(lib().fxc_tfpws(*args, out.data_ptr(), _stream_ptr(linear.device)), "fxc_tfpws")
        return out
    spec = StackSpec(float(power), int(nvoices))
    check(lib().fxc_tfpws_spec(*args, ctypes.byref(spec), out.data_ptr(),
                               _stream_ptr(linear.device)), "fxc_tfpws_spec")
    return out


def linear_stack(records, offsets, out=None):
    """Linear stack of every pair: float64 running sum of its rows (row order)
    divided by the row count, rounded once to float32.  records [offsets[-1], n]."""
    _require(records, torch.float32, "records")
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    npairs = off.shape[0] - 1
    if records.ndim != 2 or records.shape[0] != off[-1] or npairs < 1:
        raise ValueError("records must be [offsets[-1], n]")
    n = records.shape[1]
    if out is None:
        out = torch.empty((npairs, n), dtype=torch.float32, device=records.device)
    _require(out, torch.float32, "out")
    if tuple(out.shape) != (npairs, n):
        raise ValueError("out must be [npairs, n]")
    check(lib().fxc_linear_stack(records.data_ptr(), off.ctypes.data, npairs, n, out.data_ptr(),
                                 _stream_ptr(records.device)), "fxc_linear_stack")
    return out


def rotate_rtz(enz, trig, out=None):
    """ENZ -> RTZ rotation of every pair: enz [npairs, 9, n] (EE EN EZ NE NN NZ
    ZE ZN ZZ), trig [npairs, 4] float (sin az, cos az, sin baz, cos baz) ->
    [npairs, 9, n] (RR RT RZ TR TT TZ ZR ZT ZZ).  out may be enz."""
    _require(enz, torch.float32, "enz")
    _require(trig, torch.float32, "trig")
    if enz.ndim != 3 or enz.shape[1] != 9 or enz.shape[0] < 1:
        raise ValueError("enz must be [npairs, 9, n]")
    npairs, _, n = enz.shape
    if tuple(trig.shape) != (npairs, 4):
        raise ValueError("trig must be [npairs, 4]")
    if out is None:
        out = torch.empty_like(enz)
    _require(out, torch.float32, "out")
    if tuple(out.shape) != tuple(enz.shape):
        raise ValueError("out must be [npairs, 9, n]")
    check(lib().fxc_rotate_rtz(enz.data_ptr(), trig.data_ptr(), npairs, n, out.data_ptr(),
                               _stream_ptr(enz.device)), "fxc_rotate_rtz")
    return out


class StreamEngine:
    """Serving loop with device-resident per-pair stacks (fxc_stream_*).

    cfg: the S2sPlan arguments of one segment (npts = one segment's samples);
    channels are ids 0..nchan-1; every channel pair (a < b) of a segment is
    correlated and added to that pair's float64 linear accumulator (and, with
    history > 0, its last `history` correlations for PWS / TF-PWS)."""

    KINDS = {"linear": 0, "pws": 1, "tfpws": 2}

    def __init__(self, npts, seg, shift, nfft_filter, nfft_out, num_ch, delta, freq_low, idx,
                 responses, freq_lows, nchan, half_cc, precision="bf16x3", whiten=2,
                 normalize=3, phase_only=False, history=0, max_pairs=1 << 20, max_seg_chan=512,
                 lanes=2, pws_substack=0):
        resp = np.ascontiguousarray(responses, dtype=np.float32)
        lows = np.ascontiguousarray(freq_lows, dtype=np.float32)
        self._keep = (resp, lows)
        cfg = S2sConfig(npts, seg, shift, nfft_filter, nfft_out, num_ch, float(delta),
                        float(freq_low), (ctypes.c_int * 4)(*[int(v) for v in idx]), int(whiten),
                        int(normalize), int(bool(phase_only)), resp.shape[0], resp.ctypes.data,
                        lows.ctypes.data)
        self._h = ctypes.c_void_p()
        check(lib().fxc_stream_create(ctypes.byref(self._h), ctypes.byref(cfg), nchan, half_cc,
                                      PRECISIONS[precision], history, max_pairs, max_seg_chan,
                                      lanes), "fxc_stream_create")
        if pws_substack:
            check(lib().fxc_stream_set_pws(self._h, int(pws_substack)), "fxc_stream_set_pws")
        self.npts, self.nchan = npts, nchan
        self.ncc = 2 * min(half_cc, nfft_out // 2 - 1) + 1
        self._inflight = {}

    def submit(self, chan, traces) -> int:
        """chan: channel ids (int); traces: float32 [len(chan), npts] host array,
        kept alive here until the ticket is waited for."""
        ch = np.ascontiguousarray(chan, dtype=np.int32)
        tr = np.ascontiguousarray(traces, dtype=np.float32)
        if tr.shape != (len(ch), self.npts):
            raise ValueError("traces must be [len(chan), npts]")
        t = ctypes.c_longlong()
        check(lib().fxc_stream_submit(self._h, ch.ctypes.data, len(ch), tr.ctypes.data,
                                      ctypes.byref(t)), "fxc_stream_submit")
        self._inflight[t.value] = tr
        return t.value

    def wait(self, ticket) -> float:
        """Blocks until the segment's update is applied; returns its latency (ms)."""
        ms = ctypes.c_double()
        try:
            check(lib().fxc_stream_wait(self._h, ticket, ctypes.byref(ms)), "fxc_stream_wait")
        finally:
            self._inflight.pop(ticket, None)
        return ms.value

    def info(self):
        a, b, c, d = (ctypes.c_longlong() for _ in range(4))
        check(lib().fxc_stream_info(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                    ctypes.byref(d)), "fxc_stream_info")
        return {"pairs_used": a.value, "segments_done": b.value, "state_bytes": c.value,
                "lane_bytes": d.value}

    def read(self, src, rec, kind="linear", substack=1, band=None, device="cuda:0", power=1.0,
             nvoices=0):
        """Stacks of pairs (src[i] < rec[i]) after every submitted segment:
        float32 [npairs, 2H+1] device tensor and the record counts.  PWS comes
        from the running state when the engine keeps one (pws_substack),
        else from the history; power / nvoices select the spec mode."""
        s = np.ascontiguousarray(src, dtype=np.int32)
        r = np.ascontiguousarray(rec, dtype=np.int32)
        out = torch.empty((len(s), self.ncc), dtype=torch.float32, device=device)
        cnt = np.zeros(len(s), dtype=np.int64)
        bl = band is not None
        fmin, fmax, taper = band if bl else (0.0, 0.0, 0.0)
        args = (self._h, self.KINDS[kind], s.ctypes.data, r.ctypes.data, len(s), int(substack),
                int(bl), float(fmin), float(fmax), float(taper))
        if power == 1.0 and nvoices == 0:
            check(lib().fxc_stream_read(*args, out.data_ptr(), cnt.ctypes.data,
                                        _stream_ptr(out.device)), "fxc_stream_read")
        else:
            spec = StackSpec(float(power), int(nvoices))
            check(lib().fxc_stream_read_spec(*args, ctypes.byref(spec), out.data_ptr(),
                                             cnt.ctypes.data, _stream_ptr(out.device)),
                  "fxc_stream_read_spec")
        return out, cnt

    def close(self):
        if self._h:
            lib().fxc_stream_destroy(self._h)
            self._h = ctypes.c_void_p()
        self._inflight.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
