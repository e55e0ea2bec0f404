"""ctypes bindings of the oracle's C pieces -- test and baseline infrastructure only.

  libcpu_xc.so         oracle/cpu_xc.c (multi-threaded FP32 xc restatement)
  libcpu_stages.so     oracle/cpu_stages.c (multi-threaded FP32 sac2spec, PWS, TF-PWS,
                       linear stack)
  _ref/libref_s2s.so   the reference's fft_length.c and filter_response.c
  _ref/libref_xc.so    the reference's io/path_table.c
The _ref libraries are compiled from /root/reference by `make -C oracle ref`
in the build container; they are absent where the reference is absent.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
CPU_XC = os.path.join(HERE, "libcpu_xc.so")
CPU_STAGES = os.path.join(HERE, "libcpu_stages.so")
REF_S2S = os.path.join(HERE, "_ref", "libref_s2s.so")
REF_XC = os.path.join(HERE, "_ref", "libref_xc.so")


def have_cpu_xc() -> bool:
    return os.path.exists(CPU_XC)


def have_cpu_stages() -> bool:
    return os.path.exists(CPU_STAGES)


def have_ref() -> bool:
    return os.path.exists(REF_S2S) and os.path.exists(REF_XC)


def cpu_xc_correlate(spec: np.ndarray, src, rec, half_cc: int, dt: float,
                     nthreads: int = 1) -> np.ndarray:
    lib = ctypes.CDLL(CPU_XC)
    spec = np.ascontiguousarray(spec, dtype=np.complex64)
    nstep, nsta, nspec = spec.shape
    src = np.ascontiguousarray(src, dtype=np.int32)
    rec = np.ascontiguousarray(rec, dtype=np.int32)
    out = np.zeros((len(src), 2 * half_cc + 1), dtype=np.float32)
    fn = lib.cpu_xc_correlate
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                   ctypes.c_void_p, ctypes.c_long, ctypes.c_int, ctypes.c_float, ctypes.c_void_p,
                   ctypes.c_int]
    rc = fn(spec.ctypes.data, nstep, nsta, nspec, src.ctypes.data, rec.ctypes.data, len(src),
            half_cc, float(dt), out.ctypes.data, int(nthreads))
    if rc != 0:
        raise RuntimeError("cpu_xc_correlate failed")
    return out


def _vp(a):
    return a.ctypes.data


def cpu_s2s_process(traces: np.ndarray, steps, plan, whiten: int = 2, normalize: int = 3,
                    nch: int = 1, nthreads: int = 1) -> np.ndarray:
    """traces [files, npts] (a group's channels adjacent), steps: window indices,
    plan: oracle.sac2spec.Plan -> spectra [nsteps, files, nspec] complex64."""
    lib = ctypes.CDLL(CPU_STAGES)
    tr = np.ascontiguousarray(traces, dtype=np.float32)
    st = np.ascontiguousarray(steps, dtype=np.int32)
    idx = np.ascontiguousarray(plan.idx, dtype=np.int32)
    resp = np.ascontiguousarray(plan.responses[0][: plan.nfft_filter // 2 + 1], dtype=np.float32)
    out = np.zeros((len(st), tr.shape[0], plan.nspec), dtype=np.complex64)
    fn = lib.cpu_s2s_process
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                   ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                   ctypes.c_float, ctypes.c_float, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    rc = fn(_vp(tr), tr.shape[0], tr.shape[1], _vp(st), len(st), plan.seg, plan.shift,
            plan.nfft_filter, plan.nfft_out, int(nch), float(plan.delta), float(plan.freq_low),
            _vp(idx), int(whiten), int(normalize), _vp(resp), _vp(out), int(nthreads))
    if rc != 0:
        raise RuntimeError("cpu_s2s_process failed (unsupported mode or length)")
    return out


def _stack_call(name, pre, w, off, lin, nthreads):
    lib = ctypes.CDLL(CPU_STAGES)
    pre = np.ascontiguousarray(pre, dtype=np.float32)
    w = np.ascontiguousarray(w, dtype=np.float32)
    off = np.ascontiguousarray(off, dtype=np.int64)
    lin = np.ascontiguousarray(lin, dtype=np.float32)
    npairs, n = lin.shape
    out = np.zeros((npairs, n), dtype=np.float32)
    fn = getattr(lib, name)
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                   ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
    if fn(_vp(pre), _vp(w), _vp(off), _vp(lin), npairs, n, _vp(out), int(nthreads)) != 0:
        raise RuntimeError(f"{name} failed (unsupported length {n})")
    return out


def cpu_pws(pre, w, off, lin, nthreads: int = 1) -> np.ndarray:
    """pre [G, n] group means, w [G], off [npairs+1] group offsets, lin [npairs, n]."""
    return _stack_call("cpu_pws", pre, w, off, lin, nthreads)


def cpu_tfpws(pre, w, off, lin, nthreads: int = 1) -> np.ndarray:
    """pre [G, n] group sums, w [G], off [npairs+1] group offsets, lin [npairs, n]."""
    return _stack_call("cpu_tfpws", pre, w, off, lin, nthreads)


def cpu_linear_stack(rec, off, nthreads: int = 1) -> np.ndarray:
    """rec [R, n] rows of every pair, off [npairs+1] row offsets -> means [npairs, n]."""
    lib = ctypes.CDLL(CPU_STAGES)
    rec = np.ascontiguousarray(rec, dtype=np.float32)
    off = np.ascontiguousarray(off, dtype=np.int64)
    npairs, n = len(off) - 1, rec.shape[1]
    out = np.zeros((npairs, n), dtype=np.float32)
    fn = lib.cpu_linear_stack
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                   ctypes.c_void_p, ctypes.c_int]
    fn(_vp(rec), _vp(off), npairs, n, _vp(out), int(nthreads))
    return out


class _Butterworth(ctypes.Structure):
    _fields_ = [("freq_low", ctypes.c_float), ("freq_high", ctypes.c_float),
                ("b", ctypes.c_double * 5), ("a", ctypes.c_double * 5)]


class _FilterResp(ctypes.Structure):
    _fields_ = [("freq_low", ctypes.c_float), ("response", ctypes.POINTER(ctypes.c_float))]


def ref_s2s():
    lib = ctypes.CDLL(REF_S2S)
    lib.segment_length_from_seconds.restype = ctypes.c_int
    lib.segment_length_from_seconds.argtypes = [ctypes.c_float, ctypes.c_int, ctypes.c_float]
    lib.next_even_smooth_2357_length.restype = ctypes.c_int
    lib.next_even_smooth_2357_length.argtypes = [ctypes.c_int]
    lib.readButterworthFilters.restype = ctypes.POINTER(_Butterworth)
    lib.readButterworthFilters.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]
    lib.processButterworthFilters.restype = ctypes.POINTER(_FilterResp)
    lib.processButterworthFilters.argtypes = [ctypes.POINTER(_Butterworth), ctypes.c_int,
                                              ctypes.c_float, ctypes.c_int]
    lib.estimateButterworthFilterPadding.restype = ctypes.c_int
    lib.estimateButterworthFilterPadding.argtypes = [ctypes.POINTER(_Butterworth), ctypes.c_int,
                                                     ctypes.c_int]
    return lib


def ref_filters(path: str, nfft: int, max_padding: int):
    """(count, padding, responses[count][nfft/2+1]) from the reference C."""
    lib = ref_s2s()
    cnt = ctypes.c_int(0)
    f = lib.readButterworthFilters(path.encode(), ctypes.byref(cnt))
    pad = lib.estimateButterworthFilterPadding(f, cnt.value, max_padding)
    resp = lib.processButterworthFilters(f, cnt.value, ctypes.c_float(1.0), nfft)
    out = np.array([[resp[i].response[k] for k in range(nfft // 2 + 1)]
                    for i in range(cnt.value)], dtype=np.float32)
    return cnt.value, pad, out


class _PathRec(ctypes.Structure):
    _fields_ = [("path_id", ctypes.c_int), ("src_gnsl_id", ctypes.c_int),
                ("rec_gnsl_id", ctypes.c_int), ("great_circle_deg", ctypes.c_float),
                ("distance_km", ctypes.c_float), ("azimuth_deg", ctypes.c_float),
                ("back_azimuth_deg", ctypes.c_float)]


class _PathTable(ctypes.Structure):
    _fields_ = [("records", ctypes.POINTER(_PathRec)), ("count", ctypes.c_size_t)]


def ref_path_table(path: str):
    """Loads allowed_paths.tsv with the reference loader; returns a lookup function."""
    lib = ctypes.CDLL(REF_XC)
    lib.load_allowed_path_table.restype = ctypes.c_int
    lib.load_allowed_path_table.argtypes = [ctypes.c_char_p, ctypes.POINTER(_PathTable)]
    lib.find_allowed_path_canonical.restype = ctypes.POINTER(_PathRec)
    lib.find_allowed_path_canonical.argtypes = [ctypes.POINTER(_PathTable), ctypes.c_int,
                                                ctypes.c_int]
    tab = _PathTable()
    if lib.load_allowed_path_table(path.encode(), ctypes.byref(tab)) != 0:
        raise RuntimeError("reference path table load failed")

    def find(a: int, b: int):
        r = lib.find_allowed_path_canonical(ctypes.byref(tab), a, b)
        if not r:
            return None
        v = r.contents
        return (v.path_id, v.src_gnsl_id, v.rec_gnsl_id, v.great_circle_deg, v.distance_km,
                v.azimuth_deg, v.back_azimuth_deg)

    return find, tab.count
