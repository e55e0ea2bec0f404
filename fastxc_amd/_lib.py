"""ctypes binding of libfastxc_amd.so (the C ABI declared in include/fastxc_amd.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``make -C fastxc_amd/csrc``).  There is no CPU fallback: if the shared library
is missing, every call raises.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_longlong, c_void_p

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfastxc_amd.so")

FXC_PREC_FP32 = 0
FXC_PREC_BF16X3 = 1
FXC_PREC_FP16 = 2
FXC_PREC_PLAIN = 3
FXC_PREC_FP16X3 = 4
PRECISIONS = {"fp32": FXC_PREC_FP32, "bf16x3": FXC_PREC_BF16X3, "fp16": FXC_PREC_FP16,
              "plain": FXC_PREC_PLAIN, "fp16x3": FXC_PREC_FP16X3}

_lib = None


class S2sConfig(ctypes.Structure):
    """fxc_s2s_config (include/fastxc_amd.h)."""
    _fields_ = [("npts", c_int), ("seg", c_int), ("shift", c_int), ("nfft_filter", c_int),
                ("nfft_out", c_int), ("num_ch", c_int), ("delta", c_float),
                ("freq_low", c_float), ("idx", c_int * 4), ("whiten", c_int),
                ("normalize", c_int), ("phase_only", c_int), ("filter_count", c_int),
                ("responses", c_void_p), ("freq_lows", c_void_p)]


class StackSpec(ctypes.Structure):
    """fxc_stack_spec: |W| exponent and TF-PWS voice count (0 = every voice)."""
    _fields_ = [("power", c_float), ("nvoices", c_int)]


class FxcError(RuntimeError):
    pass


# name -> (restype, argtypes)
_SIGNATURES = {
    "fxc_last_error": (c_char_p, []),
    "fxc_release_plans": (None, []),
    "fxc_fft_c2c": (c_int, [c_void_p, c_longlong, c_void_p, c_longlong, c_int, c_longlong, c_int,
                            c_float, c_void_p]),
    "fxc_fft_r2c": (c_int, [c_void_p, c_longlong, c_int, c_void_p, c_longlong, c_int, c_longlong,
                            c_float, c_void_p]),
    "fxc_fft_c2r": (c_int, [c_void_p, c_longlong, c_void_p, c_longlong, c_int, c_int, c_longlong,
                            c_float, c_void_p]),
    "fxc_xc_plan_create": (c_int, [POINTER(c_void_p), c_int, c_int, c_float, c_int, c_int]),
    "fxc_xc_plan_destroy": (c_int, [c_void_p]),
    "fxc_xc_plan_info": (c_int, [c_void_p, POINTER(c_int), POINTER(c_int), POINTER(c_int)]),
    "fxc_xc_set_chunk_rows": (c_int, [c_void_p, c_longlong]),
    "fxc_xc_set_pairs": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_longlong]),
    "fxc_xc_pairs_info": (c_int, [c_void_p, POINTER(c_longlong), POINTER(c_longlong),
                                  POINTER(c_longlong)]),
    "fxc_xc_cross_spectra": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "fxc_xc_lags": (c_int, [c_void_p, c_void_p, c_longlong, c_void_p, c_void_p]),
    "fxc_xc_correlate": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "fxc_xc_batch_create": (c_int, [POINTER(c_void_p), c_int, c_float, c_int, c_int, c_int,
                                    c_void_p, c_void_p, c_int, c_longlong]),
    "fxc_xc_batch_destroy": (c_int, [c_void_p]),
    "fxc_xc_batch_info": (c_int, [c_void_p, POINTER(c_longlong), POINTER(c_longlong),
                                  POINTER(c_longlong), POINTER(c_int), POINTER(c_int)]),
    "fxc_xc_batch_set_grouping": (c_int, [c_void_p, c_int]),
    "fxc_xc_batch_set_rows": (c_int, [c_void_p, c_longlong]),
    "fxc_xc_batch_correlate": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "fxc_xc_batch_tune": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p,
                                  POINTER(ctypes.c_double)]),
    "fxc_xc_batch_capture": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "fxc_xc_batch_replay": (c_int, [c_void_p, c_void_p]),
    "fxc_xc_set_timing": (c_int, [c_void_p, c_int]),
    "fxc_xc_set_band": (c_int, [c_void_p, c_int, c_int]),
    "fxc_pws": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "fxc_tfpws": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_float,
                          c_float, c_float, c_float, c_void_p, c_void_p]),
    "fxc_pws_spec": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p,
                             c_void_p, c_void_p]),
    "fxc_tfpws_spec": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int,
                               c_float, c_float, c_float, c_float, c_void_p, c_void_p, c_void_p]),
    "fxc_stack_prepare": (c_int, [c_int]),
    "fxc_pws_noalloc": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p,
                                c_void_p]),
    "fxc_linear_stack": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "fxc_rotate_rtz": (c_int, [c_void_p, c_void_p, c_longlong, c_int, c_void_p, c_void_p]),
    "fxc_plan_paths": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int,
                               c_int, c_double, c_double, c_int, c_void_p, c_int, c_int,
                               POINTER(c_void_p), POINTER(c_longlong), POINTER(c_longlong)]),
    "fxc_plan_paths_fetch": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "fxc_plan_paths_write": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_char_p, c_char_p,
                                     c_int]),
    "fxc_plan_paths_free": (c_int, [c_void_p]),
    "fxc_s2s_plan_create": (c_int, [POINTER(c_void_p), c_void_p]),
    "fxc_s2s_plan_destroy": (c_int, [c_void_p]),
    "fxc_s2s_set_sums": (c_int, [c_void_p, c_int]),
    "fxc_s2s_plan_info": (c_int, [c_void_p, POINTER(c_int), POINTER(c_int), POINTER(c_int),
                                  POINTER(c_int)]),
    "fxc_s2s_process": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p]),
    "fxc_xc_get_timing": (c_int, [c_void_p, POINTER(ctypes.c_double), POINTER(c_longlong)]),
    "fxc_stream_create": (c_int, [POINTER(c_void_p), c_void_p, c_int, c_int, c_int, c_int,
                                  c_longlong, c_int, c_int]),
    "fxc_stream_submit": (c_int, [c_void_p, c_void_p, c_int, c_void_p, POINTER(c_longlong)]),
    "fxc_stream_wait": (c_int, [c_void_p, c_longlong, POINTER(ctypes.c_double)]),
    "fxc_stream_info": (c_int, [c_void_p, POINTER(c_longlong), POINTER(c_longlong),
                                POINTER(c_longlong), POINTER(c_longlong)]),
    "fxc_stream_read": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_longlong, c_int, c_int,
                                c_float, c_float, c_float, c_void_p, c_void_p, c_void_p]),
    "fxc_stream_set_pws": (c_int, [c_void_p, c_int]),
    "fxc_stream_read_spec": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_longlong, c_int, c_int,
                                     c_float, c_float, c_float, c_void_p, c_void_p, c_void_p,
                                     c_void_p]),
    "fxc_stream_destroy": (c_int, [c_void_p]),
}


def exported_symbols():
    """Names the C ABI exports (checked against include/fastxc_amd.h by tests)."""
    return sorted(_SIGNATURES)


def lib():
    """Loads the shared library once; raises FxcError when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FxcError(f"{LIB_PATH} not built; run __graft_entry__.build() first")
    handle = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(handle, name)
        fn.restype = res
        fn.argtypes = args
    _lib = handle
    return _lib


def check(status: int, what: str = "") -> None:
    if status != 0:
        msg = lib().fxc_last_error()
        raise FxcError(f"{what} failed ({status}): {msg.decode() if msg else ''}")
