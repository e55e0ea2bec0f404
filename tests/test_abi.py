"""The C ABI library loads and exports every symbol include/fastxc_amd.h declares (CPU only)."""
import ctypes
import os
import re

from fastxc_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    hdr_dir = os.path.join(ROOT, "include")
    for fn in os.listdir(hdr_dir):
        if fn.endswith(".h"):
            text = open(os.path.join(hdr_dir, fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(fxc_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_lib.LIB_PATH), "library not built"
    handle = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(_declared()) if not hasattr(handle, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(_lib.exported_symbols()) == _declared()


def test_batch_create_rejects_bad_shapes_before_any_device_call():
    """fxc_xc_batch_create checks its arguments on the host and returns the
    documented status codes (1 argument, 4 unsupported) with a message,
    before it touches a device."""
    lib = _lib.lib()
    h = ctypes.c_void_p()

    def create(nsta, nstep, precision=1, gap_mode=0, max_rows=0):
        a = (ctypes.c_int * len(nsta))(*nsta)
        b = (ctypes.c_int * len(nstep))(*nstep)
        return lib.fxc_xc_batch_create(ctypes.byref(h), 1025, ctypes.c_float(0.1), 100, precision,
                                       len(nsta), a, b, gap_mode, max_rows)

    assert create([10], [24], precision=0) == 4  # FP32: the batched form is split-BF16 only
    assert b"split-BF16" in lib.fxc_last_error()
    assert create([10, 5], [24, 129]) == 1  # more than 128 windows
    assert b"segment 1" in lib.fxc_last_error()
    assert create([10, 0], [24, 6]) == 1  # a segment without stations
    assert create([10], [24], gap_mode=2) == 1  # unknown gap mode
    assert create([10], [24], max_rows=-1) == 1
    assert create([], []) == 1  # no segments
    assert not h.value
