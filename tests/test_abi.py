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
