"""SAC records as the stages store them: a 632-byte header, then float32 data,
both in the header's own byte order.

Host-side restatement of the parts of fastxc/io/sac_binary.py the stacks and
the rotation use: byte-order detection from nvhdr (6 or 7), else a plausible
npts and finite delta, else the native order (:104-128); headers kept as
bytes and only the fields a caller sets rewritten (:141-216); record reads
that check the index's size against the header (:234-253); encoding in the
header's byte order (:262-281).
"""
from __future__ import annotations

import struct
import sys
from pathlib import Path
from typing import BinaryIO

import numpy as np

SAC_HEADER_BYTES = 632
ITIME, IXY = 1, 4
SAC_NULL = -12345.0
_I_BASE = 70 * 4

# word offsets of the fields used here (70 floats, then 40 ints, then text)
FLOAT_WORD = {"delta": 0, "depmin": 1, "depmax": 2, "user0": 40, "stla": 31, "stlo": 32,
              "evla": 35, "evlo": 36, "depmen": 56, "dist": 50, "az": 51, "baz": 52,
              "gcarc": 53, "cmpaz": 57, "cmpinc": 58}
INT_WORD = {"nzyear": 0, "nzjday": 1, "nzhour": 2, "nzmin": 3, "nzsec": 4, "nzmsec": 5,
            "nvhdr": 6, "npts": 9, "iftype": 15}
TEXT_FIELD = {"kstnm": (440, 8), "kevnm": (448, 16), "kcmpnm": (600, 8), "knetwk": (608, 8)}


def byte_order(raw: bytes) -> str:
    """'<' or '>' from nvhdr (6 or 7), else from a plausible npts and delta."""
    native = "<" if sys.byteorder == "little" else ">"
    off_v = _I_BASE + 4 * INT_WORD["nvhdr"]
    if struct.unpack_from("<i", raw, off_v)[0] in (6, 7):
        return "<"
    if struct.unpack_from(">i", raw, off_v)[0] in (6, 7):
        return ">"
    ok = {}
    for e in "<>":
        n = struct.unpack_from(e + "i", raw, _I_BASE + 4 * INT_WORD["npts"])[0]
        d = struct.unpack_from(e + "f", raw, 0)[0]
        ok[e] = 0 < n < 1_000_000_000 and np.isfinite(d)
    if ok["<"] != ok[">"]:
        return "<" if ok["<"] else ">"
    return native


class SacHeader:
    """A SAC header kept as bytes in its own byte order."""

    def __init__(self, raw: bytes, endian: str | None = None):
        if len(raw) != SAC_HEADER_BYTES:
            raise ValueError(f"SAC header must be {SAC_HEADER_BYTES} bytes, got {len(raw)}")
        self.raw = bytearray(raw)
        self.endian = endian or byte_order(self.raw)

    def copy(self) -> "SacHeader":
        return SacHeader(bytes(self.raw), self.endian)

    def f(self, word) -> float:
        w = FLOAT_WORD[word] if isinstance(word, str) else word
        return struct.unpack_from(self.endian + "f", self.raw, 4 * w)[0]

    def set_f(self, word, v: float) -> None:
        w = FLOAT_WORD[word] if isinstance(word, str) else word
        struct.pack_into(self.endian + "f", self.raw, 4 * w, float(v))

    def i(self, k) -> int:
        k = INT_WORD[k] if isinstance(k, str) else k
        return struct.unpack_from(self.endian + "i", self.raw, _I_BASE + 4 * k)[0]

    def set_i(self, k, v: int) -> None:
        k = INT_WORD[k] if isinstance(k, str) else k
        struct.pack_into(self.endian + "i", self.raw, _I_BASE + 4 * k, int(v))

    def text(self, field: str) -> str:
        off, size = TEXT_FIELD[field]
        return bytes(self.raw[off:off + size]).decode("ascii", errors="ignore").strip()

    def set_text(self, field: str, value: str) -> None:
        """ASCII (unknown characters as '?'), cut to the field, blank padded."""
        off, size = TEXT_FIELD[field]
        self.raw[off:off + size] = value.encode("ascii", errors="replace")[:size].ljust(size, b" ")

    @property
    def npts(self) -> int:
        return self.i("npts")

    @property
    def data_count(self) -> int:
        n = self.npts
        if n <= 0:
            raise ValueError(f"Invalid SAC npts: {n}")
        return 2 * n if self.i("iftype") == IXY else n

    @property
    def dtype(self) -> np.dtype:
        return np.dtype(self.endian + "f4")


def _read_exact(fh: BinaryIO, size: int, label: str) -> bytes:
    raw = fh.read(size)
    if len(raw) != size:
        raise EOFError(f"Expected {size} bytes for {label}, got {len(raw)}")
    return raw


def read_record(fh: BinaryIO, path, offset: int = 0, size: int | None = None):
    """(header, float32 data) of the record at offset; size, when given, must
    equal the header's record size."""
    if offset < 0:
        raise ValueError(f"{path}: SAC record offset must be >= 0, got {offset}")
    fh.seek(offset)
    hd = SacHeader(_read_exact(fh, SAC_HEADER_BYTES, "SAC record header"))
    full = SAC_HEADER_BYTES + 4 * hd.data_count
    if size is not None and int(size) != full:
        raise ValueError(f"{path}: expected SAC record size {size}, header implies {full}")
    body = _read_exact(fh, 4 * hd.data_count, "SAC record data")
    return hd, np.frombuffer(body, dtype=hd.dtype).astype(np.float32)


def read_sac(path):
    with Path(path).open("rb") as fh:
        return read_record(fh, path)


def encode_record(hd: SacHeader, data: np.ndarray) -> bytes:
    arr = np.asarray(data, dtype=np.float32)
    if arr.size != hd.data_count:
        raise ValueError(f"SAC header expects {hd.data_count} float samples, got {arr.size}")
    return bytes(hd.raw) + arr.astype(hd.dtype, copy=False).tobytes(order="C")


def write_sac(path, hd: SacHeader, data: np.ndarray) -> None:
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    blob = encode_record(hd, data)
    with path.open("wb") as fh:
        fh.write(blob)
