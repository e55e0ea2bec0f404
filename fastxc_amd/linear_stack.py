"""Cross-timestamp linear stack of SourcePack records on the GPU.

Host-side mirror of fastxc/operators/stacking/linear_stack.py (the reference
runs it in Python): the same entry point ``linear_stack_sourcepack_indexes``
(:300-384), output naming ``linear_output_path`` (:68-73) and
``linear_sourcepack_dir`` (:76-77), the same ``StackResult`` records and
errors.  The float64 sum of each group runs in the FP64-accumulator kernel
behind ``fxc_linear_stack`` (include/fastxc_amd.h); the file handling stays on
the host.

Behaviour kept from the reference:
  - indexes are merged k-way on (path_id zero-filled to 8 digits when
    numeric, component_slot or the E/N/Z rank fallback) (:158-223);
  - records of a group are read in (timestamp, record_path, record_offset)
    order (:187); every record must have the first one's sample count, and a
    delta within 1e-7 of it (:228-247);
  - the output header is the first record's with npts = sample count,
    iftype = ITIME, user0 = record count and the reference date
    2010-214 00:00:00.000 (:80-92); data keep the first header's byte order;
  - outputs go to ``<stack_dir>/linearstack_sourcepack/STACK.tmp`` and replace
    ``STACK`` when complete: ``linearstack.pack``, ``sourcepack_index.tsv``
    (csv.DictWriter rows, timestamp "STACK", dt "%.9g", storage_kind
    "linearstack_pack") and ``_SUCCESS`` (:325-384).
There is no CPU fallback: without the built library the call raises FxcError.
"""
from __future__ import annotations

import csv
import heapq
import shutil
from dataclasses import dataclass
from pathlib import Path
from typing import BinaryIO, Iterable, Iterator, Sequence

import numpy as np

from .sacrec import ITIME, SAC_HEADER_BYTES, SacHeader as _Header

SOURCEPACK_INDEX_HEADER = [
    "timestamp", "path_id", "component_slot", "source_key", "receiver_key", "src_id", "rec_id",
    "src_network", "src_station", "src_location", "src_component", "rec_network", "rec_station",
    "rec_location", "rec_component", "npts", "dt", "dist", "az", "baz", "record_path",
    "record_offset", "bytes", "storage_kind", "final_pair_path"]

_REQUIRED = ("source_key", "receiver_key", "record_path", "record_offset", "bytes",
             "src_component", "rec_component")
_TEMPLATE_FIELDS = ("path_id", "component_slot", "source_key", "receiver_key", "src_id", "rec_id",
                    "src_network", "src_station", "src_location", "src_component", "rec_network",
                    "rec_station", "rec_location", "rec_component", "dist", "az", "baz")

# word offsets in the 632-byte header (70 floats, then 40 ints)
_F_DELTA, _F_USER0 = 0, 40
_I_NZ = range(0, 6)  # nzyear nzjday nzhour nzmin nzsec nzmsec
_I_NPTS, _I_IFTYPE = 9, 15

# Rows staged per GPU call are bounded by bytes and by the kernel's pair limit.
_BATCH_BYTES = 256 << 20
_MAX_PAIRS = 65535


@dataclass(frozen=True)
class StackResult:
    input_path: Path
    output_path: Path
    segment_count: int
    sample_count: int


@dataclass(frozen=True)
class _Record:
    row: dict
    path: Path
    offset: int
    nbytes: int


def _iter_index(index: Path) -> Iterator[_Record]:
    with index.open("r", encoding="utf-8-sig", newline="") as fh:
        reader = csv.DictReader(fh, delimiter="\t")
        if reader.fieldnames is None:
            return
        missing = set(_REQUIRED).difference(reader.fieldnames)
        if missing:
            raise ValueError(f"Source index missing fields {sorted(missing)}: {index}")
        for row in reader:
            p = Path(row["record_path"]).expanduser()
            if not p.is_absolute():
                p = index.parent / p
            yield _Record(row, p, int(row["record_offset"]), int(row["bytes"]))


def _group_key(row: dict) -> tuple[str, int]:
    pid = row.get("path_id", "").strip()
    if pid:
        pid = pid.zfill(8) if pid.isdigit() else pid
    else:
        try:
            pid = "%08d" % (int(row["src_id"]) * 10000 + int(row["rec_id"]))
        except (KeyError, ValueError):
            pid = ""
    slot = row.get("component_slot", "").strip()
    if slot.isdigit():
        return pid, int(slot)

    def rank(c: str) -> int:
        return {"E": 0, "1": 0, "R": 0, "N": 1, "2": 1, "T": 1, "Z": 2, "3": 2}.get(
            c.upper()[-1:] if c else "", 100)

    return pid, rank(row.get("src_component", "")) * 3 + rank(row.get("rec_component", ""))


def _groups(indexes: Sequence[Path]) -> Iterator[list[_Record]]:
    """k-way merge of the (sorted) indexes into runs of equal group key."""
    heap: list = []
    seq = 0
    streams = [_iter_index(p) for p in indexes]
    for it in streams:
        rec = next(it, None)
        if rec is not None:
            heap.append((_group_key(rec.row), seq, rec, it))
            seq += 1
    heapq.heapify(heap)
    while heap:
        key = heap[0][0]
        run = []
        while heap and heap[0][0] == key:
            _, _, rec, it = heapq.heappop(heap)
            run.append(rec)
            nxt = next(it, None)
            if nxt is not None:
                heapq.heappush(heap, (_group_key(nxt.row), seq, nxt, it))
                seq += 1
        run.sort(key=lambda r: (r.row.get("timestamp", ""), r.path.as_posix(), r.offset))
        yield run


def _final_pair(rec: _Record) -> str:
    row = rec.row
    fp = row.get("final_pair_path")
    if fp:
        return fp
    return "%s-%s.%s-%s.%s-%s.sac" % (row["src_network"], row["rec_network"], row["src_station"],
                                      row["rec_station"], row["src_component"],
                                      row["rec_component"])


def linear_output_path(final_pair_path: str | Path, stack_dir: str | Path) -> Path:
    """<stack_dir>/linearstack/<net>.<sta>/<net>.<sta>.<cmp>.linearstack.sac"""
    parts = Path(final_pair_path).name.split(".")
    if parts and parts[-1].lower() == "sac":
        parts = parts[:-1]
    if len(parts) < 3:
        raise ValueError(f"Unexpected final pair path: {Path(final_pair_path).name}")
    pair = f"{parts[0]}.{parts[1]}"
    return (Path(stack_dir).expanduser().resolve() / "linearstack" / pair /
            f"{pair}.{parts[2]}.linearstack.sac")


def linear_sourcepack_dir(stack_dir: str | Path) -> Path:
    return Path(stack_dir).expanduser().resolve() / "linearstack_sourcepack" / "STACK"


class _Reader:
    """Open record files, kept for the whole run."""

    def __init__(self):
        self.handles: dict[Path, BinaryIO] = {}

    def read(self, rec: _Record) -> tuple[_Header, np.ndarray]:
        fh = self.handles.get(rec.path)
        if fh is None:
            fh = rec.path.open("rb")
            self.handles[rec.path] = fh
        if rec.offset < 0:
            raise ValueError(f"{rec.path}: SAC record offset must be >= 0, got {rec.offset}")
        fh.seek(rec.offset)
        raw = fh.read(SAC_HEADER_BYTES)
        if len(raw) != SAC_HEADER_BYTES:
            raise EOFError(f"Expected {SAC_HEADER_BYTES} bytes for SAC record header, got {len(raw)}")
        hd = _Header(raw)
        size = SAC_HEADER_BYTES + 4 * hd.data_count
        if rec.nbytes != size:
            raise ValueError(f"{rec.path}: expected SAC record size {rec.nbytes}, header implies {size}")
        body = fh.read(4 * hd.data_count)
        if len(body) != 4 * hd.data_count:
            raise EOFError(f"Expected {4 * hd.data_count} bytes for SAC record data, got {len(body)}")
        return hd, np.frombuffer(body, dtype=hd.dtype).astype(np.float32)

    def close(self):
        for fh in self.handles.values():
            fh.close()
        self.handles.clear()


class _Batch:
    """Groups of one sample count staged for one fxc_linear_stack call."""

    def __init__(self, device):
        self.device = device
        self.clear()

    def clear(self):
        self.groups: list = []   # (records, header)
        self.rows: list = []
        self.offsets = [0]
        self.n = None

    def fits(self, n: int, rows: int) -> bool:
        if self.n is None:
            return True
        return (n == self.n and len(self.groups) < _MAX_PAIRS and
                4 * n * (self.offsets[-1] + rows) <= _BATCH_BYTES)

    def add(self, records, header, data_rows):
        self.n = header.data_count
        self.groups.append((records, header))
        self.rows.extend(data_rows)
        self.offsets.append(self.offsets[-1] + len(data_rows))

    def run(self) -> np.ndarray:
        import torch

        from . import ops

        x = torch.from_numpy(np.stack(self.rows)).to(self.device)
        out = ops.linear_stack(x, np.asarray(self.offsets, dtype=np.int64))
        return out.cpu().numpy()


def linear_stack_sourcepack_indexes(indexes: Iterable[str | Path], stack_dir: str | Path, *,
                                    dry_run: bool = False, device=None) -> list[StackResult]:
    """Linear stack of every (path, component slot) group of the SourcePack
    indexes into ``<stack_dir>/linearstack_sourcepack/STACK``; one StackResult
    per group in output order.  ``device``: the torch device (default cuda:0)."""
    indexes = [Path(p).expanduser().resolve() for p in indexes]
    out_dir = linear_sourcepack_dir(stack_dir)
    pack_path = out_dir / "linearstack.pack"
    index_path = out_dir / "sourcepack_index.tsv"
    if dry_run:
        return [StackResult(Path(_final_pair(g[0])), pack_path, len(g), 0) for g in _groups(indexes)]

    tmp_dir = out_dir.with_name(out_dir.name + ".tmp")
    if tmp_dir.exists():
        shutil.rmtree(tmp_dir)
    tmp_dir.mkdir(parents=True, exist_ok=True)
    results: list[StackResult] = []
    reader = _Reader()
    batch = _Batch(device if device is not None else "cuda:0")
    try:
        with (tmp_dir / pack_path.name).open("wb") as pack, \
                (tmp_dir / index_path.name).open("w", encoding="utf-8", newline="") as idx:
            writer = csv.DictWriter(idx, fieldnames=SOURCEPACK_INDEX_HEADER, delimiter="\t")
            writer.writeheader()

            def flush():
                if not batch.groups:
                    return
                stacked = batch.run()
                for (records, hd), y in zip(batch.groups, stacked):
                    final_pair = _final_pair(records[0])
                    out_hd = _Header(bytes(hd.raw))
                    out_hd.set_i(_I_NPTS, hd.data_count)
                    out_hd.set_i(_I_IFTYPE, ITIME)
                    out_hd.set_f(_F_USER0, float(len(records)))
                    for k, v in zip(_I_NZ, (2010, 214, 0, 0, 0, 0)):
                        out_hd.set_i(k, v)
                    offset = pack.tell()
                    blob = bytes(out_hd.raw) + y.astype(out_hd.dtype).tobytes()
                    pack.write(blob)
                    lin_path = linear_output_path(final_pair, stack_dir)
                    row = {k: "" for k in SOURCEPACK_INDEX_HEADER}
                    for k in _TEMPLATE_FIELDS:
                        row[k] = records[0].row.get(k, "")
                    row.update(timestamp="STACK", npts=str(out_hd.npts),
                               dt=f"{out_hd.f(_F_DELTA):.9g}", record_path=pack_path.as_posix(),
                               record_offset=str(offset), bytes=str(len(blob)),
                               storage_kind="linearstack_pack",
                               final_pair_path=lin_path.as_posix())
                    writer.writerow(row)
                    results.append(StackResult(Path(final_pair), lin_path, len(records),
                                               hd.data_count))
                batch.clear()

            for records in _groups(indexes):
                first, data0 = reader.read(records[0])
                n, delta = first.data_count, first.f(_F_DELTA)
                rows = [data0]
                for rec in records[1:]:
                    hd, d = reader.read(rec)
                    if hd.data_count != n:
                        raise ValueError(f"{rec.path}: data_count mismatch ({hd.data_count} vs {n})")
                    if not np.isclose(hd.f(_F_DELTA), delta, rtol=0.0, atol=1.0e-7):
                        raise ValueError(f"{rec.path}: delta mismatch ({hd.f(_F_DELTA)} vs {delta})")
                    rows.append(d)
                if not batch.fits(n, len(rows)):
                    flush()
                batch.add(records, first, rows)
            flush()
    finally:
        reader.close()

    if not results:
        shutil.rmtree(tmp_dir, ignore_errors=True)
        return []
    (tmp_dir / "_SUCCESS").write_text(
        f"records\t{len(results)}\npack\t{pack_path}\nindex\t{index_path}\n", encoding="utf-8")
    if out_dir.exists():
        shutil.rmtree(out_dir)
    tmp_dir.replace(out_dir)
    return results
