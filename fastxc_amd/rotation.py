"""ENZ -> RTZ rotation of stacked correlations, the arithmetic on the GPU.

Host-side mirror of fastxc/operators/rotation/python_rotate.py (the reference
runs it in numpy): the same entry points ``rotate_sourcepack_stack`` /
``rotate_linearstack_sourcepack`` (:264-390), ``rotate_sac_files`` /
``rotate_from_lists`` / ``rotate_many_from_lists`` /
``discover_rotate_list_pairs`` (:128-262), ``build_rotate_matrix`` and
``rotate_enz_to_rtz`` (:34-104), the same ``RotateResult`` records, files
and errors.  The 9x9 map of every pair runs in the ``fxc_rotate_rtz`` kernel
(include/fastxc_amd.h), many pairs per call; file handling stays on the host.

Behaviour kept from the reference:
  - a stack index is grouped by (path_id, source_key, receiver_key); groups
    run in that key order; a group needs all nine tags E-E ... Z-Z (the
    component names as written, the last row of a tag wins) or it is
    skipped and counted (:277-318, :436-470);
  - every record must have the Z-Z record's npts, and a delta within 1e-7
    of it (:325-333);
  - the geometry comes from the Z-Z header: Rudoe from evlo/evla/stlo/stla
    when all four are set, else its az/baz (and gcarc/dist) (:109-125);
  - outputs keep the Z-Z header with gcarc/az/baz/dist, cmpaz = baz -+ 180,
    cmpinc = 90 and kcmpnm RR ... ZZ (:336-346); they go to
    ``<out>/stack/rtz_<method>_sourcepack/STACK.tmp``, which replaces
    ``STACK`` when complete: ``rtz_<method>.pack``, ``sourcepack_index.tsv``
    (csv.DictWriter rows, timestamp "ROTATE", components R/T/Z, slot with
    R ranked as E and T as N, "%.9g" numbers, storage_kind
    "rtz_<method>_pack") and ``_SUCCESS`` (:348-390, :473-520).
There is no CPU fallback: without the built library the calls raise FxcError.
"""
from __future__ import annotations

import csv
import shutil
from dataclasses import dataclass
from pathlib import Path
from typing import BinaryIO, Iterable, Sequence

import numpy as np

from .geodesy import distkm_az_baz_rudoe
from .sacrec import SAC_NULL, SacHeader, encode_record, read_record, read_sac, write_sac

ENZ_ORDER = ("E-E", "E-N", "E-Z", "N-E", "N-N", "N-Z", "Z-E", "Z-N", "Z-Z")
RTZ_ORDER = ("R-R", "R-T", "R-Z", "T-R", "T-T", "T-Z", "Z-R", "Z-T", "Z-Z")
_KCMPNM = tuple(t.replace("-", "") for t in RTZ_ORDER)

SOURCEPACK_INDEX_HEADER = [
    "timestamp", "path_id", "component_slot", "source_key", "receiver_key", "src_id", "rec_id",
    "src_network", "src_station", "src_location", "src_component", "rec_network", "rec_station",
    "rec_location", "rec_component", "npts", "dt", "dist", "az", "baz", "record_path",
    "record_offset", "bytes", "storage_kind", "final_pair_path"]
_REQUIRED = {"source_key", "receiver_key", "record_path", "record_offset", "bytes",
             "src_component", "rec_component"}
_COPIED = ("path_id", "source_key", "receiver_key", "src_id", "rec_id", "src_network",
           "src_station", "src_location", "rec_network", "rec_station", "rec_location")

# pairs staged per kernel call are bounded by bytes (input + output)
_BATCH_BYTES = 256 << 20


@dataclass(frozen=True)
class RotateResult:
    input_list: Path
    output_list: Path
    output_paths: tuple
    sample_count: int


def _trig(azimuth_deg: float, back_azimuth_deg: float) -> np.ndarray:
    """float32 (sin az, cos az, sin baz, cos baz) of the float32 radian angles."""
    az = np.float32(np.deg2rad(azimuth_deg))
    baz = np.float32(np.deg2rad(back_azimuth_deg))
    return np.array([np.sin(az, dtype=np.float32), np.cos(az, dtype=np.float32),
                     np.sin(baz, dtype=np.float32), np.cos(baz, dtype=np.float32)], np.float32)


def build_rotate_matrix(azimuth_deg: float, back_azimuth_deg: float) -> np.ndarray:
    """9x9 float32 map of the ENZ tensor (rows of ENZ_ORDER) to RTZ (RTZ_ORDER):
    source basis R = E sin az + N cos az, T = E cos az - N sin az; receiver
    basis R = -E sin baz - N cos baz, T = -E cos baz + N sin baz; Z unchanged.
    Every entry is one float32 product, the values the kernel forms."""
    sa, ca, sb, cb = (np.float32(v) for v in _trig(azimuth_deg, back_azimuth_deg))
    src = {"R": (sa, ca), "T": (ca, -sa)}      # coefficients of (E, N)
    rec = {"R": (-sb, -cb), "T": (-cb, sb)}
    m = np.zeros((9, 9), np.float32)
    col = {c: k for k, c in enumerate(ENZ_ORDER)}
    for r, tag in enumerate(RTZ_ORDER):
        x, y = tag.split("-")
        for i, ci in enumerate("EN"):
            for j, cj in enumerate("EN"):
                if x != "Z" and y != "Z":
                    m[r, col[f"{ci}-{cj}"]] = np.float32(src[x][i] * rec[y][j])
            if x != "Z" and y == "Z":
                m[r, col[f"{ci}-Z"]] = src[x][i]
            if x == "Z" and y != "Z":
                m[r, col[f"Z-{ci}"]] = rec[y][i]
    m[8, 8] = 1.0
    return m


def _rotate_batch(enz: np.ndarray, trig: np.ndarray, device=None) -> np.ndarray:
    """enz [P, 9, n] float32, trig [P, 4] -> [P, 9, n] through fxc_rotate_rtz."""
    import torch

    from . import ops

    dev = torch.device(device if device is not None else "cuda:0")
    x = torch.from_numpy(np.ascontiguousarray(enz, dtype=np.float32)).to(dev)
    t = torch.from_numpy(np.ascontiguousarray(trig, dtype=np.float32)).to(dev)
    ops.rotate_rtz(x, t, out=x)
    return x.cpu().numpy()


def rotate_enz_to_rtz(enz_data: np.ndarray, azimuth_deg: float, back_azimuth_deg: float, *,
                      device=None) -> np.ndarray:
    enz = np.asarray(enz_data, dtype=np.float32)
    if enz.ndim != 2 or enz.shape[0] != 9:
        raise ValueError(f"ENZ data must have shape (9, npts), got {enz.shape}")
    return _rotate_batch(enz[None], _trig(azimuth_deg, back_azimuth_deg)[None], device)[0]


def _is_defined(value: float) -> bool:
    return bool(np.isfinite(value)) and abs(value - SAC_NULL) > 1.0e-3


def _geometry_from_header(hd: SacHeader) -> tuple:
    ll = [hd.f(k) for k in ("evlo", "evla", "stlo", "stla")]
    if all(_is_defined(v) for v in ll):
        return distkm_az_baz_rudoe(*ll)
    gcarc, az, baz, dist = (hd.f(k) for k in ("gcarc", "az", "baz", "dist"))
    if _is_defined(az) and _is_defined(baz):
        return gcarc, az, baz, dist
    raise ValueError("Cannot determine azimuth/back-azimuth from SAC header")


def _check_shapes(labels, headers):
    template = headers[8]
    npts, delta = template.npts, template.f("delta")
    for label, hd in zip(labels, headers):
        if hd.npts != npts:
            raise ValueError(f"{label}: npts mismatch ({hd.npts} vs {npts})")
        if not np.isclose(hd.f("delta"), delta, rtol=0.0, atol=1.0e-7):
            raise ValueError(f"{label}: delta mismatch ({hd.f('delta')} vs {delta})")


def _rotated_header(template: SacHeader, geometry) -> SacHeader:
    gcarc, az, baz, dist = geometry
    hd = template.copy()
    hd.set_f("gcarc", float(gcarc))
    hd.set_f("az", float(az))
    hd.set_f("baz", float(baz))
    hd.set_f("dist", float(dist))
    hd.set_f("cmpaz", float(baz + 180.0 if baz - 180.0 < 0.0 else baz - 180.0))
    hd.set_f("cmpinc", 90.0)
    return hd


class _Pending:
    """Pairs read and checked, waiting for one kernel call."""

    def __init__(self, device):
        self.device = device
        self.items: list = []     # caller's payload per pair
        self.data: list = []      # [9, n] float32 per pair
        self.trig: list = []
        self.n = None

    def fits(self, n: int) -> bool:
        return self.n is None or (n == self.n and
                                  (len(self.items) + 1) * 9 * n * 8 <= _BATCH_BYTES)

    def add(self, item, data, trig):
        self.n = data.shape[1]
        self.items.append(item)
        self.data.append(data)
        self.trig.append(trig)

    def run(self):
        if not self.items:
            return []
        out = _rotate_batch(np.stack(self.data), np.stack(self.trig), self.device)
        done = list(zip(self.items, out))
        self.items, self.data, self.trig, self.n = [], [], [], None
        return done


def _read_pair_files(inputs):
    headers, traces = [], []
    for path in inputs:
        hd, data = read_sac(path)
        headers.append(hd)
        traces.append(data)
    _check_shapes(inputs, headers)
    template = headers[8]
    geometry = _geometry_from_header(template)
    return template, geometry, np.stack(traces, axis=0).astype(np.float32, copy=False)


def _write_pair_files(template, geometry, rtz, outputs):
    base = _rotated_header(template, geometry)
    for name, out_path, row in zip(_KCMPNM, outputs, rtz):
        hd = base.copy()
        hd.set_text("kcmpnm", name)
        write_sac(out_path, hd, row)


def rotate_sac_files(input_paths: Sequence, output_paths: Sequence, *, dry_run: bool = False,
                     device=None) -> RotateResult:
    if len(input_paths) != 9 or len(output_paths) != 9:
        raise ValueError("rotate_sac_files expects exactly 9 input and 9 output paths")
    inputs = tuple(Path(p).expanduser().resolve() for p in input_paths)
    outputs = tuple(Path(p).expanduser().resolve() for p in output_paths)
    if dry_run:
        return RotateResult(Path("<direct>"), Path("<direct>"), outputs, 0)
    template, geometry, enz = _read_pair_files(inputs)
    rtz = _rotate_batch(enz[None], _trig(geometry[1], geometry[2])[None], device)[0]
    _write_pair_files(template, geometry, rtz, outputs)
    return RotateResult(Path("<direct>"), Path("<direct>"), outputs, template.npts)


def _read_list(path, expected: int = 9) -> list:
    path = Path(path).expanduser().resolve()
    rows = [Path(line.strip()).expanduser().resolve()
            for line in path.read_text().splitlines() if line.strip()]
    if len(rows) != expected:
        raise ValueError(f"{path}: expected {expected} paths, got {len(rows)}")
    return rows


def rotate_from_lists(input_list, output_list, *, dry_run: bool = False,
                      device=None) -> RotateResult:
    return rotate_many_from_lists([(input_list, output_list)], dry_run=dry_run,
                                  device=device)[0]


def discover_rotate_list_pairs(output_dir) -> list:
    root = Path(output_dir).expanduser().resolve() / "rotate_list"
    if not root.is_dir():
        return []
    pairs = []
    for in_list in sorted(root.glob("*/*/enz_list.txt")):
        out_list = in_list.with_name("rtz_list.txt")
        if out_list.is_file():
            pairs.append((in_list, out_list))
    return pairs


def rotate_many_from_lists(list_pairs: Iterable, *, max_workers: int = 1, dry_run: bool = False,
                           device=None) -> list:
    """Every (enz_list.txt, rtz_list.txt) pair; the files are read and written
    on the host, the rotations batched into kernel calls.  Results in task
    order, or sorted by input list when max_workers > 1 (the reference's
    thread-pool order)."""
    tasks = [(Path(a).expanduser().resolve(), Path(b).expanduser().resolve())
             for a, b in list_pairs]
    if not tasks:
        return []
    results = {}
    pending = _Pending(device)

    def flush():
        for (k, in_list, out_list, template, geometry, outputs), rtz in pending.run():
            _write_pair_files(template, geometry, rtz, outputs)
            results[k] = RotateResult(in_list, out_list, outputs, template.npts)

    for k, (in_list, out_list) in enumerate(tasks):
        outputs = tuple(_read_list(out_list))
        if dry_run:
            results[k] = RotateResult(in_list, out_list, outputs, 0)
            continue
        template, geometry, enz = _read_pair_files(_read_list(in_list))
        if not pending.fits(enz.shape[1]):
            flush()
        pending.add((k, in_list, out_list, template, geometry, outputs), enz,
                    _trig(geometry[1], geometry[2]))
    flush()
    ordered = [results[k] for k in range(len(tasks))]
    if max(1, int(max_workers)) > 1 and len(tasks) > 1:
        ordered.sort(key=lambda r: str(r.input_list))
    return ordered


def _row_path_id(row: dict) -> str:
    pid = (row.get("path_id") or "").strip()
    if pid:
        return pid.zfill(8) if pid.isdigit() else pid
    try:
        return "%08d" % (int(row["src_id"]) * 10000 + int(row["rec_id"]))
    except (KeyError, ValueError, TypeError):
        return ""


def _slot(src_component: str, rec_component: str) -> int:
    def rank(c: str) -> int:
        return {"E": 0, "1": 0, "R": 0, "N": 1, "2": 1, "T": 1, "Z": 2, "3": 2}.get(
            c.upper()[-1:] if c else "", 100)
    return rank(src_component) * 3 + rank(rec_component)


def _read_groups(index_path: Path) -> dict:
    groups: dict = {}
    with index_path.open("r", encoding="utf-8-sig", newline="") as fh:
        reader = csv.DictReader(fh, delimiter="\t")
        if reader.fieldnames is None:
            return groups
        missing = _REQUIRED.difference(reader.fieldnames)
        if missing:
            raise ValueError(f"Sourcepack index missing fields {sorted(missing)}: {index_path}")
        for row in reader:
            groups.setdefault((_row_path_id(row), row["source_key"], row["receiver_key"]),
                              []).append(row)
    return groups


def _legacy_rtz_path(output_dir: Path, legacy: str, row: dict, tag: str) -> Path:
    pair = f"{row['src_network']}-{row['rec_network']}.{row['src_station']}-{row['rec_station']}"
    return output_dir / "stack" / legacy / pair / f"{pair}.{tag}.sac"


class _Records:
    """Record files opened once for the whole run."""

    def __init__(self):
        self.handles: dict = {}

    def read(self, row: dict):
        path = Path(row["record_path"]).expanduser()
        fh = self.handles.get(path)
        if fh is None:
            fh = path.open("rb")
            self.handles[path] = fh
        return read_record(fh, path, int(row["record_offset"]), int(row["bytes"]))

    def close(self):
        for fh in self.handles.values():
            fh.close()
        self.handles.clear()


def rotate_sourcepack_stack(output_dir, method: str, *, dry_run: bool = False,
                            device=None) -> list:
    """Rotates every complete ENZ pair of ``<output_dir>/stack/<method>_sourcepack/
    STACK/sourcepack_index.tsv`` into ``<output_dir>/stack/rtz_<method>_sourcepack/
    STACK``; one RotateResult per rotated pair, in output order."""
    output_dir = Path(output_dir).expanduser().resolve()
    method = method.strip().lower()
    legacy = f"rtz_{method}"
    index_path = output_dir / "stack" / f"{method}_sourcepack" / "STACK" / "sourcepack_index.tsv"
    if not index_path.is_file():
        raise FileNotFoundError(f"{method} sourcepack index not found: {index_path}")
    groups = sorted(_read_groups(index_path).items(),
                    key=lambda kv: (_row_path_id(kv[1][0]), kv[0][1], kv[0][2]))
    out_dir = output_dir / "stack" / f"rtz_{method}_sourcepack" / "STACK"
    pack_path = out_dir / f"rtz_{method}.pack"
    out_index = out_dir / "sourcepack_index.tsv"
    if dry_run:
        return [RotateResult(index_path, out_index,
                             tuple(_legacy_rtz_path(output_dir, legacy, rows[0], t)
                                   for t in RTZ_ORDER), 0)
                for _, rows in groups]

    tmp_dir = out_dir.with_name(out_dir.name + ".tmp")
    if tmp_dir.exists():
        shutil.rmtree(tmp_dir)
    tmp_dir.mkdir(parents=True, exist_ok=True)
    results: list = []
    skipped = 0
    records = _Records()
    pending = _Pending(device)
    try:
        with (tmp_dir / pack_path.name).open("wb") as pack, \
                (tmp_dir / out_index.name).open("w", encoding="utf-8", newline="") as idx:
            writer = csv.DictWriter(idx, fieldnames=SOURCEPACK_INDEX_HEADER, delimiter="\t")
            writer.writeheader()

            def flush():
                for (first, template, geometry), rtz in pending.run():
                    results.append(_write_stack_pair(pack, writer, first, template, geometry,
                                                     rtz, output_dir, legacy, pack_path,
                                                     f"rtz_{method}_pack", index_path,
                                                     out_index))

            for _, rows in groups:
                by_tag = {f"{r['src_component']}-{r['rec_component']}": r for r in rows}
                if any(t not in by_tag for t in ENZ_ORDER):
                    skipped += 1
                    continue
                ordered = [by_tag[t] for t in ENZ_ORDER]
                read = [records.read(r) for r in ordered]
                headers = [hd for hd, _ in read]
                _check_shapes([r["record_path"] for r in ordered], headers)
                template = headers[8]
                geometry = _geometry_from_header(template)
                enz = np.stack([d for _, d in read], axis=0)
                if not pending.fits(enz.shape[1]):
                    flush()
                pending.add((rows[0], template, geometry), enz, _trig(geometry[1], geometry[2]))
            flush()
    finally:
        records.close()

    (tmp_dir / "_SUCCESS").write_text(
        f"pairs\t{len(results)}\nskipped\t{skipped}\npack\t{pack_path}\nindex\t{out_index}\n",
        encoding="utf-8")
    if out_dir.exists():
        shutil.rmtree(out_dir)
    tmp_dir.replace(out_dir)
    return results


def _write_stack_pair(pack: BinaryIO, writer, first: dict, template: SacHeader, geometry, rtz,
                      output_dir: Path, legacy: str, pack_path: Path, storage_kind: str,
                      index_path: Path, out_index: Path) -> RotateResult:
    gcarc, az, baz, dist = geometry
    base = _rotated_header(template, geometry)
    outputs = []
    for k, tag in enumerate(RTZ_ORDER):
        sc, rc = tag.split("-")
        hd = base.copy()
        hd.set_text("kcmpnm", _KCMPNM[k])
        offset = pack.tell()
        blob = encode_record(hd, rtz[k])
        pack.write(blob)
        final = _legacy_rtz_path(output_dir, legacy, first, tag)
        row = {f: "" for f in SOURCEPACK_INDEX_HEADER}
        for f in _COPIED:
            row[f] = first.get(f, "")
        row.update(timestamp="ROTATE", path_id=_row_path_id(first),
                   component_slot=str(_slot(sc, rc)), src_component=sc, rec_component=rc,
                   npts=str(hd.npts), dt=f"{hd.f('delta'):.9g}", dist=f"{dist:.9g}",
                   az=f"{az:.9g}", baz=f"{baz:.9g}", record_path=pack_path.as_posix(),
                   record_offset=str(offset), bytes=str(len(blob)), storage_kind=storage_kind,
                   final_pair_path=final.as_posix())
        writer.writerow(row)
        outputs.append(final)
    return RotateResult(index_path, out_index, tuple(outputs), template.npts)


def rotate_linearstack_sourcepack(output_dir, *, dry_run: bool = False, device=None) -> list:
    return rotate_sourcepack_stack(output_dir, "linearstack", dry_run=dry_run, device=device)
