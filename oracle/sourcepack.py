"""SourcePack index builder restated in Python -- test infrastructure only.

  read_xcpack          read_xcpack_tsv, fastxc/io/sourcepack.py:120-154
                       (required fields, pack paths relative to the TSV's
                       folder, offset >= 0, bytes > 0, sequence counted over
                       every row read)
  index_row            _index_row, fastxc/operators/sourcepack/builder.py:309-337
  sort_key             _sort_key / _component_rank / _component_slot /
                       _path_id_text, builder.py:401-429
  safe_name            _safe_name, builder.py:432-434
  build_timestamp      build_sourcepack_timestamp + _build_one_timestamp,
                       builder.py:242-306 (TSVs named after the timestamp, else all)
  build_all            build_sourcepack + _discover_expected_timestamps +
                       _read_inventory_timestamps, builder.py:163-239, 340-382

Pinning (tests/test_host_sourcepack.py): a hand-written three-record index
(rows, order, slots, path ids, keys and _SUCCESS text worked out by hand);
the reference's own materializer test case (a lone .done marker yields a
header-only index and both markers, fastxc/operators/sourcepack/tests/
test_async_materializer.py:6-24); the sort is a permutation of the input
rows, in non-decreasing key order, stable for identical keys.
"""
from __future__ import annotations

import csv
import os
import re

HEADER = ["timestamp", "path_id", "component_slot", "source_key", "receiver_key", "src_id",
          "rec_id", "src_network", "src_station", "src_location", "src_component",
          "rec_network", "rec_station", "rec_location", "rec_component", "npts", "dt", "dist",
          "az", "baz", "record_path", "record_offset", "bytes", "storage_kind",
          "final_pair_path"]

REQUIRED = {"timestamp", "pack_path", "offset", "bytes", "src_id", "rec_id", "src_network",
            "src_station", "src_location", "src_component", "rec_network", "rec_station",
            "rec_location", "rec_component", "npts", "dt", "dist", "az", "baz"}

_RANK = {"E": 0, "1": 0, "N": 1, "2": 1, "Z": 2, "3": 2, "R": 3, "T": 4}


def component_slot(src_comp: str, rec_comp: str) -> int:
    def rank(c):
        return _RANK.get(c.upper()[-1:], 100) if c else 100
    return rank(src_comp) * 3 + rank(rec_comp)


def path_id_text(row: dict) -> str:
    pid = (row.get("path_id") or "").strip()
    if pid:
        return pid.zfill(8) if pid.isdigit() else pid
    return "%08d" % (int(row["src_id"]) * 10000 + int(row["rec_id"]))


def safe_name(text: str) -> str:
    return re.sub(r"[^0-9A-Za-z._:-]+", "-", text).strip("-") or "item"


def _key(row, prefix):
    return ".".join(row[prefix + k] for k in ("_network", "_station", "_location"))


def read_xcpack(tsv: str, first_seq: int) -> list[dict]:
    out = []
    folder = os.path.dirname(tsv)
    with open(tsv, encoding="utf-8-sig", newline="") as fh:
        rd = csv.DictReader(fh, delimiter="\t")
        if rd.fieldnames is None:
            return out
        if REQUIRED - set(rd.fieldnames):
            raise ValueError("missing fields in %s" % tsv)
        for row in rd:
            p = os.path.expanduser(row["pack_path"])
            if not os.path.isabs(p):
                p = os.path.join(folder, p)
            p = _pathlib_text(p)
            if not os.path.isfile(p):
                raise FileNotFoundError(p)
            off, nb = int(row["offset"]), int(row["bytes"])
            if off < 0 or nb <= 0:
                raise ValueError("bad offset/bytes in %s" % tsv)
            out.append(dict(seq=first_seq + len(out), row=row, path=p, offset=off, bytes=nb))
    return out


def _pathlib_text(p: str) -> str:
    # pathlib's spelling: "." parts and doubled or trailing slashes dropped, ".." kept
    parts = [x for x in p.split("/") if x not in ("", ".")]
    return ("/" if p.startswith("/") else "") + "/".join(parts)


def sort_key(r: dict):
    row = r["row"]
    return (path_id_text(row), component_slot(row["src_component"], row["rec_component"]),
            _key(row, "src"), _key(row, "rec"), r["seq"])


def index_row(r: dict) -> dict:
    row = r["row"]
    return {
        "timestamp": row["timestamp"], "path_id": path_id_text(row),
        "component_slot": str(component_slot(row["src_component"], row["rec_component"])),
        "source_key": _key(row, "src"), "receiver_key": _key(row, "rec"),
        **{k: row[k] for k in HEADER[5:20]},
        "record_path": r["path"], "record_offset": str(r["offset"]), "bytes": str(r["bytes"]),
        "storage_kind": "xcpack", "final_pair_path": row.get("final_pair_path") or "",
    }


def write_one(out_dir: str, ts: str, recs: list[dict], sort: bool = True) -> dict:
    """Writes <out>/<safe ts>/{sourcepack_index.tsv,_SUCCESS}; returns counts."""
    target = os.path.join(out_dir, safe_name(ts))
    os.makedirs(target, exist_ok=True)
    recs = sorted(recs, key=sort_key if sort else (lambda r: r["seq"]))
    sources = {_key(r["row"], "src") for r in recs}
    nbytes = sum(r["bytes"] for r in recs)
    with open(os.path.join(target, "sourcepack_index.tsv"), "w", encoding="utf-8",
              newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=HEADER, delimiter="\t")
        w.writeheader()
        for r in recs:
            w.writerow(index_row(r))
    with open(os.path.join(target, "_SUCCESS"), "w", encoding="utf-8") as fh:
        fh.write("timestamp\t%s\nrecords\t%d\nsources\t%d\nbytes_indexed\t%d\n"
                 % (ts, len(recs), len(sources), nbytes))
    return dict(records=len(recs), sources=len(sources), bytes=nbytes)


def _tsvs(xcpack_dir):
    return sorted(os.path.join(xcpack_dir, n) for n in os.listdir(xcpack_dir)
                  if n.endswith(".tsv") and os.path.isfile(os.path.join(xcpack_dir, n)))


def build_timestamp(xcpack_dir: str, out_dir: str, ts: str, sort: bool = True) -> dict:
    allt = _tsvs(xcpack_dir)
    sel = [p for p in allt if os.path.basename(p).startswith(ts)] or allt
    recs, seq = [], 0
    for p in sel:
        got = read_xcpack(p, seq)
        seq += len(got)
        recs += [r for r in got if r["row"]["timestamp"] == ts]
    return write_one(out_dir, ts, recs, sort)


def manifest_timestamps(xcpack_dir: str) -> set:
    parent = os.path.dirname(xcpack_dir)
    ws = os.path.dirname(parent) if os.path.basename(parent) == "ncf" else parent
    man = os.path.join(ws, "manifest", "timestamp_index.tsv")
    if not os.path.isfile(man):
        return set()
    with open(man, encoding="utf-8-sig", newline="") as fh:
        rd = csv.DictReader(fh, delimiter="\t")
        if not rd.fieldnames or "timestamp" not in rd.fieldnames:
            return set()
        return {(r.get("timestamp") or "").strip() for r in rd} - {""}


def build_all(xcpack_dir: str, out_dir: str, sort: bool = True) -> dict:
    by_ts, seq = {}, 0
    for p in _tsvs(xcpack_dir):
        got = read_xcpack(p, seq)
        seq += len(got)
        for r in got:
            by_ts.setdefault(r["row"]["timestamp"], []).append(r)
    stamps = set(by_ts) | manifest_timestamps(xcpack_dir)
    stamps |= {n[:-5] for n in os.listdir(xcpack_dir)
               if n.endswith(".done") and not n.startswith("_")}
    if not stamps:
        raise FileNotFoundError("no timestamps in %s" % xcpack_dir)
    tot = dict(timestamps=0, records=0, sources=0, bytes=0)
    for ts in sorted(stamps):
        c = write_one(out_dir, ts, by_ts.get(ts, []), sort)
        tot["timestamps"] += 1
        for k in ("records", "sources", "bytes"):
            tot[k] += c[k]
    return tot
