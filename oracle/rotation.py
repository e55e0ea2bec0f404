"""ENZ -> RTZ rotation restated in numpy float64 -- test infrastructure only.

  rotate_matrix     build_rotate_matrix, fastxc/operators/rotation/python_rotate.py:34-95
                    (entries listed row by row, each a float32 product of the
                    float32 sines and cosines of the float32 radian angles)
  rotate64          rotate_enz_to_rtz, :98-104, the product in float64
  expected_stack    rotate_sourcepack_stack, :264-390 with _read_rotation_groups,
                    _rotation_group_sort_key, _legacy_rtz_path, _row_path_id,
                    _component_slot/_component_rank and _rotated_index_row
                    (:436-520): the index text, the record headers, the float64
                    rotated data and the _SUCCESS text the reference writes,
                    computed without writing anything.  The geometry function
                    (Rudoe, fastxc/inventory/planner.py:833) is passed in.

Pinning (tests/test_rotation.py): rotate_matrix against the reference's own
explicit per-component formulas (fastxc/operators/rotation/
verify_rotation_example.py:14-33, its seed and angles); orthogonality of the
matrix (a Kronecker product of two orthogonal 3x3 bases), so the rotation
keeps every sample's energy; az = 0, baz = 180 maps RR to NN and TT to EE;
the index, headers and _SUCCESS of a hand-worked two-pair stack (one pair
complete, one missing N-Z).
"""
from __future__ import annotations

import csv
import io
import struct
from pathlib import Path

import numpy as np

ENZ_ORDER = ("E-E", "E-N", "E-Z", "N-E", "N-N", "N-Z", "Z-E", "Z-N", "Z-Z")
RTZ_ORDER = ("R-R", "R-T", "R-Z", "T-R", "T-T", "T-Z", "Z-R", "Z-T", "Z-Z")
HEADER = ["timestamp", "path_id", "component_slot", "source_key", "receiver_key", "src_id",
          "rec_id", "src_network", "src_station", "src_location", "src_component",
          "rec_network", "rec_station", "rec_location", "rec_component", "npts", "dt", "dist",
          "az", "baz", "record_path", "record_offset", "bytes", "storage_kind",
          "final_pair_path"]


def trig32(az, baz):
    a = np.float32(np.deg2rad(az))
    b = np.float32(np.deg2rad(baz))
    return (np.sin(a, dtype=np.float32), np.cos(a, dtype=np.float32),
            np.sin(b, dtype=np.float32), np.cos(b, dtype=np.float32))


def rotate_matrix(az, baz) -> np.ndarray:
    sa, ca, sb, cb = trig32(az, baz)
    m = np.zeros((9, 9), np.float32)
    m[0, [0, 1, 3, 4]] = [-sa * sb, -sa * cb, -ca * sb, -ca * cb]
    m[1, [0, 1, 3, 4]] = [-sa * cb, sa * sb, -ca * cb, ca * sb]
    m[2, [2, 5]] = [sa, ca]
    m[3, [0, 1, 3, 4]] = [-ca * sb, -ca * cb, sa * sb, sa * cb]
    m[4, [0, 1, 3, 4]] = [-ca * cb, ca * sb, sa * cb, -sa * sb]
    m[5, [2, 5]] = [ca, -sa]
    m[6, [6, 7]] = [-sb, -cb]
    m[7, [6, 7]] = [-cb, sb]
    m[8, 8] = 1.0
    return m


def rotate64(enz, az, baz) -> np.ndarray:
    return rotate_matrix(az, baz).astype(np.float64) @ np.asarray(enz, np.float64)


# --------------------------------------------------------------- headers --
_F = {"delta": 0, "stla": 31, "stlo": 32, "evla": 35, "evlo": 36, "dist": 50, "az": 51,
      "baz": 52, "gcarc": 53, "cmpaz": 57, "cmpinc": 58}


def _endian(raw: bytes) -> str:
    if struct.unpack_from("<i", raw, 280 + 24)[0] in (6, 7):
        return "<"
    return ">" if struct.unpack_from(">i", raw, 280 + 24)[0] in (6, 7) else "<"


def _getf(raw, e, name):
    return struct.unpack_from(e + "f", raw, 4 * _F[name])[0]


def _defined(v):
    return np.isfinite(v) and abs(v + 12345.0) > 1e-3


def _read(row):
    with open(Path(row["record_path"]).expanduser(), "rb") as fh:
        fh.seek(int(row["record_offset"]))
        raw = fh.read(int(row["bytes"]))
    e = _endian(raw[:632])
    npts = struct.unpack_from(e + "i", raw, 280 + 36)[0]
    return raw[:632], e, npts, np.frombuffer(raw[632:], e + "f4").astype(np.float64)


def _path_id(row):
    pid = (row.get("path_id") or "").strip()
    if pid:
        return pid.zfill(8) if pid.isdigit() else pid
    try:
        return "%08d" % (int(row["src_id"]) * 10000 + int(row["rec_id"]))
    except (KeyError, ValueError, TypeError):
        return ""


def _slot(sc, rc):
    rank = {"E": 0, "1": 0, "R": 0, "N": 1, "2": 1, "T": 1, "Z": 2, "3": 2}
    return (rank.get(sc.upper()[-1:], 100) if sc else 100) * 3 + \
        (rank.get(rc.upper()[-1:], 100) if rc else 100)


def expected_stack(output_dir, method, geometry_fn):
    """dict(index=text, records=[(header bytes, float64 data)], success=text,
    pairs=n) of rotate_sourcepack_stack(output_dir, method)."""
    out = Path(output_dir).expanduser().resolve()
    method = method.strip().lower()
    src_index = out / "stack" / f"{method}_sourcepack" / "STACK" / "sourcepack_index.tsv"
    groups = {}
    with open(src_index, encoding="utf-8-sig", newline="") as fh:
        for row in csv.DictReader(fh, delimiter="\t"):
            groups.setdefault((_path_id(row), row["source_key"], row["receiver_key"]),
                              []).append(row)
    stack = out / "stack" / f"rtz_{method}_sourcepack" / "STACK"
    pack = stack / f"rtz_{method}.pack"
    text = io.StringIO(newline="")
    w = csv.DictWriter(text, fieldnames=HEADER, delimiter="\t")
    w.writeheader()
    records, pairs, skipped, offset = [], 0, 0, 0
    for key in sorted(groups, key=lambda k: (_path_id(groups[k][0]), k[1], k[2])):
        rows = groups[key]
        tags = {}
        for r in rows:
            tags[r["src_component"] + "-" + r["rec_component"]] = r
        if any(t not in tags for t in ENZ_ORDER):
            skipped += 1
            continue
        got = [_read(tags[t]) for t in ENZ_ORDER]
        raw, e, npts, _ = got[8]
        delta = _getf(raw, e, "delta")
        for (r2, e2, n2, _), t in zip(got, ENZ_ORDER):
            assert n2 == npts and abs(_getf(r2, e2, "delta") - delta) <= 1e-7
        ll = [_getf(raw, e, k) for k in ("evlo", "evla", "stlo", "stla")]
        if all(_defined(v) for v in ll):
            gcarc, az, baz, dist = geometry_fn(*ll)
        else:
            gcarc, az, baz, dist = (_getf(raw, e, k) for k in ("gcarc", "az", "baz", "dist"))
        rtz = rotate64(np.stack([g[3] for g in got]), az, baz)
        base = bytearray(raw)
        for name, v in (("gcarc", gcarc), ("az", az), ("baz", baz), ("dist", dist),
                        ("cmpaz", baz + 180.0 if baz - 180.0 < 0.0 else baz - 180.0),
                        ("cmpinc", 90.0)):
            struct.pack_into(e + "f", base, 4 * _F[name], float(v))
        first = rows[0]
        pair = (f"{first['src_network']}-{first['rec_network']}."
                f"{first['src_station']}-{first['rec_station']}")
        for k, tag in enumerate(RTZ_ORDER):
            hd = bytearray(base)
            hd[600:608] = tag.replace("-", "").encode().ljust(8, b" ")
            nbytes = 632 + 4 * rtz.shape[1]
            records.append((bytes(hd), rtz[k]))
            sc, rc = tag.split("-")
            row = {f: "" for f in HEADER}
            for f in ("path_id", "source_key", "receiver_key", "src_id", "rec_id", "src_network",
                      "src_station", "src_location", "rec_network", "rec_station",
                      "rec_location"):
                row[f] = first.get(f, "")
            row.update(timestamp="ROTATE", path_id=_path_id(first),
                       component_slot=str(_slot(sc, rc)), src_component=sc, rec_component=rc,
                       npts=str(npts), dt=f"{delta:.9g}", dist=f"{dist:.9g}", az=f"{az:.9g}",
                       baz=f"{baz:.9g}", record_path=pack.as_posix(), record_offset=str(offset),
                       bytes=str(nbytes), storage_kind=f"rtz_{method}_pack",
                       final_pair_path=(out / "stack" / f"rtz_{method}" / pair /
                                        f"{pair}.{tag}.sac").as_posix())
            w.writerow(row)
            offset += nbytes
        pairs += 1
    success = (f"pairs\t{pairs}\nskipped\t{skipped}\npack\t{pack}\n"
               f"index\t{stack / 'sourcepack_index.tsv'}\n")
    return dict(index=text.getvalue(), records=records, success=success, pairs=pairs,
                skipped=skipped)
