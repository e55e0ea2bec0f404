"""File formats of the stage boundary, restated for the tests (writers of
inputs, readers of outputs).  Test infrastructure only.

  StepPack v3      native/sac2spec/io/stepack_writer.c:13-70 (header 232 B,
                   NSLC entry 128 B, TSV columns), :433-535 (one record per batch)
  xcpack           native/xc/pack_io/pack_writer.cu:47-65 (names), :151-156 and
                   :275-330 (TSV columns and row format); records
                   [SAC header 632 B][npts float32] (native/xc/io/sac_record.cu)
  allowed_paths    fastxc/inventory/planner.py:368-397 (9 columns)
  SourcePack       native/pws/io/sourcepack_io.cu (25-column index, records
                   [SAC header][npts float32] at record_offset of record_path)
  SAC header       70 float32, 40 int32, 192 bytes of text
"""
from __future__ import annotations

import os
import struct

import numpy as np

STEPACK_TSV_COLUMNS = [
    "timestamp", "worker_id", "batch_seq", "start_group", "group_count", "pack_path", "offset",
    "bytes", "version", "layout", "nstep", "nslc_start", "nslc_count", "batch_nslc_count", "nspec",
    "dt", "df", "nslc_table_bytes", "payload_offset", "payload_bytes", "batch_payload_bytes",
    "step_bytes", "pitch_step_bytes", "nslc_step_bytes", "first_nsl_id", "last_nsl_id"]

XCPACK_TSV_COLUMNS = [
    "timestamp", "worker_id", "anchor_block", "target_begin_block", "target_end_block",
    "block_size", "anchor_begin", "anchor_end", "target_begin", "target_end", "pack_path",
    "offset", "bytes", "src_id", "rec_id", "src_network", "src_station", "src_location",
    "src_component", "rec_network", "rec_station", "rec_location", "rec_component", "npts", "dt",
    "dist", "az", "baz", "final_pair_path"]

HEADER_FMT = "<8sIIIIQQQIIIIffQQQQ64s64s"
NSLC_FMT = "<IIff16s32s16s16s32s"
assert struct.calcsize(HEADER_FMT) == 232 and struct.calcsize(NSLC_FMT) == 128


def _fixed(text: str, size: int) -> bytes:
    b = text.encode()[: size - 1]
    return b + b"\0" * (size - len(b))


def _g9(x: float) -> str:
    return "%.9g" % x


def write_stepack(root, worker_id, batch_seq, nodes, spectra, dt, df, start_group=0,
                  group_count=None):
    """nodes: list of dicts (timestamp, nsl_id (text), gnsl_id, stla, stlo, network, station,
    location, component); spectra complex64 [nstep, len(nodes), nspec]."""
    os.makedirs(root, exist_ok=True)
    spectra = np.ascontiguousarray(spectra, dtype=np.complex64)
    nstep, n, nspec = spectra.shape
    assert n == len(nodes)
    group_count = n if group_count is None else group_count
    num_ch = n // group_count if group_count and n % group_count == 0 else 0
    table_bytes = n * 128
    nslc_step = nspec * 8
    pitch = n * nslc_step
    payload_bytes = nstep * pitch
    payload_offset = 232 + table_bytes
    record_bytes = payload_offset + payload_bytes
    hdr = struct.pack(HEADER_FMT, b"FXCSTPK\0", 3, 232, 128, 2, batch_seq, start_group,
                      group_count, worker_id, nstep, n, nspec, dt, df, table_bytes,
                      payload_offset, payload_bytes, pitch,
                      _fixed(nodes[0]["timestamp"], 64), b"\0" * 64)
    table = b"".join(
        struct.pack(NSLC_FMT, i, nd["gnsl_id"], nd["stla"], nd["stlo"], _fixed(nd["network"], 16),
                    _fixed(nd["station"], 32), _fixed(nd["location"], 16),
                    _fixed(nd["component"], 16), b"\0" * 32)
        for i, nd in enumerate(nodes))
    root = os.path.abspath(root)
    pack = os.path.join(root, "w%03d.b%06d.stepack" % (worker_id, batch_seq))
    with open(pack, "wb") as f:
        f.write(hdr)
        f.write(table)
        f.write(spectra.tobytes())
    rows = ["\t".join(STEPACK_TSV_COLUMNS)]
    b = 0
    while b < n:
        e = b + 1
        while e < n and nodes[e]["timestamp"] == nodes[b]["timestamp"]:
            e += 1
        run = e - b
        sg, gc = start_group, group_count
        if num_ch and b % num_ch == 0 and run % num_ch == 0:
            sg, gc = start_group + b // num_ch, run // num_ch
        vals = [nodes[b]["timestamp"], worker_id, batch_seq, sg, gc, pack, 0, record_bytes, 3, 2,
                nstep, b, run, n, nspec, _g9(float(np.float32(dt))), _g9(float(np.float32(df))),
                table_bytes, payload_offset, nstep * run * nslc_step, payload_bytes,
                run * nslc_step, pitch, nslc_step, nodes[b]["nsl_id"], nodes[e - 1]["nsl_id"]]
        rows.append("\t".join(str(v) for v in vals))
        b = e
    with open(os.path.join(root, "w%03d.b%06d.tsv" % (worker_id, batch_seq)), "w") as f:
        f.write("\n".join(rows) + "\n")
    return pack


def read_stepack(pack):
    with open(pack, "rb") as f:
        data = f.read()
    h = struct.unpack_from(HEADER_FMT, data, 0)
    (magic, version, hsize, esize, layout, batch_seq, start_group, group_count, worker, nstep,
     n, nspec, dt, df, table_bytes, poff, pbytes, pitch, first_ts, _res) = h
    nodes = []
    for i in range(n):
        e = struct.unpack_from(NSLC_FMT, data, 232 + 128 * i)
        nodes.append(dict(file_index=e[0], gnsl_id=e[1], stla=e[2], stlo=e[3],
                          network=e[4].rstrip(b"\0").decode(), station=e[5].rstrip(b"\0").decode(),
                          location=e[6].rstrip(b"\0").decode(),
                          component=e[7].rstrip(b"\0").decode()))
    spec = np.frombuffer(data, dtype=np.complex64, count=nstep * n * nspec, offset=poff)
    hdr = dict(magic=magic, version=version, header_size=hsize, nslc_entry_size=esize,
               layout=layout, batch_seq=batch_seq, start_group=start_group,
               group_count=group_count, worker_id=worker, nstep=nstep, nslc_count=n, nspec=nspec,
               dt=dt, df=df, nslc_table_bytes=table_bytes, payload_offset=poff,
               payload_bytes=pbytes, pitch_step_bytes=pitch,
               first_timestamp=first_ts.rstrip(b"\0").decode())
    return hdr, nodes, spec.reshape(nstep, n, nspec)


def write_allowed_paths(path, records):
    """records: (src_id, rec_id, gcarc, dist_km, az, baz) with src_id <= rec_id."""
    with open(path, "w") as f:
        f.write("path_id\tsrc_nsl_id\trec_nsl_id\tsrc_nsl\trec_nsl\tgreat_circle_deg\tdistance_km"
                "\tazimuth_deg\tback_azimuth_deg\n")
        for s, r, g, d, az, baz in records:
            f.write("%d\t%d\t%d\tN%d\tN%d\t%.6f\t%.6f\t%.6f\t%.6f\n" % (s * 10000 + r, s, r, s, r,
                                                                    g, d, az, baz))


SAC_FLOATS = ["delta", "depmin", "depmax", "scale", "odelta", "b", "e", "o", "a", "internal1"] + \
    ["t%d" % i for i in range(10)] + ["f"] + ["resp%d" % i for i in range(10)] + \
    ["stla", "stlo", "stel", "stdp", "evla", "evlo", "evel", "evdp", "mag"] + \
    ["user%d" % i for i in range(10)] + \
    ["dist", "az", "baz", "gcarc", "internal2", "internal3", "depmen", "cmpaz", "cmpinc",
     "xminimum", "xmaximum", "yminimum", "ymaximum"] + ["unused_f%d" % i for i in range(7)]
SAC_INTS = ["nzyear", "nzjday", "nzhour", "nzmin", "nzsec", "nzmsec", "nvhdr", "norid", "nevid",
            "npts", "internal7", "internal8", "unused13", "unused14", "unused15", "iftype", "idep",
            "iztype", "unused16", "iinst", "istreg", "ievreg", "ievtyp", "iqual", "isynth"] + \
    ["unused%d" % i for i in range(17, 27)] + ["leven", "lpspol", "lovrok", "lcalda", "unused27"]
SAC_TEXT = [("kstnm", 8), ("kevnm", 16), ("khole", 8), ("ko", 8), ("ka", 8)] + \
    [("kt%d" % i, 8) for i in range(10)] + \
    [("kf", 8), ("kuser0", 8), ("kuser1", 8), ("kuser2", 8), ("kcmpnm", 8), ("knetwk", 8),
     ("kdatrd", 8), ("kinst", 8)]
assert len(SAC_FLOATS) == 70 and len(SAC_INTS) == 40


def parse_sac_header(b: bytes) -> dict:
    assert len(b) >= 632
    fl = struct.unpack_from("<70f", b, 0)
    it = struct.unpack_from("<40i", b, 280)
    out = dict(zip(SAC_FLOATS, fl))
    out.update(zip(SAC_INTS, it))
    off = 440
    for name, size in SAC_TEXT:
        out[name] = b[off:off + size].decode(errors="replace")
        off += size
    return out


def read_xcpack_dir(xcpack_dir):
    """All records of all TSV sidecars: list of (row dict, header dict, data float32)."""
    out = []
    for name in sorted(os.listdir(xcpack_dir)):
        if not name.endswith(".tsv"):
            continue
        with open(os.path.join(xcpack_dir, name)) as f:
            lines = f.read().splitlines()
        assert lines[0].split("\t") == XCPACK_TSV_COLUMNS, name
        for line in lines[1:]:
            row = dict(zip(XCPACK_TSV_COLUMNS, line.split("\t")))
            with open(row["pack_path"], "rb") as pf:
                pf.seek(int(row["offset"]))
                blob = pf.read(int(row["bytes"]))
            hd = parse_sac_header(blob)
            data = np.frombuffer(blob, dtype=np.float32, offset=632)
            out.append((row, hd, data))
    return out


def sac_null_bytes() -> bytearray:
    """The reference's null SAC header: -12345 everywhere, nvhdr 6, text "-12345"."""
    b = bytearray(632)
    struct.pack_into("<70f", b, 0, *([-12345.0] * 70))
    ints = [-12345] * 40
    ints[6] = 6
    struct.pack_into("<40i", b, 280, *ints)
    off = 440
    for _name, size in SAC_TEXT:
        b[off:off + size] = (b"-12345" + b" " * (size - 6))
        off += size
    return b


def sac_bytes(data, delta, stla=0.0, stlo=0.0, kstnm="", knetwk="", kcmpnm="", b=0.0):
    """One SAC record: 632-byte header then float32 samples."""
    data = np.ascontiguousarray(data, dtype=np.float32)
    h = sac_null_bytes()
    struct.pack_into("<f", h, 0, delta)                                   # delta
    struct.pack_into("<ff", h, 5 * 4, b, b + (len(data) - 1) * delta)    # b, e
    struct.pack_into("<ff", h, SAC_FLOATS.index("stla") * 4, stla, stlo)
    struct.pack_into("<i", h, 280 + 4 * SAC_INTS.index("npts"), len(data))
    struct.pack_into("<i", h, 280 + 4 * SAC_INTS.index("iftype"), 1)
    struct.pack_into("<i", h, 280 + 4 * SAC_INTS.index("leven"), 1)
    for name, value in (("kstnm", kstnm), ("knetwk", knetwk), ("kcmpnm", kcmpnm)):
        off = 440 + sum(s for n, s in SAC_TEXT[:[n for n, _ in SAC_TEXT].index(name)])
        size = dict(SAC_TEXT)[name]
        h[off:off + size] = value.encode()[:size].ljust(size, b" ")
    return bytes(h) + data.tobytes()


def write_sac(path, data, delta, stla=0.0, stlo=0.0, kstnm="", knetwk="", kcmpnm=""):
    with open(path, "wb") as f:
        f.write(sac_bytes(data, delta, stla, stlo, kstnm, knetwk, kcmpnm))


SOURCEPACK_COLUMNS = [
    "timestamp", "path_id", "component_slot", "source_key", "receiver_key", "src_id", "rec_id",
    "src_network", "src_station", "src_location", "src_component", "rec_network", "rec_station",
    "rec_location", "rec_component", "npts", "dt", "dist", "az", "baz", "record_path",
    "record_offset", "bytes", "storage_kind", "final_pair_path"]


def read_sourcepack_index(path):
    """Rows of a sourcepack_index.tsv as dicts, with each record's header and samples."""
    with open(path) as f:
        lines = f.read().splitlines()
    assert lines[0].split("\t") == SOURCEPACK_COLUMNS, lines[0]
    out = []
    for line in lines[1:]:
        row = dict(zip(SOURCEPACK_COLUMNS, line.split("\t")))
        with open(row["record_path"], "rb") as pf:
            pf.seek(int(row["record_offset"]))
            blob = pf.read(int(row["bytes"]))
        out.append((row, parse_sac_header(blob), np.frombuffer(blob, dtype=np.float32, offset=632)))
    return out
