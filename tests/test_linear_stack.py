"""Linear stack (fastxc_amd.linear_stack, the mirror of
fastxc/operators/stacking/linear_stack.py): grouping, naming and errors on
the CPU; the FP64-accumulator kernel and the file-to-file operator on the GPU,
bit-exact against the oracle's float64 running sum."""
import csv
import os
import struct

import numpy as np
import pytest

from fastxc_amd import linear_stack as ls
from oracle import stack
from tests import fxformats as ff


def _write_shards(tmp, groups, n=201, delta=0.1, seed=0, nshards=2, slot_text=True):
    """groups: list of (path_id text, slot, src, rec, comps, count).  Records
    of a group spread over shards and timestamps; each shard index sorted by
    the group key.  Returns (index paths, {(key): records [T, n] in timestamp order})."""
    rng = np.random.default_rng(seed)
    shards = [[] for _ in range(nshards)]
    truth = {}
    for pid, slot, s, r, (cs, cr), cnt in groups:
        recs = []
        for k in range(cnt):
            ts = "2021.%03d.0000" % (10 + k)
            x = rng.standard_normal(n).astype(np.float32) * np.float32(10.0 ** rng.integers(-3, 4))
            recs.append((ts, x))
            shards[k % nshards].append((pid, slot, s, r, cs, cr, ts, x))
        truth[(pid, slot)] = np.stack([x for _, x in sorted(recs, key=lambda z: z[0])])
    paths = []
    for w, rows in enumerate(shards):
        rows = [rows[i] for i in rng.permutation(len(rows))]
        pack = os.path.join(tmp, "xc.w%03d.pack" % w)
        body = []
        with open(pack, "wb") as f:
            for pid, slot, s, r, cs, cr, ts, x in rows:
                off = f.tell()
                blob = ff.sac_bytes(x, delta, kstnm="R%d" % r, knetwk="YY", kcmpnm=cr)
                f.write(blob)
                fpp = "/ncf/YY.S%d/YY.S%d/YY-YY.S%d-S%d.%s-%s.sac" % (s, r, s, r, cs, cr)
                body.append((ls._group_key({"path_id": pid, "component_slot": str(slot)}),
                             [ts, pid, slot if slot_text else "", "YY.S%d" % s, "YY.S%d" % r, s, r,
                              "YY", "S%d" % s, "", cs, "YY", "S%d" % r, "", cr, n, "%g" % delta,
                              5.0 * r, 40.0 + s, 220.0 + r, pack, off, len(blob), "xc_pack", fpp]))
        body.sort(key=lambda z: z[0])
        idx = os.path.join(tmp, "xc.w%03d.tsv" % w)
        with open(idx, "w") as f:
            f.write("\t".join(ff.SOURCEPACK_COLUMNS) + "\n")
            for _, row in body:
                f.write("\t".join(str(v) for v in row) + "\n")
        paths.append(idx)
    return paths, truth


GROUPS = [("10002", 8, 1, 2, ("BHZ", "BHZ"), 6),
          ("10002", 0, 1, 2, ("BHE", "BHE"), 3),
          ("100004", 8, 1, 4, ("BHZ", "BHZ"), 11),   # numeric order: after 20003
          ("20003", 8, 2, 3, ("BHZ", "BHZ"), 1),
          ("30030", 8, 3, 30, ("BHZ", "BHZ"), 24)]


def test_dry_run_groups_in_reference_order(tmp_path):
    idx, truth = _write_shards(str(tmp_path), GROUPS)
    res = ls.linear_stack_sourcepack_indexes(idx, tmp_path / "stack", dry_run=True)
    # path ids zero-filled to 8 digits, so numeric order; then component slot
    want = sorted(truth, key=lambda k: (k[0].zfill(8), k[1]))
    got = [(r.input_path.name.split(".")[1], r.segment_count) for r in res]
    assert [truth[k].shape[0] for k in want] == [c for _, c in got]
    sta = {(pid, slot): "S%d-S%d" % (s, r) for pid, slot, s, r, _, _ in GROUPS}
    assert [sta[k] for k in want] == [g for g, _ in got]
    assert all(r.output_path == ls.linear_sourcepack_dir(tmp_path / "stack") / "linearstack.pack"
               for r in res)
    assert not (tmp_path / "stack").exists()


def test_component_slot_fallback_rank():
    assert ls._group_key({"path_id": "12", "component_slot": "",
                          "src_component": "BHN", "rec_component": "BHZ"}) == ("00000012", 5)
    assert ls._group_key({"path_id": "", "src_id": "3", "rec_id": "7", "component_slot": "2"}) == \
        ("00030007", 2)
    assert ls._group_key({"path_id": "ABC", "component_slot": "x",
                          "src_component": "Q", "rec_component": "R"}) == ("ABC", 300)


def test_output_naming():
    p = ls.linear_output_path("/a/b/XX-YY.S1-S2.BHZ-BHZ.sac", "/w/stack")
    assert p == ls.linear_sourcepack_dir("/w/stack").parent.parent / "linearstack" / \
        "XX-YY.S1-S2" / "XX-YY.S1-S2.BHZ-BHZ.linearstack.sac"
    with pytest.raises(ValueError):
        ls.linear_output_path("bad.sac", "/w")


def test_byte_order_detection():
    le = bytes(ff.sac_null_bytes())
    from fastxc_amd.sacrec import byte_order

    assert byte_order(le) == "<"
    f = struct.unpack("<70f", le[:280])
    i = struct.unpack("<40i", le[280:440])
    be = struct.pack(">70f", *f) + struct.pack(">40i", *i) + le[440:]
    assert byte_order(be) == ">"


def test_missing_index_fields(tmp_path):
    idx = tmp_path / "bad.tsv"
    idx.write_text("timestamp\tpath_id\tsource_key\n2021.010.0000\t10002\tA\n")
    with pytest.raises(ValueError, match="missing fields"):
        ls.linear_stack_sourcepack_indexes([idx], tmp_path / "stack", dry_run=True)


@pytest.mark.gpu
def test_fp64_kernel_bit_exact(gpu):
    import torch
    from fastxc_amd import ops

    rng = np.random.default_rng(3)
    counts = [1, 2, 7, 300, 1000]
    n = 2001
    rows = []
    for c in counts:
        scale = (10.0 ** rng.integers(-4, 5, size=(c, 1))).astype(np.float32)
        rows.append((rng.standard_normal((c, n)) * scale).astype(np.float32))
    off = np.concatenate([[0], np.cumsum(counts)])
    x = torch.from_numpy(np.concatenate(rows)).to(gpu)
    out = ops.linear_stack(x, off).cpu().numpy()
    for p, r in enumerate(rows):
        assert np.array_equal(out[p], stack.linear_stack(r)), p


@pytest.mark.gpu
def test_file_to_file_bit_exact(gpu, tmp_path):
    n, delta = 201, 0.1
    idx, truth = _write_shards(str(tmp_path), GROUPS, n=n, delta=delta)
    res = ls.linear_stack_sourcepack_indexes(idx, tmp_path / "stack", device=gpu)
    out_dir = ls.linear_sourcepack_dir(tmp_path / "stack")
    assert not out_dir.with_name("STACK.tmp").exists()
    with open(out_dir / "sourcepack_index.tsv", newline="") as f:
        rows = list(csv.DictReader(f, delimiter="\t"))
    with open(out_dir / "sourcepack_index.tsv", "rb") as f:
        assert f.readline().endswith(b"\r\n")      # csv.DictWriter line ends
    want_keys = sorted(truth, key=lambda k: (k[0].zfill(8), k[1]))
    assert [(r["path_id"], int(r["component_slot"])) for r in rows] == want_keys
    assert len(res) == len(rows)
    for r, row in zip(res, rows):
        key = (row["path_id"], int(row["component_slot"]))
        recs = truth[key]
        with open(row["record_path"], "rb") as f:
            f.seek(int(row["record_offset"]))
            blob = f.read(int(row["bytes"]))
        hd = ff.parse_sac_header(blob)
        data = np.frombuffer(blob, np.float32, offset=632)
        assert np.array_equal(data, stack.linear_stack(recs)), key
        assert hd["user0"] == recs.shape[0] and hd["npts"] == n and hd["iftype"] == 1
        assert (hd["nzyear"], hd["nzjday"], hd["nzhour"], hd["nzmsec"]) == (2010, 214, 0, 0)
        assert hd["kstnm"].startswith("R")
        assert row["timestamp"] == "STACK" and row["storage_kind"] == "linearstack_pack"
        assert row["dt"] == "%.9g" % np.float32(delta) and row["npts"] == str(n)
        assert row["final_pair_path"] == str(r.output_path)
        assert row["final_pair_path"].endswith(".linearstack.sac")
        assert r.segment_count == recs.shape[0] and r.sample_count == n
    succ = (out_dir / "_SUCCESS").read_text().splitlines()
    assert succ == ["records\t%d" % len(rows), "pack\t%s" % (out_dir / "linearstack.pack"),
                    "index\t%s" % (out_dir / "sourcepack_index.tsv")]


def test_mismatched_records_raise(tmp_path):
    """Checked while reading, before any device work (host logic)."""
    idx, _ = _write_shards(str(tmp_path), [("10002", 8, 1, 2, ("BHZ", "BHZ"), 3)], n=50)
    # rewrite one record's header with another delta
    with open(idx[1]) as f:
        row = f.read().splitlines()[1].split("\t")
    path, off = row[20], int(row[21])
    with open(path, "r+b") as f:
        f.seek(off)
        f.write(struct.pack("<f", 0.2))
    with pytest.raises(ValueError, match="delta mismatch"):
        ls.linear_stack_sourcepack_indexes(idx, tmp_path / "stack", device="cuda:0")
