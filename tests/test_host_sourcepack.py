"""sourcepack_index (the native SourcePack indexer) against the Python
restatement of the reference builder (oracle/sourcepack.py): byte-identical
index files and markers.  CPU only; xc_fast's --sourcepack-dir run is in
tests/test_host_xc.py (GPU)."""
import os
import subprocess

import numpy as np
import pytest

from oracle import sourcepack as osp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SPI = os.path.join(ROOT, "bin", "sourcepack_index")
XC = os.path.join(ROOT, "bin", "xc_fast")

pytestmark = pytest.mark.skipif(not os.path.exists(SPI), reason="bin/sourcepack_index not built")

XC_COLUMNS = ["timestamp", "worker_id", "anchor_block", "target_begin_block", "target_end_block",
              "block_size", "anchor_begin", "anchor_end", "target_begin", "target_end",
              "pack_path", "offset", "bytes", "src_id", "rec_id", "src_network", "src_station",
              "src_location", "src_component", "rec_network", "rec_station", "rec_location",
              "rec_component", "npts", "dt", "dist", "az", "baz", "final_pair_path"]


def run(args):
    return subprocess.run([SPI] + args, capture_output=True, text=True, timeout=120)


def write_tsv(path, rows, columns=XC_COLUMNS):
    with open(path, "w", encoding="utf-8", newline="") as fh:
        fh.write("\t".join(columns) + "\n")
        for r in rows:
            fh.write("\t".join(str(r.get(c, "")) for c in columns) + "\n")


def row(ts, src, rec, sc, rc, pack, offset, nbytes, net="XX", loc="", **extra):
    r = dict(timestamp=ts, worker_id=0, anchor_block=0, target_begin_block=0, target_end_block=0,
             block_size=4, anchor_begin=0, anchor_end=4, target_begin=0, target_end=4,
             pack_path=pack, offset=offset, bytes=nbytes, src_id=src, rec_id=rec,
             src_network=net, src_station="S%02d" % src, src_location=loc, src_component=sc,
             rec_network=net, rec_station="S%02d" % rec, rec_location=loc, rec_component=rc,
             npts=401, dt=0.1, dist=12.5, az=30, baz=210,
             final_pair_path="/ncf/%s/S%02d-S%02d.%s-%s.sac" % (ts, src, rec, sc, rc))
    r.update(extra)
    return r


def tree(root):
    """{relative path: bytes} of every file under root."""
    out = {}
    for d, _, files in os.walk(root):
        for f in files:
            p = os.path.join(d, f)
            with open(p, "rb") as fh:
                out[os.path.relpath(p, root)] = fh.read()
    return out


def test_known_answer_index(tmp_path):
    """Three records worked out by hand: path ids from src*10000+rec, slots
    rank(src)*3+rank(rec) with E 0, N 1, Z 2, T 4, sorted by (path id, slot)."""
    xdir = tmp_path / "ncf" / "xcpack"
    xdir.mkdir(parents=True)
    pack = xdir / "T1.i000000.j000000.w000.p000.xcpack"
    pack.write_bytes(b"\0" * 300)
    write_tsv(xdir / "T1.i000000.j000000.w000.p000.tsv",
              [row("T1", 2, 3, "BHZ", "BHZ", pack.name, 0, 100),
               row("T1", 1, 2, "BHE", "BHN", pack.name, 100, 100),
               row("T1", 1, 2, "BHZ", "BHT", pack.name, 200, 100)])
    P = str(pack)
    want = ("\t".join(osp.HEADER) + "\r\n"
            + "T1\t00010002\t1\tXX.S01.\tXX.S02.\t1\t2\tXX\tS01\t\tBHE\tXX\tS02\t\tBHN\t401\t0.1"
              "\t12.5\t30\t210\t%s\t100\t100\txcpack\t/ncf/T1/S01-S02.BHE-BHN.sac\r\n" % P
            + "T1\t00010002\t10\tXX.S01.\tXX.S02.\t1\t2\tXX\tS01\t\tBHZ\tXX\tS02\t\tBHT\t401\t0.1"
              "\t12.5\t30\t210\t%s\t200\t100\txcpack\t/ncf/T1/S01-S02.BHZ-BHT.sac\r\n" % P
            + "T1\t00020003\t8\tXX.S02.\tXX.S03.\t2\t3\tXX\tS02\t\tBHZ\tXX\tS03\t\tBHZ\t401\t0.1"
              "\t12.5\t30\t210\t%s\t0\t100\txcpack\t/ncf/T1/S02-S03.BHZ-BHZ.sac\r\n" % P)
    marker = "timestamp\tT1\nrecords\t3\nsources\t2\nbytes_indexed\t300\n"
    osp.build_all(str(xdir), str(tmp_path / "oracle"))
    r = run(["-I", str(tmp_path / "ncf"), "-O", str(tmp_path / "sp")])
    assert r.returncode == 0, r.stderr
    assert "Built 1 timestamp source index(es), 2 source(s), 3 record(s)." in r.stdout
    for out in ("oracle", "sp"):
        assert (tmp_path / out / "T1" / "sourcepack_index.tsv").read_bytes() == want.encode()
        assert (tmp_path / out / "T1" / "_SUCCESS").read_text() == marker
    # the batch build writes no global marker (only the materializer does)
    assert not (tmp_path / "sp" / "_SUCCESS").exists()


def test_lone_done_marker_gives_header_only_index(tmp_path):
    """The reference's materializer case: ncf/xcpack holds only
    2020.001.0000.done (fastxc/operators/sourcepack/tests/test_async_materializer.py:6-24)."""
    xdir = tmp_path / "ncf" / "xcpack"
    xdir.mkdir(parents=True)
    (xdir / "2020.001.0000.done").write_text("")
    (xdir / "_SUCCESS.done").write_text("")            # names starting with _ are skipped
    r = run(["-I", str(xdir), "-O", str(tmp_path / "sp")])
    assert r.returncode == 0, r.stderr
    osp.build_all(str(xdir), str(tmp_path / "oracle"))
    assert tree(tmp_path / "sp") == tree(tmp_path / "oracle")
    assert sorted(os.listdir(tmp_path / "sp")) == ["2020.001.0000"]
    idx = (tmp_path / "sp" / "2020.001.0000" / "sourcepack_index.tsv").read_text()
    assert idx == "\t".join(osp.HEADER) + "\n"   # read back with universal newlines
    assert (tmp_path / "sp" / "2020.001.0000" / "_SUCCESS").read_text() == (
        "timestamp\t2020.001.0000\nrecords\t0\nsources\t0\nbytes_indexed\t0\n")


def _workspace(tmp_path, seed=0):
    """Three timestamps over four TSVs: mixed components (E/N/Z/1/2/3/R/T and
    unknown), duplicate keys, a path_id column in one TSV (short digits and
    text), relative, ./ and .. pack paths, a field the csv writer must quote,
    a timestamp that needs a safe name,
    a manifest timestamp without records and a .done-only timestamp."""
    rng = np.random.default_rng(seed)
    ws = tmp_path / "ws"
    xdir = ws / "ncf" / "xcpack"
    (xdir / "sub").mkdir(parents=True)
    stamps = ["20200502T04:00", "20200502T05:00", "2020 05#02 x"]
    comps = ["BHE", "BHN", "BHZ", "HH1", "HH2", "HH3", "BHR", "BHT", "BHX", "bhz", ""]
    for k, ts in enumerate(stamps):
        for part in range(2 if k < 2 else 1):
            stem = "%s.i%06d.j000000.w000.p%03d" % (ts, k, part)
            pack = xdir / (stem + ".xcpack")
            nrec = int(rng.integers(5, 40))
            pack.write_bytes(b"\0" * (nrec * 64))
            rows = []
            for i in range(nrec):
                s, q = sorted(rng.choice(12, size=2, replace=False) + 1)
                if rng.random() < 0.15 and rows:
                    s, q = int(rows[-1]["src_id"]), int(rows[-1]["rec_id"])   # duplicates
                spath = [pack.name, "./" + pack.name, "sub/../" + pack.name, str(pack)][i % 4]
                r = row(ts, int(s), int(q), comps[rng.integers(len(comps))],
                        comps[rng.integers(len(comps))], spath, i * 64, 64,
                        net=["XX", "YY"][i % 2], loc=["", "00"][i % 3 == 0])
                if i == 2:
                    r["final_pair_path"] = '/ncf/x"y z'     # written quoted, quote doubled
                rows.append(r)
            cols = XC_COLUMNS
            if k == 1 and part == 1:
                cols = XC_COLUMNS + ["path_id"]
                for i, r in enumerate(rows):
                    r["path_id"] = ["123", "P7", "", "000123456789"][i % 4]
            write_tsv(xdir / (stem + ".tsv"), rows, cols)
    (ws / "manifest").mkdir()
    with open(ws / "manifest" / "timestamp_index.tsv", "w") as fh:
        fh.write("timestamp\tfiles\n20200502T04:00\t3\n20200503T00:00\t0\n \t1\n")
    (xdir / "20200504T00:00.done").write_text("DONE\n")
    (xdir / "_SUCCESS").write_text("DONE\n")
    return ws, xdir, stamps


@pytest.mark.parametrize("sort", [True, False])
def test_build_all_matches_oracle(tmp_path, sort):
    ws, xdir, stamps = _workspace(tmp_path)
    args = ["-I", str(ws / "ncf"), "-O", str(tmp_path / "sp")] + ([] if sort else ["--keep-order"])
    prog = tmp_path / "progress" / "sp.tsv"
    r = run(args + ["--progress", str(prog)])
    assert r.returncode == 0, r.stderr
    tot = osp.build_all(str(xdir), str(tmp_path / "oracle"), sort=sort)
    got, want = tree(tmp_path / "sp"), tree(tmp_path / "oracle")
    assert sorted(got) == sorted(want)
    for k in want:
        assert got[k] == want[k], k
    names = sorted(os.listdir(tmp_path / "sp"))
    assert names == sorted(["20200502T04:00", "20200502T05:00", "20200503T00:00",
                            "20200504T00:00", "2020-05-02-x"])
    assert tot["timestamps"] == 5
    assert ("Built 5 timestamp source index(es), %d source(s), %d record(s)."
            % (tot["sources"], tot["records"])) in r.stdout
    assert "overall\tDONE\t5\t5\ttimestamps\tsource index ready" in prog.read_text()


def test_sorted_index_is_a_stable_keyed_permutation(tmp_path):
    """Order property independent of the restatement: every xcpack row appears
    once, keys never decrease, identical keys keep their read order."""
    ws, xdir, stamps = _workspace(tmp_path, seed=3)
    assert run(["-I", str(xdir), "-O", str(tmp_path / "sp")]).returncode == 0
    ts = "20200502T05:00"
    rows = []
    for p in sorted(xdir.glob(ts + "*.tsv")):
        rows += osp.read_xcpack(str(p), len(rows))
    idx = (tmp_path / "sp" / ts / "sourcepack_index.tsv").read_text().splitlines()[1:]
    got = [l.split("\t") for l in idx]
    H = {c: i for i, c in enumerate(osp.HEADER)}
    assert len(got) == len(rows)
    ident = [(g[H["record_path"]], int(g[H["record_offset"]])) for g in got]
    assert sorted(ident) == sorted((r["path"], r["offset"]) for r in rows)
    seq = {(r["path"], r["offset"]): r["seq"] for r in rows}
    keys = [(g[H["path_id"]], int(g[H["component_slot"]]), g[H["source_key"]],
             g[H["receiver_key"]], seq[i]) for g, i in zip(got, ident)]
    assert keys == sorted(keys)


def test_single_timestamp_matches_oracle(tmp_path):
    ws, xdir, stamps = _workspace(tmp_path, seed=1)
    for ts in stamps + ["20200504T00:00"]:
        r = run(["-I", str(xdir), "-O", str(tmp_path / "sp"), "--timestamp", ts])
        assert r.returncode == 0, r.stderr
        osp.build_timestamp(str(xdir), str(tmp_path / "oracle"), ts)
    assert tree(tmp_path / "sp") == tree(tmp_path / "oracle")


def test_rebuild_replaces_old_index(tmp_path):
    ws, xdir, stamps = _workspace(tmp_path, seed=2)
    out = tmp_path / "sp"
    (out / "20200502T04:00").mkdir(parents=True)
    (out / "20200502T04:00" / "stale.txt").write_text("x")
    (out / ".20200502T04:00.tmp").mkdir()
    assert run(["-I", str(ws / "ncf"), "-O", str(out)]).returncode == 0
    assert sorted(os.listdir(out / "20200502T04:00")) == ["_SUCCESS", "sourcepack_index.tsv"]
    assert not (out / ".20200502T04:00.tmp").exists()


def test_invalid_inputs_fail(tmp_path):
    xdir = tmp_path / "xcpack"
    xdir.mkdir()
    pack = xdir / "T.xcpack"
    pack.write_bytes(b"\0" * 64)
    assert run(["-I", str(tmp_path / "nowhere"), "-O", str(tmp_path / "o")]).returncode == 1
    assert run(["-I", str(xdir)]).returncode == 1
    assert run(["-I", str(xdir), "-O", str(tmp_path / "o"), "--bogus"]).returncode == 1
    # no timestamps at all
    assert run(["-I", str(xdir), "-O", str(tmp_path / "o")]).returncode == 1
    cases = {
        "missing_field": (XC_COLUMNS[:-5] + XC_COLUMNS[-4:], row("T", 1, 2, "Z", "Z", pack.name, 0, 64)),
        "missing_pack": (XC_COLUMNS, row("T", 1, 2, "Z", "Z", "gone.xcpack", 0, 64)),
        "negative_offset": (XC_COLUMNS, row("T", 1, 2, "Z", "Z", pack.name, -1, 64)),
        "zero_bytes": (XC_COLUMNS, row("T", 1, 2, "Z", "Z", pack.name, 0, 0)),
        "text_offset": (XC_COLUMNS, row("T", 1, 2, "Z", "Z", pack.name, "1x", 64)),
    }
    for name, (cols, r) in cases.items():
        tsv = xdir / "T.bad.tsv"
        write_tsv(tsv, [r], cols)
        res = run(["-I", str(xdir), "-O", str(tmp_path / "o")])
        assert res.returncode == 1, name
        with pytest.raises((ValueError, FileNotFoundError)):
            osp.build_all(str(xdir), str(tmp_path / "o2"))
        tsv.unlink()


@pytest.mark.skipif(not os.path.exists(XC), reason="bin/xc_fast not built")
@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="needs a host without a GPU (xc_fast must fail)")
def test_failed_xc_builds_done_timestamps_but_withholds_global_success(tmp_path):
    """xc_fast fails (no GPU here) after the materializer started: the final
    scan still indexes the .done timestamps already present, a stale global
    _SUCCESS is removed and the progress sidecar says FAILED
    (fastxc/operators/sourcepack/tests/test_async_materializer.py:27-49)."""
    from tests.test_host_xc import make_inputs

    sp_dir, paths, *_ = make_inputs(str(tmp_path))
    out = tmp_path / "ncf"
    (out / "xcpack").mkdir(parents=True)
    (out / "xcpack" / "2020.001.0000.done").write_text("DONE\n")
    spdir = tmp_path / "sourcepack"
    spdir.mkdir()
    (spdir / "_SUCCESS").write_text("stale\n")
    prog = tmp_path / "progress" / "sourcepack_progress.tsv"
    r = subprocess.run([XC, "-I", sp_dir, "-P", paths, "-O", str(out), "-C", "20",
                        "--sourcepack-dir", str(spdir), "--sourcepack-progress", str(prog)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 1
    assert (spdir / "2020.001.0000" / "_SUCCESS").is_file()
    assert (spdir / "2020.001.0000" / "sourcepack_index.tsv").is_file()
    assert not (spdir / "_SUCCESS").exists()
    assert "\tFAILED\t" in prog.read_text()


def test_python_build_sourcepack_mirror(tmp_path):
    """fastxc_amd.sourcepack.build_sourcepack: the reference's signature and
    result record, default output beside xcpack, errors for a missing input."""
    from fastxc_amd.sourcepack import build_sourcepack

    ws, xdir, stamps = _workspace(tmp_path, seed=4)
    res = build_sourcepack(ws / "ncf", progress_file=tmp_path / "p.tsv")
    tot = osp.build_all(str(xdir), str(tmp_path / "oracle"))
    assert res.output_dir == (ws / "ncf" / "sourcepack").resolve()
    assert len(res.index_paths) == tot["timestamps"] == 5
    assert (res.record_count, res.source_count, res.bytes_indexed) == (
        tot["records"], tot["sources"], tot["bytes"])
    assert tree(res.output_dir) == tree(tmp_path / "oracle")
    res2 = build_sourcepack(xdir, tmp_path / "kept", sort_within_source=False)
    assert res2.record_count == tot["records"]
    with pytest.raises(FileNotFoundError):
        build_sourcepack(tmp_path / "nowhere")
