"""ncf_pws / ncf_tfpws executables: command line (CPU) and the file-to-file
run over SourcePack inputs against the FP64 oracle (GPU).

Inputs follow the reference's SourcePack layout (native/pws/io/sourcepack_io.cu:
25-column index, records [SAC header][npts float32]); pairs are grouped by
(path_id as text, component_slot), records of a pair sorted by timestamp
(SourcePackGroupReader, native/pws/io/sourcepack_io.cu).  Tolerances as in
tests/test_stack_gpu.py: relative L2 1e-5 (PWS), 1e-4 (TF-PWS).
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import stack
from oracle.xc import rel_l2
from tests import fxformats as ff

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PWS = os.path.join(ROOT, "bin", "ncf_pws")
TFPWS = os.path.join(ROOT, "bin", "ncf_tfpws")

pytestmark = pytest.mark.skipif(not (os.path.exists(PWS) and os.path.exists(TFPWS)),
                                reason="bin/ncf_pws, bin/ncf_tfpws not built")


def run(exe, args):
    return subprocess.run([exe] + args, capture_output=True, text=True, timeout=300)


# pair: (path_id text, component_slot, src, rec, comps, number of records)
PAIRS = [("10002", 0, 1, 2, ("BHZ", "BHZ"), 7),
         ("10002", 4, 1, 2, ("BHN", "BHE"), 5),
         ("20003", 0, 2, 3, ("BHZ", "BHZ"), 4),
         ("100004", 0, 1, 4, ("BHZ", "BHZ"), 9),   # sorts before "20003" as text
         ("30003", 0, 3, 3, ("BHZ", "BHZ"), 1)]    # a single record


def _trace(rng, n):
    t = np.arange(n)
    sig = np.cos(2 * np.pi * 0.07 * (t - n // 2)) * np.exp(-((t - n // 2) / (0.08 * n)) ** 2)
    return (sig + 0.6 * rng.standard_normal(n)).astype(np.float32)


def make_sourcepack(tmp, n=345, delta=0.1, seed=0):
    """Two SourcePack shards (each a pack + index sorted by path_id text, slot)
    holding the records of PAIRS spread over timestamps; the record order in
    the files is shuffled so the executable has to sort by timestamp."""
    rng = np.random.default_rng(seed)
    shards = {0: [], 1: []}
    truth = {}
    for pid, slot, s, r, (cs, cr), cnt in PAIRS:
        recs = []
        for k in range(cnt):
            ts = "2020.%03d.0000" % (100 + k)
            recs.append((ts, _trace(rng, n)))
            shards[k % 2].append((pid, slot, s, r, cs, cr, ts, recs[-1][1]))
        truth[(pid, slot)] = np.stack([x for _, x in sorted(recs, key=lambda z: z[0])])
    index_paths = []
    for w, rows in shards.items():
        order = rng.permutation(len(rows))
        rows = [rows[i] for i in order]
        pack = os.path.join(tmp, "in.w%03d.pack" % w)
        lines = ["\t".join(ff.SOURCEPACK_COLUMNS)]
        body = []
        with open(pack, "wb") as f:
            for pid, slot, s, r, cs, cr, ts, x in rows:
                off = f.tell()
                blob = ff.sac_bytes(x, delta, kstnm="S%02d" % r, knetwk="XX", kcmpnm=cr,
                                    b=-(n // 2) * delta)
                f.write(blob)
                fpp = "XX.S%02d-XX.S%02d/XX.S%02d-XX.S%02d.%s-%s.sac" % (s, r, s, r, cs, cr)
                body.append((pid, slot, [ts, pid, slot, "XX.S%02d" % s, "XX.S%02d" % r, s, r,
                                         "XX", "S%02d" % s, "", cs, "XX", "S%02d" % r, "", cr, n,
                                         "%g" % delta, 11.5 * r, 30.0 + s, 210.0 + r, pack, off,
                                         len(blob), "xc_pack", fpp]))
        body.sort(key=lambda z: (z[0], z[1]))      # stable: shard order within a group
        lines += ["\t".join(str(v) for v in row) for _, _, row in body]
        idx = os.path.join(tmp, "in.w%03d.tsv" % w)
        with open(idx, "w") as f:
            f.write("\n".join(lines) + "\n")
        index_paths.append(idx)
    lst = os.path.join(tmp, "sourcepack.list")
    with open(lst, "w") as f:
        f.write("# sourcepack indexes\n\n" + "\n".join(index_paths) + "\n")
    return lst, truth


@pytest.mark.parametrize("exe", [PWS, TFPWS])
def test_usage_and_required_arguments(tmp_path, exe):
    r = run(exe, [])
    assert r.returncode == 1 and "Usage" in r.stderr
    r = run(exe, ["--sourcepack-list", "x"])
    assert r.returncode == 1 and "event=missing_sourcepack_io" in r.stderr
    lst = tmp_path / "empty.list"
    lst.write_text("# nothing\n\n")
    r = run(exe, ["--sourcepack-list", str(lst), "--output-sourcepack", str(tmp_path / "o")])
    assert r.returncode == 1 and "event=sourcepack_list_empty" in r.stderr
    idx = tmp_path / "bad.tsv"
    idx.write_text("timestamp\tpath_id\n")
    lst.write_text(str(idx) + "\n")
    r = run(exe, ["--sourcepack-list", str(lst), "--output-sourcepack", str(tmp_path / "o")])
    assert r.returncode == 1 and "event=sourcepack_index_missing_field" in r.stderr


def test_tfpws_option_checks(tmp_path):
    base = ["-I", "x", "-O", str(tmp_path)]
    for sel in ("000", "0010", "012", "110"):
        r = run(TFPWS, base + ["-S", sel])
        assert r.returncode == 1 and "event=invalid_output_selection" in r.stderr, sel
    for band in ("0.3/0.1", "abc", "-1/2"):
        r = run(TFPWS, base + ["-F", band])
        assert r.returncode == 1 and "event=invalid_band" in r.stderr, band
    r = run(TFPWS, base + ["-T", "-1"])
    assert r.returncode == 1 and "event=invalid_band_taper" in r.stderr
    r = run(PWS, ["--sourcepack-list", "x", "--output-sourcepack", "y", "-H", "0"])
    assert r.returncode == 1 and "event=invalid_staged_group_limit" in r.stderr


def _check_outputs(out, kind, truth, n, delta, B, band=None, tol=1e-5):
    rows = ff.read_sourcepack_index(os.path.join(out, "sourcepack_index.tsv"))
    keys = [(row["path_id"], int(row["component_slot"])) for row, _, _ in rows]
    assert keys == sorted(truth), keys
    assert all(row["timestamp"] == "STACK" for row, _, _ in rows)
    for row, hd, data in rows:
        recs = truth[(row["path_id"], int(row["component_slot"]))]
        T = recs.shape[0]
        if kind == "pws":
            g, w, lin = stack.pws_prepare(recs, B)
            want = stack.pws(g, w, lin)
            G = len(w)
            assert (hd["user0"], hd["user1"], hd["user2"]) == (T, G, 1 if B < 2 else B)
            assert hd["kuser0"].startswith("NTRACE") and hd["kuser2"].startswith("SUBSTK")
            assert (hd["nzyear"], hd["nzjday"]) == (2010, 214)
            assert hd["depmax"] == data.max() and hd["depmin"] == data.min()
            assert row["storage_kind"] == "pws_pack"
            assert row["final_pair_path"].endswith(".pws.sac")
        else:
            s, w, lin = stack.tfpws_prepare(recs, B)
            want = stack.tfpws(s, w, lin, band=band, delta=delta)
            assert hd["user0"] == -12345.0 and hd["depmax"] == -12345.0   # header passed through
            assert row["storage_kind"] == "tfpws_pack"
            assert row["final_pair_path"].endswith(".tfpws.sac")
        assert hd["npts"] == n and int(row["npts"]) == n and row["dt"] == "%g" % np.float32(delta)
        assert np.isclose(hd["b"], -(n // 2) * np.float32(delta))
        scale = np.linalg.norm(lin)
        if np.linalg.norm(want) < 1e-9 * scale:
            # one group: every TF-PWS voice's time sum is removed by G(0, f) = 0,
            # so the stack is zero up to rounding
            err = np.linalg.norm(data) / scale
        else:
            err = rel_l2(data, want)
        print(f"{kind} pair={row['path_id']}/{row['component_slot']} T={T} rel_l2={err:.3e}",
              flush=True)
        assert err < tol
    with open(os.path.join(out, "_SUCCESS")) as f:
        succ = f.read().splitlines()
    assert succ[0] == "records\t%d" % len(truth) and succ[1] == "workers\t1"
    assert succ[2] == "index\t" + os.path.join(out, "sourcepack_index.tsv")
    assert succ[3] == "pack\t" + os.path.join(out, "%s.w000.pack" % kind)


@pytest.mark.gpu
@pytest.mark.parametrize("B,n", [(1, 345), (3, 401)])
def test_pws_file_to_file_matches_oracle(tmp_path, B, n):
    delta = 0.1
    lst, truth = make_sourcepack(str(tmp_path), n=n, delta=delta, seed=B)
    out = str(tmp_path / "pws")
    prog = str(tmp_path / "pws_progress.tsv")
    r = run(PWS, ["--sourcepack-list", lst, "--output-sourcepack", out, "-B", str(B),
                  "--progress", prog, "-H", "2"])
    assert r.returncode == 0, r.stderr
    _check_outputs(out, "pws", truth, n, delta, B)
    with open(prog) as f:
        lines = f.read().splitlines()
    assert lines[1].split("\t")[:5] == ["overall", "DONE", "5", "5", "groups"]


@pytest.mark.gpu
@pytest.mark.parametrize("B,band", [(2, None), (1, (0.5, 2.5, 0.3))])
def test_tfpws_file_to_file_matches_oracle(tmp_path, B, band):
    n, delta = 201, 0.1
    lst, truth = make_sourcepack(str(tmp_path), n=n, delta=delta, seed=10 + B)
    out = str(tmp_path / "tfpws")
    args = ["-I", lst, "-O", out, "-B", str(B), "-S", "001"]
    if band:
        args += ["-F", "%g/%g" % band[:2], "-T", "%g" % band[2]]
    r = run(TFPWS, args)
    assert r.returncode == 0, r.stderr
    _check_outputs(out, "tfpws", truth, n, delta, B, band=band, tol=1e-4)
