"""xc_fast executable: command line, StepPack input, task order (CPU) and the
full file-to-file run against the oracle (GPU)."""
import os
import subprocess

import numpy as np
import pytest

from oracle import xc
from tests import fxformats as ff

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
XC = os.path.join(ROOT, "bin", "xc_fast")

pytestmark = pytest.mark.skipif(not os.path.exists(XC), reason="bin/xc_fast not built")


def run(args, **kw):
    return subprocess.run([XC] + args, capture_output=True, text=True, timeout=300, **kw)


def make_inputs(tmp, nsta=6, nch=1, nstep=3, nspec=883, ts="2020.123.0400", seed=0,
                drop=((1, 4),), shuffle_batches=True):
    """Two StepPack batches with the stations split across them (batch 1 holds
    the lower ids to exercise the merge + stable sort), and an allowed-path
    table with every pair except `drop`."""
    rng = np.random.default_rng(seed)
    comps = ["BHE", "BHN", "BHZ"][-nch:] if nch > 1 else ["BHZ"]
    ids = list(range(1, nsta + 1))
    nodes_all = []
    for sid in ids:
        for c in comps:
            nodes_all.append(dict(timestamp=ts, nsl_id="XX.S%02d." % sid, gnsl_id=sid,
                                  stla=30.0 + sid, stlo=100.0 - sid, network="XX",
                                  station="S%02d" % sid, location="", component=c))
    spec = (rng.standard_normal((nstep, len(nodes_all), nspec)) +
            1j * rng.standard_normal((nstep, len(nodes_all), nspec))).astype(np.complex64)
    half = (nsta // 2) * nch
    parts = [(nodes_all[half:], spec[:, half:]), (nodes_all[:half], spec[:, :half])]
    if not shuffle_batches:
        parts = parts[::-1]
    sp_dir = os.path.join(tmp, "stepack")
    for b, (nd, sp) in enumerate(parts):
        ff.write_stepack(sp_dir, 0, b, nd, sp, 0.1, 1.0 / (2 * (nspec - 1) * 0.1),
                         start_group=0, group_count=len(nd) // nch)
    recs = []
    for a in ids:
        for b in ids:
            if a <= b and (a, b) not in drop:
                recs.append((a, b, 0.01 * (a + b), 1.1 * (b - a), 10.0 * a, 20.0 * b))
    paths = os.path.join(tmp, "allowed_paths.tsv")
    ff.write_allowed_paths(paths, recs)
    allowed = {a * 10000 + b for a, b, *_ in recs}
    return sp_dir, paths, nodes_all, spec, allowed


def test_usage_and_required_arguments(tmp_path):
    r = run([])
    assert r.returncode == 1 and "Usage" in r.stderr
    r = run(["--help"])
    assert r.returncode == 0
    r = run(["-O", str(tmp_path), "-C", "10"])
    assert r.returncode == 1 and "event=required_argument_missing" in r.stderr
    r = run(["-O", str(tmp_path), "-C", "10", "-P", "x"])
    assert r.returncode == 1 and "event=timestamp_input_missing" in r.stderr
    r = run(["-I", "x", "-O", str(tmp_path), "-C", "10", "-P", "x", "--bogus"])
    assert r.returncode == 1 and "event=argument_unknown" in r.stderr


def test_precision_environment_default(tmp_path):
    env = dict(os.environ, FASTXC_XC_PRECISION="bogus")
    r = run(["-I", "x", "-P", "y", "-O", str(tmp_path), "-C", "10"], env=env)
    assert r.returncode == 1 and 'option="--precision" value="bogus"' in r.stderr
    # the command line wins over the environment
    sp_dir, paths, *_ = make_inputs(str(tmp_path), nsta=3)
    r = run(["-I", sp_dir, "-P", paths, "-O", str(tmp_path / "o"), "-C", "5", "--dry-run",
             "--precision", "bf16x3"], env=env)
    assert r.returncode == 0, r.stderr


@pytest.mark.parametrize("nch,block", [(1, 0), (1, 3), (3, 0), (3, 5)])
def test_dry_run_task_order_matches_oracle(tmp_path, nch, block):
    sp_dir, paths, nodes, spec, allowed = make_inputs(str(tmp_path), nsta=7, nch=nch)
    args = ["-I", sp_dir, "-P", paths, "-O", str(tmp_path / "out"), "-C", "20", "--dry-run"]
    if block:
        args += ["--block-files", str(block)]
    r = run(args)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().split("\n")[1:]
    got = [tuple(int(v) for v in ln.split("\t")[1:]) for ln in lines]
    # rows sorted by gnsl id (stable): the oracle sees them in id order
    gnsl = sorted(n["gnsl_id"] for n in nodes)
    rows = len(gnsl)
    bs = min(rows, 4096) if not block else min(rows, block)
    want = []
    for job in xc.build_jobs(rows, bs):
        tasks, _ = xc.build_tasks_for_job(gnsl, allowed, job)
        want += [(job.anchor_block, job.target_begin_block, t.src_meta, t.rec_meta, t.src_local,
                  t.rec_local, t.path_id, int(t.is_autocorr)) for t in tasks]
    assert got == want


def test_stepack_tsv_merge_sorts_rows_by_id(tmp_path):
    # the second batch holds the lower ids; the merged rows come out sorted
    sp_dir, paths, nodes, spec, allowed = make_inputs(str(tmp_path), nsta=6)
    r = run(["-I", sp_dir, "-P", paths, "-O", str(tmp_path / "o"), "-C", "20", "--dry-run"])
    assert r.returncode == 0
    rows = [ln.split("\t") for ln in r.stdout.strip().split("\n")[1:]]
    assert rows[0][0] == "20200502T04:00"  # 2020.123 -> May 2
    srcs = {int(x[3]) for x in rows}
    assert min(srcs) == 0


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [("fp32", 1e-5), ("bf16x3", 1e-3)])
def test_file_to_file_matches_oracle(gpu, tmp_path, precision, tol):
    nsta, nstep, nspec, cclen, dt = 6, 3, 883, 30.0, 0.1
    sp_dir, paths, nodes, spec, allowed = make_inputs(str(tmp_path), nsta=nsta, nstep=nstep,
                                                      nspec=nspec)
    out = tmp_path / "ncf"
    prog = tmp_path / "progress.tsv"
    r = run(["-I", sp_dir, "-P", paths, "-O", str(out), "-C", str(cclen), "--progress", str(prog),
             "--precision", precision])
    print(r.stderr[-2000:])
    assert r.returncode == 0
    xdir = out / "xcpack"
    assert (xdir / "_SUCCESS").read_text() == "DONE\n"
    assert (xdir / "20200502T04:00.done").read_text() == "DONE\n"
    assert "overall\tDONE\t1\t1\ttimestamps" in prog.read_text()
    recs = ff.read_xcpack_dir(str(xdir))
    # rows in id order: spectra re-ordered accordingly (nodes are already in id order)
    order = sorted(range(len(nodes)), key=lambda i: nodes[i]["gnsl_id"])
    S = spec[:, order]
    shape = xc.finalize_shape(nspec, nstep, dt, cclen)
    H = shape.half_cc
    job = xc.build_jobs(len(order), len(order))[0]
    tasks, _ = xc.build_tasks_for_job([nodes[i]["gnsl_id"] for i in order], allowed, job)
    assert len(recs) == len(tasks)
    src = np.array([t.src_meta for t in tasks])
    rec = np.array([t.rec_meta for t in tasks])
    want = xc.correlate(S, src, rec, shape.nfft, H, dt)
    got = np.stack([d for _, _, d in recs])
    assert got.shape == (len(tasks), 2 * H + 1)
    err = xc.rel_l2(got, want)
    print(f"precision={precision} rel_l2={err:.3e}")
    assert err < tol
    for (row, hd, _), t in zip(recs, tasks):
        s, q = nodes[order[t.src_meta]], nodes[order[t.rec_meta]]
        assert int(row["src_id"]) == s["gnsl_id"] and int(row["rec_id"]) == q["gnsl_id"]
        assert int(row["npts"]) == 2 * H + 1 and float(row["dt"]) == pytest.approx(dt)
        assert row["final_pair_path"] == "%s/XX.%s/XX.%s/XX-XX.%s-%s.BHZ-BHZ.sac" % (
            out, s["station"], q["station"], s["station"], q["station"])
        assert hd["npts"] == 2 * H + 1 and hd["delta"] == pytest.approx(dt)
        assert hd["b"] == pytest.approx(-H * dt) and hd["e"] == pytest.approx(H * dt)
        assert hd["user0"] == s["gnsl_id"] and hd["user1"] == q["gnsl_id"]
        assert hd["user2"] == t.path_id
        assert hd["kcmpnm"].rstrip() == "BHZ-BHZ" and hd["kevnm"].rstrip() == "XX." + s["station"]
        assert hd["kstnm"].rstrip() == q["station"]
        assert hd["stla"] == pytest.approx(q["stla"]) and hd["evla"] == pytest.approx(s["stla"])
        assert (hd["nzyear"], hd["nzjday"], hd["nzhour"], hd["nzmin"]) == (2020, 123, 4, 0)
        assert hd["nvhdr"] == 6 and hd["iftype"] == 1 and hd["leven"] == 1


@pytest.mark.gpu
def test_sourcepack_dir_indexes_each_timestamp(gpu, tmp_path):
    """--sourcepack-dir: the index built beside the run equals the restated
    reference builder over the finished xcpack output, byte for byte, and the
    global _SUCCESS / progress sidecar follow the materializer's finish()."""
    from oracle import sourcepack as osp

    sp_dir, paths, nodes, spec, allowed = make_inputs(str(tmp_path), nsta=5, nch=3, nstep=2,
                                                      nspec=257)
    out = tmp_path / "ncf"
    spdir = tmp_path / "sourcepack"
    prog = tmp_path / "progress" / "sourcepack_progress.tsv"
    r = run(["-I", sp_dir, "-P", paths, "-O", str(out), "-C", "10", "--sourcepack-dir",
             str(spdir), "--sourcepack-progress", str(prog)])
    print(r.stderr[-2000:])
    assert r.returncode == 0
    tot = osp.build_all(str(out / "xcpack"), str(tmp_path / "oracle"))
    from tests.test_host_sourcepack import tree
    got, want = tree(spdir), tree(tmp_path / "oracle")
    assert got.pop("_SUCCESS") == ("timestamps\t1\nrecords\t%d\nsources\t%d\nbytes_indexed\t%d\n"
                                   % (tot["records"], tot["sources"], tot["bytes"])).encode()
    assert got == want
    assert tot["records"] == len(ff.read_xcpack_dir(str(out / "xcpack"))) > 0
    assert "overall\tDONE\t1\t" in prog.read_text()
