"""sac2spec executable: command line and plan derivation (CPU), file-to-file
spectra and the sac2spec -> xc_fast chain against the oracle (GPU)."""
import os
import subprocess

import numpy as np
import pytest

from oracle import sac2spec as s2s
from oracle import xc
from tests import fxformats as ff

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
S2S = os.path.join(ROOT, "bin", "sac2spec")
XC = os.path.join(ROOT, "bin", "xc_fast")

pytestmark = pytest.mark.skipif(not os.path.exists(S2S), reason="bin/sac2spec not built")


def run(exe, args):
    return subprocess.run([exe] + args, capture_output=True, text=True, timeout=300)


def make_sac_inputs(tmp, nsta, npts, delta, nch=1, ts="2020.123.0400", seed=0, seconds=None):
    rng = np.random.default_rng(seed)
    comps = ["BHZ"] if nch == 1 else ["BHE", "BHN", "BHZ"]
    rows, traces = [], []
    os.makedirs(os.path.join(tmp, "sac"), exist_ok=True)
    for s in range(1, nsta + 1):
        for c in comps:
            n = npts if seconds is None else seconds
            x = rng.standard_normal(n).astype(np.float32)
            p = os.path.join(tmp, "sac", "S%02d.%s.sac" % (s, c))
            ff.write_sac(p, x, delta, stla=40.0 + s, stlo=120.0 - s, kstnm="S%02d" % s,
                         knetwk="XX", kcmpnm=c)
            rows.append((ts, str(s), "XX", "S%02d" % s, "--", c, p))
            traces.append(x)
    index = os.path.join(tmp, "sac_index.tsv")
    with open(index, "w") as f:
        f.write("timestamp\tnsl_id\tnetwork\tstation\tlocation\tcomponent\tsac_path\n")
        for r in rows:
            f.write("\t".join(r) + "\n")
    return index, rows, traces


def test_usage_and_argument_errors(tmp_path):
    assert run(S2S, []).returncode == 1
    assert run(S2S, ["--help"]).returncode == 0
    r = run(S2S, ["-I", "x", "-O", str(tmp_path)])
    assert r.returncode == 1 and "-L sac_len/win_len/shift_len/xcorr_lag is required" in r.stderr
    r = run(S2S, ["-I", "x", "-O", str(tmp_path), "-B", "f", "-L", "1/2/3"])
    assert r.returncode == 1 and "Invalid -L value" in r.stderr
    r = run(S2S, ["-I", "x", "-O", str(tmp_path), "-B", "f", "-L", "10/2/1/0", "-W", "5/0"])
    assert r.returncode == 1 and "whiten_type must be 0, 1, 2, or 3" in r.stderr
    r = run(S2S, ["-I", "x", "-O", str(tmp_path), "-B", "f", "-L", "10/2/1/0", "-Z", "1"])
    assert r.returncode == 1 and "Unknown option -Z" in r.stderr


@pytest.mark.parametrize("L,skip,bands", [("86400/3600/3600/100", None, "0.1/2"),
                                          ("86400/3600/1800/100", "3,7,-1", "0.1/2 0.5/1"),
                                          ("7200/600/300/50", "0,-1", "0.2/1.5")])
def test_dry_run_plan_matches_oracle(tmp_path, L, skip, bands):
    index, rows, _ = make_sac_inputs(str(tmp_path), 2, 100, 0.1)
    filt = str(tmp_path / "filter.txt")
    s2s.write_filter_file(filt, 0.1, bands)
    args = ["-I", index, "-O", str(tmp_path / "ws"), "-C", "1", "-B", filt, "-L", L, "--dry-run"]
    if skip:
        args += ["-Q", skip]
    r = run(S2S, args)
    assert r.returncode == 0, r.stderr
    got = dict(ln.split("\t", 1) for ln in r.stdout.strip().split("\n"))
    sac_len, win, shift, lag = (float(v) for v in L.split("/"))
    p = s2s.plan_lengths(sac_len, 0.1, win, shift, lag, s2s.read_filters(filt))
    assert int(got["npts"]) == p.npts and int(got["seg"]) == p.seg
    assert int(got["shift"]) == p.shift and int(got["nstep"]) == p.nstep
    assert int(got["nfft_filter"]) == p.nfft_filter and int(got["nfft_out"]) == p.nfft_out
    assert int(got["nspec"]) == p.nspec
    assert tuple(int(v) for v in got["idx"].split(",")) == p.idx
    skipped = [int(v) for v in skip.split(",")[:-1]] if skip else []
    assert [int(v) for v in got["valid_steps"].split(",")] == \
        [s for s in range(p.nstep) if s not in skipped]
    assert int(got["bands"]) == len(p.responses)
    for i, want in enumerate(p.responses):
        vals = np.array([float(v) for v in got["band%d" % i].split("\t")[1].split(",")])
        w = want[: p.nfft_filter // 2 + 1]
        assert np.max(np.abs(vals - w) / np.maximum(np.abs(w), 1e-30)) < 1e-6


def _oracle_spectra(traces, plan, nch, whiten, normalize, valid=None):
    x = np.stack(traces).astype(np.float64)
    win = s2s.cut_windows(x, plan, valid)                   # [steps, files, seg]
    steps, files, seg = win.shape
    fr = win.reshape(steps * files // nch, nch, seg)
    out = s2s.process_batch(fr, plan, whiten=whiten, normalize=normalize)
    return out.reshape(steps, files, -1)


@pytest.mark.gpu
@pytest.mark.parametrize("nch,whiten,normalize", [(1, 2, 3), (3, 3, 1), (1, 0, 2)])
def test_file_to_file_and_xc_chain(gpu, tmp_path, nch, whiten, normalize):
    nsta, npts, delta = 5, 4000, 0.1
    index, rows, traces = make_sac_inputs(str(tmp_path), nsta, npts, delta, nch=nch)
    filt = str(tmp_path / "filter.txt")
    s2s.write_filter_file(filt, delta, "0.2/2" if normalize != 1 else "0.2/2 0.5/1.5")
    ws = tmp_path / "ws"
    L = "400/100/100/20"
    r = run(S2S, ["-I", index, "-O", str(ws), "-C", str(nch), "-B", filt, "-L", L,
                  "-W", "%d/0" % whiten, "-N", str(normalize), "-T", "4"])
    print(r.stderr[-3000:])
    assert r.returncode == 0
    sp = ws / "stepack"
    assert (sp / "_SUCCESS").read_text() == "DONE\n"
    assert (sp / "2020.123.0400" / "_SUCCESS").read_text() == \
        "timestamp\t2020.123.0400\ngroups\t%d\n" % nsta
    prog = (ws / "progress" / "sac2spec_progress.tsv").read_text().split("\n")
    assert prog[1].startswith("overall\tDONE\t%d\t%d\tsac_rows" % (nsta * nch, nsta * nch))
    packs = sorted(p for p in os.listdir(sp) if p.endswith(".stepack"))
    assert packs == ["w000.b000000.stepack"]
    hdr, nodes, spec = ff.read_stepack(str(sp / packs[0]))
    plan = s2s.plan_lengths(400, delta, 100, 100, 20, s2s.read_filters(filt))
    assert (hdr["nstep"], hdr["nspec"], hdr["nslc_count"]) == (plan.nstep, plan.nspec, nsta * nch)
    assert hdr["dt"] == pytest.approx(delta) and hdr["df"] == pytest.approx(1 / (plan.nfft_out * delta))
    assert [n["gnsl_id"] for n in nodes] == [int(r_[1]) for r_ in rows]
    assert nodes[0]["stla"] == pytest.approx(41.0) and nodes[0]["location"] == "--"
    want = _oracle_spectra(traces, plan, nch, whiten, normalize)
    err = xc.rel_l2(spec, want)
    print(f"sac2spec nch={nch} whiten={whiten} normalize={normalize} rel_l2={err:.3e}")
    assert err < 1e-5
    # the chain: xc_fast on these StepPacks against the oracle on the oracle spectra
    recs = [(a, b) for a in range(1, nsta + 1) for b in range(a, nsta + 1)]
    paths = str(tmp_path / "allowed_paths.tsv")
    ff.write_allowed_paths(paths, [(a, b, 0.1, 10.0, 1.0, 2.0) for a, b in recs])
    out = tmp_path / "ncf"
    r = run(XC, ["-I", str(ws), "-P", paths, "-O", str(out), "-C", "20"])
    print(r.stderr[-2000:])
    assert r.returncode == 0
    got = ff.read_xcpack_dir(str(out / "xcpack"))
    gnsl = [n["gnsl_id"] for n in nodes]
    job = xc.build_jobs(len(gnsl), len(gnsl))[0]
    tasks, _ = xc.build_tasks_for_job(gnsl, {a * 10000 + b for a, b in recs}, job)
    shape = xc.finalize_shape(plan.nspec, plan.nstep, delta, 20.0)
    ref = xc.correlate(want, np.array([t.src_meta for t in tasks]),
                       np.array([t.rec_meta for t in tasks]), shape.nfft, shape.half_cc, delta)
    cc = np.stack([d for _, _, d in got])
    err = xc.rel_l2(cc, ref)
    print(f"chain rel_l2={err:.3e}")
    assert err < 1e-5


@pytest.mark.gpu
def test_sac_lengths_and_delta_tolerance(gpu, tmp_path):
    """read_sac_buffer (native/sac2spec/io/sacio.c:158-211): a file longer than
    npts = llround(sac_len/dt) is truncated, a shorter one zero-padded; the
    delta check (io/parallel_io.c:12-73) accepts |delta - dt| <= max(1e-5 dt,
    1e-6) and rejects the run otherwise."""
    delta, npts = 0.1, 4000
    rng = np.random.default_rng(21)
    # shorter files end on a window boundary (1000 samples): a window that
    # is part data, part zero padding would hand RUN-ABS the prefilter's
    # decaying tail, which it scales up to O(1) from values near FP32
    # rounding, so no FP32 path can match FP64 there
    lens = [npts, npts + 517, npts - 2000, 1000, npts + 1]
    deltas = [delta, delta * (1 + 8e-6), delta, delta * (1 - 9e-6), delta]
    os.makedirs(tmp_path / "sac", exist_ok=True)
    rows, traces = [], []
    for s, (n, d) in enumerate(zip(lens, deltas), start=1):
        x = rng.standard_normal(n).astype(np.float32)
        p = str(tmp_path / "sac" / ("S%02d.BHZ.sac" % s))
        ff.write_sac(p, x, d, stla=40.0 + s, stlo=120.0 - s, kstnm="S%02d" % s, knetwk="XX",
                     kcmpnm="BHZ")
        rows.append(("2020.123.0400", str(s), "XX", "S%02d" % s, "--", "BHZ", p))
        t = np.zeros(npts, np.float32)
        t[: min(n, npts)] = x[:npts]
        traces.append(t)

    def write_index(path, rws):
        with open(path, "w") as f:
            f.write("timestamp\tnsl_id\tnetwork\tstation\tlocation\tcomponent\tsac_path\n")
            for r_ in rws:
                f.write("\t".join(r_) + "\n")

    index = str(tmp_path / "sac_index.tsv")
    write_index(index, rows)
    filt = str(tmp_path / "filter.txt")
    s2s.write_filter_file(filt, delta, "0.2/2")
    L = "400/100/100/20"
    ws = tmp_path / "ws"
    r = run(S2S, ["-I", index, "-O", str(ws), "-C", "1", "-B", filt, "-L", L, "-W", "2/0",
                  "-N", "3", "-T", "2"])
    print(r.stderr[-3000:])
    assert r.returncode == 0
    assert "truncated" in r.stderr and "zero_padded" in r.stderr
    sp = ws / "stepack"
    packs = sorted(p for p in os.listdir(sp) if p.endswith(".stepack"))
    _, _, spec = ff.read_stepack(str(sp / packs[0]))
    plan = s2s.plan_lengths(400, delta, 100, 100, 20, s2s.read_filters(filt))
    assert plan.npts == npts
    want = _oracle_spectra(traces, plan, 1, 2, 3)
    err = xc.rel_l2(spec, want)
    print(f"sac lengths rel_l2={err:.3e}")
    assert err < 1e-5
    # one file 2e-5 off in delta: the run fails and names the file
    bad = str(tmp_path / "sac" / "BAD.BHZ.sac")
    ff.write_sac(bad, rng.standard_normal(npts).astype(np.float32), delta * (1 + 2e-5),
                 kstnm="BAD", knetwk="XX", kcmpnm="BHZ")
    index2 = str(tmp_path / "sac_index2.tsv")
    write_index(index2, rows[:2] + [("2020.123.0400", "9", "XX", "BAD", "--", "BHZ", bad)])
    r = run(S2S, ["-I", index2, "-O", str(tmp_path / "ws2"), "-C", "1", "-B", filt, "-L", L,
                  "-W", "2/0", "-N", "3", "-T", "2"])
    assert r.returncode != 0
    assert "sac_delta_mismatch" in r.stderr and "BAD.BHZ.sac" in r.stderr
