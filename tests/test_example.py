"""The reference's bundled example (a committed subset: two stations x E/N/Z
x two days at 1 Hz, tests/golden/example/) through the stage executables,
against the FP64 oracle's outputs written by tests/golden/make_example_golden.py
(golden.npz).  Settings are the example's own (example/config.ini):
RUN-ABS-MF over the bands 0.2/0.4 and 0.1/0.2 Hz, whitening after, one-hour
windows without overlap, max_lag 100 s (201 lags); stacks over the two days
(PWS / TF-PWS with sub-stacks of one record).  Tolerances: relative L2 1e-5
for the correlations and the linear stack, 1e-4 for PWS and TF-PWS."""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
EX = os.path.join(HERE, "golden", "example")
STATIONS = ["A7K2", "B4Q8"]
DAYS = ["138", "139"]
COMPS = ["E", "N", "Z"]


def _golden():
    g = np.load(os.path.join(EX, "golden.npz"))  # no pickled objects
    keys = [(int(a), int(b), c, d) for a, b, c, d in g["keys"]]
    return keys, g


def _sac(sta, day, comp):
    return os.path.join(EX, "data", sta, "2022", f"{sta}.2022.{day}.0000.{comp}.sac")


def test_golden_is_the_oracle_on_the_committed_inputs():
    """The fixture matches what the oracle computes from the committed SAC
    files and filter file (CPU)."""
    from tests.golden import make_example_golden as mk

    keys, g = _golden()
    plan, k2, days, lin, pws, tfpws = mk.oracle_outputs(os.path.join(EX, "data"),
                                                         os.path.join(EX, "filter.txt"))
    assert [tuple(map(str, k)) for k in k2] == [tuple(map(str, k)) for k in keys]
    assert np.allclose(days, g["xc"], rtol=0, atol=1e-12 * np.abs(g["xc"]).max())
    for name, v in (("linear", lin), ("pws", pws), ("tfpws", tfpws)):
        assert np.allclose(v, g[name], rtol=0, atol=1e-12 * np.abs(g[name]).max()), name
    assert list(g["plan"]) == [plan.npts, plan.seg, plan.nfft_filter, plan.nfft_out, plan.nspec,
                               plan.nstep]


def _run(exe, args):
    r = subprocess.run([os.path.join(ROOT, "bin", exe)] + args, capture_output=True, text=True,
                       timeout=600)
    if r.returncode != 0:
        print(r.stderr[-3000:])
    return r


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.gpu
def test_example_through_the_executables(gpu, tmp_path):
    from tests import fxformats as ff

    keys, g = _golden()
    index = tmp_path / "sac_index.tsv"
    with open(index, "w") as f:
        f.write("timestamp\tnsl_id\tnetwork\tstation\tlocation\tcomponent\tsac_path\n")
        for day in DAYS:
            for s, sta in enumerate(STATIONS, start=1):
                for c in COMPS:
                    f.write(f"2022.{day}.0000\t{s}\tKS\t{sta}\t00\t{c}\t{_sac(sta, day, c)}\n")
    ws = tmp_path / "ws"
    r = _run("sac2spec", ["-I", str(index), "-O", str(ws), "-C", "3", "-B",
                          os.path.join(EX, "filter.txt"), "-L", "86400/3600/3600/100",
                          "-W", "2/0", "-N", "1", "-Q", "-1", "-T", "4"])
    assert r.returncode == 0
    paths = str(tmp_path / "allowed_paths.tsv")
    ff.write_allowed_paths(paths, [(1, 2, 0.06, 6.6, 90.0, 270.0)])
    out, spd = tmp_path / "ncf", tmp_path / "sp"
    r = _run("xc_fast", ["-I", str(ws), "-P", paths, "-O", str(out), "-C", "100",
                         "--sourcepack-dir", str(spd)])
    assert r.returncode == 0
    recs = ff.read_xcpack_dir(str(out / "xcpack"))
    stamps = sorted({row["timestamp"] for row, _, _ in recs})
    assert len(stamps) == len(DAYS) and len(recs) == len(DAYS) * len(keys)
    key = lambda row: (int(row["src_id"]), int(row["rec_id"]), row["src_component"],
                       row["rec_component"])
    for d, ts in enumerate(stamps):
        got = {key(row): data for row, _, data in recs if row["timestamp"] == ts}
        err = _rel(np.stack([got[k] for k in keys]), g["xc"][d])
        print(f"example xc day {d}: rel_l2={err:.3e}")
        assert err < 1e-5
    idx = [p / "sourcepack_index.tsv" for p in sorted(spd.iterdir())
           if (p / "sourcepack_index.tsv").exists()]
    assert len(idx) == len(DAYS)
    from fastxc_amd.linear_stack import linear_stack_sourcepack_indexes
    linear_stack_sourcepack_indexes(idx, tmp_path / "stack", device=gpu)
    lin = ff.read_sourcepack_index(
        str(tmp_path / "stack" / "linearstack_sourcepack" / "STACK" / "sourcepack_index.tsv"))
    got = {key(row): data for row, _, data in lin}
    err = _rel(np.stack([got[k] for k in keys]), g["linear"])
    print(f"example linear: rel_l2={err:.3e}")
    assert err < 1e-5
    lst = tmp_path / "sourcepacks.txt"
    lst.write_text("\n".join(str(p) for p in idx) + "\n")
    for exe, name, extra in (("ncf_pws", "pws", []), ("ncf_tfpws", "tfpws", ["-S", "001"])):
        o = str(tmp_path / name)
        r = _run(exe, ["--sourcepack-list", str(lst), "--output-sourcepack", o, "-B", "1"] + extra)
        assert r.returncode == 0
        rows = ff.read_sourcepack_index(os.path.join(o, "sourcepack_index.tsv"))
        got = {key(row): data for row, _, data in rows}
        err = _rel(np.stack([got[k] for k in keys]), g[name])
        print(f"example {name}: rel_l2={err:.3e}")
        assert err < 1e-4
