"""Inventory path planning: fastxc_amd/planner.py and its native all-pairs
loop (fxc_plan_paths, host code in libfastxc_amd.so) against the reference's
own planner cases and the plain-loop restatement in oracle/planner.py.

All CPU: the planner makes no GPU call.  The native geometry is compared bit
for bit with fastxc_amd.geodesy (pinned in tests/test_rotation.py against
independent formulas)."""
import os
import struct
from datetime import datetime

import numpy as np
import pytest

from oracle import planner as opl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
pytestmark = pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "fastxc_amd",
                                                                "libfastxc_amd.so")),
                                reason="libfastxc_amd.so not built")
TS = datetime(2023, 1, 1)


def _sac(path, lat, lon, ele=1.0, endian="<"):
    raw = bytearray(632)
    struct.pack_into(endian + "i", raw, 280 + 24, 6)
    struct.pack_into(endian + "3f", raw, 124, lat, lon, ele)
    path.write_bytes(bytes(raw))


def _info(tmp, station, lat, lon, network="XX"):
    paths = []
    for c in "ENZ":
        p = tmp / f"{station}.{c}.sac"
        _sac(p, lat, lon)
        paths.append(str(p))
    return {"network": [network] * 3, "location": ["00"] * 3, "component": list("ENZ"),
            "path": paths}


# ------------------------------------- the reference's own planner cases --
def test_ids_and_distance_filter(tmp_path):
    from fastxc_amd.planner import build_path_plan, filter_group_by_path_plan

    g = {("STA01", TS): _info(tmp_path, "STA01", 0.0, 0.0),
         ("STA02", TS): _info(tmp_path, "STA02", 0.0, 0.1),
         ("STA03", TS): _info(tmp_path, "STA03", 10.0, 10.0)}
    plan = build_path_plan(files_groups={"1": g}, distance_range="0/20",
                           azimuth_range="-1/360", allow_autocorr=False)
    assert [n.id_text for n in plan.nodes] == ["0001", "0002", "0003"]
    assert [p.path_id_text for p in plan.paths] == ["00010002"]
    assert plan.allowed_path_ids == {10002}
    assert {k[0] for k in filter_group_by_path_plan(g, "1", plan)} == {"STA01", "STA02"}


def test_single_station_autocorr_by_default(tmp_path):
    from fastxc_amd.planner import build_path_plan

    plan = build_path_plan(files_groups={"1": {("STA01", TS): _info(tmp_path, "STA01", 0, 0)}})
    assert [p.path_id_text for p in plan.paths] == ["00010001"]
    p = plan.paths[0]
    assert (p.distance_km, p.azimuth_deg, p.back_azimuth_deg) == (0.0, 0.0, 0.0)


def test_writes_path_plan_files(tmp_path):
    from fastxc_amd.planner import build_path_plan, write_path_plan

    g = {("STA01", TS): _info(tmp_path, "STA01", 0.0, 0.0),
         ("STA02", TS): _info(tmp_path, "STA02", 0.0, 0.1)}
    plan = build_path_plan(files_groups={"1": g}, allow_autocorr=False)
    out = tmp_path / "path_plan"
    write_path_plan(plan, out)
    lines = (out / "allowed_paths.tsv").read_text().splitlines()
    assert "great_circle_deg" in lines[0] and "distance_km" in lines[0]
    assert "\t\t" not in lines[1]
    f = lines[1].split("\t")
    assert f[:5] == ["00010002", "0001", "0002", "1.XX.STA01.00", "1.XX.STA02.00"]
    assert float(f[6]) == pytest.approx(11.13, abs=0.01)          # 0.1 degree at the equator
    assert (out / "allowed_path_ids.txt").read_text().strip() == "00010002"
    cat = (out / "nsl_catalog.tsv").read_text().splitlines()
    assert cat[1].split("\t")[:9] == ["0001", "1", "XX", "STA01", "00", "0", "0", "1", "E,N,Z"]


def test_external_geo_overrides_headers(tmp_path):
    from fastxc_amd.planner import build_path_plan

    g = {("STA01", TS): _info(tmp_path, "STA01", 0.0, 0.0),
         ("STA02", TS): _info(tmp_path, "STA02", 30.0, 30.0)}
    for header in ("station\tlat\tlon\tele", "station\tlatitude\tlongitude\televation"):
        ext = tmp_path / "geo.tsv"
        ext.write_text(header + "\nSTA01\t0.0\t0.0\t1.0\nSTA02\t0.0\t0.1\t2.0\n")
        plan = build_path_plan(files_groups={"1": g}, distance_range="0/20",
                               azimuth_range="-1/360", allow_autocorr=False,
                               external_geo_tsv_path=ext)
        assert (plan.external_geo_row_count, plan.external_geo_node_fill_count,
                plan.external_geo_conflict_node_count) == (2, 1, 1)
        assert [p.path_id_text for p in plan.paths] == ["00010002"]
        assert plan.nodes[1].lat == pytest.approx(0.0) and plan.nodes[1].lon == pytest.approx(0.1)
        assert plan.nodes[1].ele == pytest.approx(2.0)


def test_external_geo_matches_network_station_location(tmp_path):
    from fastxc_amd.planner import build_path_plan

    groups = {"1": {("STA01", TS): _info(tmp_path, "STA01", 0.0, 0.0, network="AA")},
              "2": {("STA01", TS): _info(tmp_path, "STA01", 30.0, 30.0, network="BB")}}
    ext = tmp_path / "geo.tsv"
    ext.write_text("station\tlat\tlon\tnetwork\tlocation\tgroup\n"
                   "STA01\t0.0\t0.1\tAA\t00\t9\nSTA01\t0.0\t0.2\tBB\t00\t9\n")
    plan = build_path_plan(files_groups=groups, distance_range="0/30", azimuth_range="-1/360",
                           allow_autocorr=False, group_pair_mode="inter",
                           external_geo_tsv_path=ext)
    by_net = {n.network: n for n in plan.nodes}
    assert by_net["AA"].lon == pytest.approx(0.1) and by_net["BB"].lon == pytest.approx(0.2)
    assert [p.path_id_text for p in plan.paths] == ["00010002"]
    ext.write_text("station\tlat\tlon\nSTA01\t0\t0\nSTA01\t1\t1\n")
    with pytest.raises(ValueError, match="duplicate station-only"):
        build_path_plan(files_groups=groups, external_geo_tsv_path=ext)


def test_timestamp_manifest(tmp_path):
    from fastxc_amd.planner import build_path_plan, write_timestamp_manifests

    g = {("STA01", TS): _info(tmp_path, "STA01", 0.0, 0.0),
         ("STA02", TS): _info(tmp_path, "STA02", 0.0, 0.1)}
    plan = build_path_plan(files_groups={"1": g}, allow_autocorr=False)
    root = tmp_path / "manifest" / "seisarray1"
    write_timestamp_manifests(g, "1", plan, root)
    assert "2023.001.0000\tby_timestamp/2023.001.0000.tsv\t6" in \
        (root / "timestamp_index.tsv").read_text()
    lst = (root / "timestamp_manifest_list.txt").read_text().splitlines()
    assert lst == [str((root / "by_timestamp" / "2023.001.0000.tsv").resolve())]
    lines = (root / "by_timestamp" / "2023.001.0000.tsv").read_text().splitlines()
    assert lines[0] == "nsl_id\tgroup\tnetwork\tstation\tlocation\tcomponent\tsac_path"
    assert len(lines) == 7 and [l.split("\t")[5] for l in lines[1:4]] == ["E", "N", "Z"]


def test_group_and_autocorr_modes(tmp_path):
    from fastxc_amd.planner import build_path_plan

    g1 = {("STA01", TS, "XX", "00"): _info(tmp_path, "STA01", 0.0, 0.0),
          ("STA02", TS, "XX", "00"): _info(tmp_path, "STA02", 0.0, 0.1)}
    g2 = {("STA03", TS, "YY", "00"): _info(tmp_path, "STA03", 0.0, 0.2)}
    counts = [len(build_path_plan(files_groups={"1": g1, "2": g2}, allow_autocorr=False,
                                  group_pair_mode=m).paths) for m in ("intra", "inter", "all")]
    assert counts == [1, 2, 3]
    ids = {m: [p.path_id_text for p in build_path_plan(files_groups={"1": g1},
                                                        autocorr_mode=m).paths]
           for m in ("off", "include", "only")}
    assert ids == {"off": ["00010002"], "include": ["00010001", "00010002", "00020002"],
                   "only": ["00010001", "00020002"]}
    with pytest.raises(ValueError, match="unsupported group_pair_mode"):
        build_path_plan(files_groups={"1": g1}, group_pair_mode="sideways")
    with pytest.raises(ValueError, match="distance_range"):
        build_path_plan(files_groups={"1": g1}, distance_range="30/10")


def test_sac_geo_endianness_and_undefined(tmp_path):
    from fastxc_amd.planner import read_station_geo_from_sac_header

    p = tmp_path / "be.sac"
    _sac(p, 45.5, -120.25, 3.0, endian=">")
    g = read_station_geo_from_sac_header(str(p))
    assert (g.lat, g.lon, g.ele, g.parse_ok) == (45.5, -120.25, 3.0, True)
    _sac(p, -12345.0, 200.0, 1.0)
    g = read_station_geo_from_sac_header(str(p))
    assert (g.lat, g.lon, g.parse_ok) == (None, None, True)
    p.write_bytes(b"\0" * 100)
    assert not read_station_geo_from_sac_header(str(p)).parse_ok


# ---------------------------------------------- native loop vs the oracle --
def _nodes(rng, n, ngroups=2, missing=0.0):
    nodes = []
    for k in range(n):
        lat = float(np.float32(rng.uniform(-60, 60)))
        lon = float(np.float32(rng.uniform(-180, 180)))
        if k % 17 == 3:
            lat = 0.0                                         # the series' 1e-10 latitude
        if k % 23 == 5:
            lat = 90.0                                        # a pole
        if rng.random() < missing:
            lat = lon = None
        nodes.append((lat, lon, f"g{rng.integers(ngroups)}"))
    if n > 4:
        nodes[4] = (nodes[2][0], nodes[2][1], nodes[4][2])   # co-located stations
    return nodes


CONFIGS = [
    dict(allow_autocorr=True, autocorr_only=False, group_mode="all", use_distance=False,
         min_km=0.0, max_km=50000.0, use_azimuth=False, ranges=((0.0, 360.0),)),
    dict(allow_autocorr=False, autocorr_only=False, group_mode="cross_group_only",
         use_distance=True, min_km=500.0, max_km=6000.0, use_azimuth=False,
         ranges=((0.0, 360.0),)),
    dict(allow_autocorr=True, autocorr_only=False, group_mode="same_group_only",
         use_distance=False, min_km=0.0, max_km=50000.0, use_azimuth=True,
         ranges=((300.0, 45.0),)),
    dict(allow_autocorr=True, autocorr_only=True, group_mode="all", use_distance=False,
         min_km=0.0, max_km=50000.0, use_azimuth=False, ranges=((0.0, 360.0),)),
    dict(allow_autocorr=False, autocorr_only=False, group_mode="all", use_distance=True,
         min_km=0.0, max_km=3000.0, use_azimuth=True, ranges=((90.0, 180.0),)),
]


def _native(nodes, cfg, threads=0):
    from fastxc_amd.planner import GeoNode, PairFilterConfig, build_valid_paths

    geo = [GeoNode(k + 1, g, "XX", f"S{k:04d}", "00", "x", lat=la, lon=lo)
           for k, (la, lo, g) in enumerate(nodes)]
    pc = PairFilterConfig(allow_autocorr=cfg["allow_autocorr"],
                          autocorr_only=cfg["autocorr_only"], group_pair_mode=cfg["group_mode"],
                          use_distance_filter=cfg["use_distance"],
                          min_distance_km=cfg["min_km"], max_distance_km=cfg["max_km"],
                          use_azimuth_filter=cfg["use_azimuth"],
                          azimuth_ranges_deg=cfg["ranges"])
    paths, checks = build_valid_paths(geo, pc, threads=threads)
    return [(p.src_gnsl_id - 1, p.rec_gnsl_id - 1,
             None if p.distance_km is None else
             (p.great_circle_deg, p.azimuth_deg, p.back_azimuth_deg, p.distance_km))
            for p in paths], checks


@pytest.mark.parametrize("k", range(len(CONFIGS)))
def test_native_loop_equals_oracle_bit_for_bit(k):
    from fastxc_amd.geodesy import distkm_az_baz_rudoe

    rng = np.random.default_rng(100 + k)
    nodes = _nodes(rng, 90, missing=0.08 if k == 0 else 0.0)
    want, wchecks = opl.valid_paths(nodes, CONFIGS[k], distkm_az_baz_rudoe)
    got, checks = _native(nodes, CONFIGS[k])
    assert checks == wchecks
    assert len(got) == len(want) > 0
    assert got == want          # indices, order and every float64 value


def test_thread_count_does_not_change_the_result():
    rng = np.random.default_rng(7)
    nodes = _nodes(rng, 300, ngroups=3)
    a = _native(nodes, CONFIGS[1], threads=1)
    b = _native(nodes, CONFIGS[1], threads=7)
    assert a == b


def test_azimuth_ranges_hand_worked():
    from fastxc_amd.planner import azimuth_in_ranges

    assert azimuth_in_ranges(10.0, ((0.0, 360.0),))
    assert azimuth_in_ranges(-30.0, ((300.0, 340.0),))          # -30 is 330
    assert azimuth_in_ranges(359.0, ((300.0, 45.0),)) and azimuth_in_ranges(45.0, ((300.0, 45.0),))
    assert not azimuth_in_ranges(46.0, ((300.0, 45.0),))
    assert azimuth_in_ranges(0.0, ((0.0, 0.0),)) and not azimuth_in_ranges(0.5, ((0.0, 0.0),))
    for az, r in ((-30.0, ((300.0, 340.0),)), (46.0, ((300.0, 45.0),)), (720.0, ((0.0, 1.0),))):
        assert azimuth_in_ranges(az, r) == opl.in_ranges(az, r)


def test_native_argument_errors():
    from fastxc_amd._lib import FxcError
    from fastxc_amd.planner import GeoNode, PairFilterConfig, build_valid_paths

    nodes = [GeoNode(k + 1, "1", "XX", f"S{k}", "00", "x", lat=0.0, lon=0.0) for k in range(10000)]
    with pytest.raises(FxcError, match="9999"):
        build_valid_paths(nodes, PairFilterConfig())
    paths, checks = build_valid_paths([], PairFilterConfig())
    assert len(paths) == 0 and checks == 0


def test_native_writer_equals_python_writer(tmp_path):
    """allowed_paths.tsv / allowed_path_ids.txt from the native set equal the
    csv.writer output of the same records, byte for byte (quoting, "%.7g",
    empty cells without coordinates, "\\r\\n")."""
    from fastxc_amd.planner import (GeoNode, PairFilterConfig, PathPlan, build_valid_paths,
                                    write_path_plan)

    rng = np.random.default_rng(3)
    nodes = []
    for k in range(120):
        lat, lon = float(rng.uniform(-80, 80)), float(rng.uniform(-180, 180))
        if k % 11 == 0:
            lat = lon = None
        sta = 'S"%03d' % k if k == 7 else ("S\t%03d" % k if k == 9 else "S%03d" % k)
        nodes.append(GeoNode(k + 1, str(k % 3), "XX", sta, "00", "x", lat=lat, lon=lon))
    paths, checks = build_valid_paths(nodes, PairFilterConfig())
    plan = PathPlan(nodes, paths, checks, 0, 0, 0, 0, 0, PairFilterConfig())
    write_path_plan(plan, tmp_path / "native")
    plan_py = PathPlan(nodes, list(paths), checks, 0, 0, 0, 0, 0, PairFilterConfig())
    write_path_plan(plan_py, tmp_path / "python")
    for name in ("allowed_paths.tsv", "allowed_path_ids.txt", "nsl_catalog.tsv"):
        assert (tmp_path / "native" / name).read_bytes() == (tmp_path / "python" / name).read_bytes()
    assert plan.allowed_path_ids == plan_py.allowed_path_ids
    assert plan.retained_gnsls == plan_py.retained_gnsls
    assert paths[-1] == list(paths)[-1] and paths[2:5] == list(paths)[2:5]

