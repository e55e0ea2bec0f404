"""Station inventory path plan, the all-pairs loop native.

Host-side mirror of fastxc/inventory/planner.py (the reference runs all of it
in Python): the same entry points ``build_path_plan`` (:104-137),
``filter_group_by_path_plan`` (:315-321), ``write_path_plan`` (:324-395),
``write_timestamp_manifests`` (:398-478) and their record types, files and
errors.  ``build_valid_paths`` -- every node pair's Rudoe geometry and the
distance / azimuth / group / autocorrelation filters, N(N+1)/2 evaluations --
runs in ``fxc_plan_paths`` (include/fastxc_amd.h), threaded host C++ with
the reference's float64 libm arithmetic, so distances, azimuths and the
retained set equal the reference's bit for bit.

Behaviour kept from the reference:
  - nodes are unique (group, network, station, location) keys sorted and
    numbered from 1 (at most 9999); the representative file is the smallest
    resolved path; station lat/lon/elevation come from that SAC header
    (endianness scored on nvhdr and plausible lat/lon; -12345 and
    out-of-range values unknown), then an external geo TSV (network +
    station + location rows, else station-only rows) overrides them
    (:140-291, :515-548, :938-1000);
  - path ids are src_id * 10000 + rec_id; pairs run src <= rec in id order;
    ranges "low/high" with low < 0 and high >= 50000 (360) switch a filter
    off (:481-512, :882-910);
  - allowed_paths.tsv / nsl_catalog.tsv / allowed_path_ids.txt and the
    per-timestamp manifests are written with the same columns, "%.7g"
    numbers and order.
"""
from __future__ import annotations

import csv
import ctypes
import math
import struct
from collections import OrderedDict
from dataclasses import dataclass, field
from pathlib import Path
from typing import Dict, Iterable, Sequence

import numpy as np

from .geodesy import distkm_az_baz_rudoe  # noqa: F401  (the reference exports it here)

GNSLKey = tuple
_SAC_UNDEFINED = -12345.0


@dataclass
class GeoNode:
    gnsl_id: int
    group: str
    network: str
    station: str
    location: str
    representative_file_path: str
    components: set = field(default_factory=set)
    lat: float | None = None
    lon: float | None = None
    ele: float | None = None
    sac_geo_parse_ok: bool = True

    @property
    def gnsl(self) -> GNSLKey:
        return (self.group, self.network, self.station, self.location)

    @property
    def id_text(self) -> str:
        return f"{self.gnsl_id:04d}"


@dataclass(frozen=True)
class PathRecord:
    path_id: int
    src_gnsl_id: int
    rec_gnsl_id: int
    src_gnsl: GNSLKey
    rec_gnsl: GNSLKey
    great_circle_deg: float | None
    distance_km: float | None
    azimuth_deg: float | None
    back_azimuth_deg: float | None

    @property
    def path_id_text(self) -> str:
        return f"{self.path_id:08d}"


@dataclass(frozen=True)
class PairFilterConfig:
    allow_autocorr: bool = True
    autocorr_only: bool = False
    group_pair_mode: str = "all"
    use_distance_filter: bool = False
    min_distance_km: float = 0.0
    max_distance_km: float = 50000.0
    use_azimuth_filter: bool = False
    azimuth_ranges_deg: tuple = ((0.0, 360.0),)


@dataclass
class PathPlan:
    nodes: list
    paths: list
    pair_checks_total: int
    missing_geo_node_count: int
    sac_header_parse_error_count: int
    external_geo_row_count: int
    external_geo_node_fill_count: int
    external_geo_conflict_node_count: int
    config: PairFilterConfig

    @property
    def node_by_gnsl(self) -> dict:
        return {n.gnsl: n for n in self.nodes}

    @property
    def retained_gnsls(self) -> set:
        if isinstance(self.paths, PathList):
            return {self.nodes[int(k)].gnsl for k in self.paths.node_indices()}
        return {p.src_gnsl for p in self.paths} | {p.rec_gnsl for p in self.paths}

    @property
    def allowed_path_ids(self) -> set:
        if isinstance(self.paths, PathList):
            return set(self.paths.path_ids().tolist())
        return {p.path_id for p in self.paths}


# ------------------------------------------------------------------ nodes --
def _group_order(name: str):
    return (int(name), name) if name.isdigit() else (999, name)


def _groups(files_groups: dict) -> OrderedDict:
    return OrderedDict((str(g), fg or {}) for g, fg in
                       sorted(files_groups.items(), key=lambda kv: _group_order(str(kv[0]))))


def _first(values, default: str) -> str:
    if isinstance(values, (list, tuple)) and values:
        return str(values[0])
    if values not in (None, ""):
        return str(values)
    return default


def _gnsl_of(group: str, key, info: dict) -> GNSLKey:
    if isinstance(key, tuple):
        station = str(key[0])
        network = str(key[2]) if len(key) >= 3 else _first(info.get("network"), "VV")
        location = str(key[3]) if len(key) >= 4 else _first(info.get("location"), "00")
    else:
        station = str(key)
        network = _first(info.get("network"), "VV")
        location = _first(info.get("location"), "00")
    return (group, network, station, location)


def collect_geo_nodes_from_groups(*, files_groups: dict) -> list:
    found: dict = {}
    for group, files_group in _groups(files_groups).items():
        for key, info in files_group.items():
            gnsl = _gnsl_of(group, key, info)
            paths = [Path(p).expanduser().resolve().as_posix() for p in info.get("path", []) or []]
            if not paths:
                continue
            rep = min(paths)
            node = found.get(gnsl)
            if node is None:
                node = found[gnsl] = GeoNode(0, *gnsl, representative_file_path=rep)
            else:
                node.representative_file_path = min(node.representative_file_path, rep)
            node.components.update(str(c) for c in info.get("component", []))
    nodes = sorted(found.values(), key=lambda n: n.gnsl)
    if len(nodes) > 9999:
        raise ValueError("GNSL node count exceeds 9999; 4-digit IDs are not enough.")
    for k, node in enumerate(nodes, start=1):
        node.gnsl_id = k
    return nodes


@dataclass(frozen=True)
class SacStationGeoResult:
    lat: float | None
    lon: float | None
    ele: float | None
    parse_ok: bool
    error_message: str | None = None


def _undefined(v: float) -> bool:
    return abs(v - _SAC_UNDEFINED) <= 1e-3


def _plausible(v: float, kind: str) -> bool:
    if _undefined(v):
        return True
    if not math.isfinite(v):
        return False
    lim = {"lat": 90.0, "lon": 180.0}.get(kind)
    return lim is None or -lim <= v <= lim


def _geo_value(v: float, kind: str) -> float | None:
    if _undefined(v) or not math.isfinite(v):
        return None
    lim = {"lat": 90.0, "lon": 180.0}.get(kind)
    if lim is not None and not -lim <= v <= lim:
        return None
    return float(v)


def _sac_endian(raw: bytes) -> str | None:
    """The byte order with the best score: nvhdr in 1..20 (2), lat and lon
    plausible (1 each); None when neither scores."""
    best, score = None, -1
    for e in "<>":
        nvhdr = struct.unpack_from(e + "i", raw, 280 + 24)[0]
        s = (2 if 0 < nvhdr <= 20 else 0) + \
            _plausible(struct.unpack_from(e + "f", raw, 124)[0], "lat") + \
            _plausible(struct.unpack_from(e + "f", raw, 128)[0], "lon")
        if s > score:
            best, score = e, s
    return best if score > 0 else None


def read_station_geo_from_sac_header(file_path: str) -> SacStationGeoResult:
    path = Path(file_path).expanduser().resolve()
    if not path.is_file():
        return SacStationGeoResult(None, None, None, False, f"file not found: {path}")
    with path.open("rb") as fh:
        raw = fh.read(632)
    if len(raw) < 632:
        return SacStationGeoResult(None, None, None, False,
                                   "file too small to contain a complete SAC header")
    e = _sac_endian(raw)
    if e is None:
        return SacStationGeoResult(None, None, None, False,
                                   "unable to determine SAC header endianness")
    lat, lon, ele = struct.unpack_from(e + "3f", raw, 124)
    return SacStationGeoResult(_geo_value(lat, "lat"), _geo_value(lon, "lon"),
                               _geo_value(ele, "ele"), True)


def enrich_nodes_from_sac_headers(nodes: Iterable[GeoNode]) -> None:
    for node in nodes:
        geo = read_station_geo_from_sac_header(node.representative_file_path)
        node.lat, node.lon, node.ele, node.sac_geo_parse_ok = geo.lat, geo.lon, geo.ele, geo.parse_ok


def _cell_float(row: dict, names, row_number: int, label: str | None):
    for name in names:
        v = row.get(name)
        if v is not None and str(v).strip() != "":
            return float(v)
    if label is None:
        return None
    raise ValueError(f"external geo row {row_number}: invalid {label!r}")


def load_external_geo(external_geo_tsv_path):
    if external_geo_tsv_path is None or str(external_geo_tsv_path).strip().upper() == "NONE":
        return {}, {}, 0
    path = Path(external_geo_tsv_path).expanduser().resolve()
    if not path.is_file():
        raise FileNotFoundError(f"external geo TSV file not found: {path}")
    by_nsl, by_station, count = {}, {}, 0
    with path.open("r", encoding="utf-8", newline="") as fh:
        reader = csv.DictReader(fh, delimiter="\t")
        if reader.fieldnames is None:
            return {}, {}, 0
        for n, row in enumerate(reader, start=2):
            station = (row.get("station") or "").strip()
            if not station:
                raise ValueError(f"external geo row {n}: missing 'station'")
            geo = {"lat": _cell_float(row, ("lat", "latitude"), n, "lat/latitude"),
                   "lon": _cell_float(row, ("lon", "longitude"), n, "lon/longitude"),
                   "ele": _cell_float(row, ("ele", "elevation"), n, None)}
            count += 1
            network = (row.get("network") or "").strip()
            location = (row.get("location") or "").strip()
            if network and location:
                key = (network, station, location)
                if key in by_nsl:
                    raise ValueError(
                        f"duplicate network/station/location external geo row for {key!r}")
                by_nsl[key] = geo
            else:
                if station in by_station:
                    raise ValueError(f"duplicate station-only external geo row for {station!r}")
                by_station[station] = geo
    return by_nsl, by_station, count


def inject_external_geo(nodes: Iterable[GeoNode], external_geo_tsv_path) -> dict:
    by_nsl, by_station, count = load_external_geo(external_geo_tsv_path)
    stats = {"external_geo_row_count": count, "external_geo_node_fill_count": 0,
             "external_geo_conflict_node_count": 0}
    if count == 0:
        return stats
    for node in nodes:
        geo = by_nsl.get((node.network, node.station, node.location)) or by_station.get(node.station)
        if geo is None:
            continue
        updated = conflicted = False
        for name in ("lat", "lon", "ele"):
            new = geo.get(name)
            if new is None:
                continue
            old = getattr(node, name)
            if old is not None and abs(old - new) > 1e-6:
                conflicted = True
            if old is None or abs(old - new) > 1e-6:
                setattr(node, name, new)
                updated = True
        stats["external_geo_node_fill_count"] += updated
        stats["external_geo_conflict_node_count"] += conflicted
    return stats


# ----------------------------------------------------------------- config --
def _range_pair(value: str, name: str):
    try:
        lo_text, hi_text = str(value).strip().split("/", 1)
        lo, hi = float(lo_text), float(hi_text)
    except Exception as exc:
        raise ValueError(f"{name} must be 'low/high', got {value!r}") from exc
    if hi < lo and lo >= 0.0:
        raise ValueError(f"{name}: lower {lo} > upper {hi}")
    return lo, hi


def _group_mode(mode, double_array: bool) -> str:
    v = (mode or "auto").strip().lower()
    if v == "auto":
        return "cross_group_only" if double_array else "all"
    for name, words in (("all", {"all", "both"}),
                        ("same_group_only", {"intra", "within", "same", "same_group_only"}),
                        ("cross_group_only", {"inter", "between", "cross", "cross_group_only"})):
        if v in words:
            return name
    raise ValueError(f"unsupported group_pair_mode: {mode!r}")


def _autocorr_mode(mode, allow: bool) -> str:
    if mode is None:
        return "include" if allow else "off"
    v = str(mode).strip().lower().replace("-", "_")
    for name, words in (("off", {"", "off", "none", "false", "no", "0"}),
                        ("include", {"include", "on", "true", "yes", "1"}),
                        ("only", {"only", "auto_only", "autocorr_only", "self", "self_only"})):
        if v in words:
            return name
    raise ValueError(f"unsupported autocorr_mode: {mode!r}")


def parse_pair_filter_config(*, distance_range: str, azimuth_range: str, double_array: bool,
                             allow_autocorr: bool, autocorr_mode=None,
                             group_pair_mode=None) -> PairFilterConfig:
    dlo, dhi = _range_pair(distance_range, "distance_range")
    alo, ahi = _range_pair(azimuth_range, "azimuth_range")
    auto = _autocorr_mode(autocorr_mode, allow_autocorr)
    use_d = not (dlo < 0.0 and dhi >= 50000.0)
    use_a = not (alo < 0.0 and ahi >= 360.0)
    return PairFilterConfig(
        allow_autocorr=auto != "off", autocorr_only=auto == "only",
        group_pair_mode=_group_mode(group_pair_mode, double_array),
        use_distance_filter=use_d, min_distance_km=max(0.0, dlo) if use_d else 0.0,
        max_distance_km=dhi if use_d else 50000.0, use_azimuth_filter=use_a,
        azimuth_ranges_deg=((max(0.0, alo), min(360.0, ahi)),) if use_a else ((0.0, 360.0),))


def azimuth_in_ranges(azimuth_deg: float, ranges_deg) -> bool:
    az = azimuth_deg % 360.0
    for lo, hi in ranges_deg:
        if lo == 0.0 and hi == 360.0:
            return True
        a, b = lo % 360.0, hi % 360.0
        if (a <= az <= b) if a <= b else (az >= a or az <= b):
            return True
    return False


# ----------------------------------------------------------------- paths --
_GROUP_MODES = {"all": 0, "same_group_only": 1, "cross_group_only": 2}


class PathList(Sequence):
    """The retained paths of ``fxc_plan_paths``: a read-only sequence of
    PathRecord built on access from the native arrays (src / rec node index,
    geometry), so millions of pairs cost no Python objects until they are
    read; ``write_path_plan`` writes the set natively."""

    def __init__(self, nodes: list, handle, count: int):
        from ._lib import check, lib

        self._nodes = nodes
        self._handle = handle
        self._lib = lib()
        self.src = np.empty(count, np.int32)
        self.rec = np.empty(count, np.int32)
        self.has_geo = np.empty(count, np.uint8)
        self.geo = np.empty((count, 4), np.float64)   # gcarc, az, baz, dist
        if count:
            check(self._lib.fxc_plan_paths_fetch(handle, self.src.ctypes.data,
                                                 self.rec.ctypes.data, self.has_geo.ctypes.data,
                                                 self.geo.ctypes.data), "fxc_plan_paths_fetch")

    def __len__(self) -> int:
        return len(self.src)

    def _record(self, k: int) -> PathRecord:
        a, b = self._nodes[int(self.src[k])], self._nodes[int(self.rec[k])]
        if self.has_geo[k]:
            gcarc, az, baz, dist = (float(v) for v in self.geo[k])
        else:
            gcarc = az = baz = dist = None
        return PathRecord(a.gnsl_id * 10000 + b.gnsl_id, a.gnsl_id, b.gnsl_id, a.gnsl, b.gnsl,
                          gcarc, dist, az, baz)

    def __getitem__(self, k):
        if isinstance(k, slice):
            return [self._record(i) for i in range(*k.indices(len(self)))]
        if k < 0:
            k += len(self)
        if not 0 <= k < len(self):
            raise IndexError(k)
        return self._record(k)

    def __iter__(self):
        for k in range(len(self)):
            yield self._record(k)

    def path_ids(self) -> np.ndarray:
        ids = np.array([n.gnsl_id for n in self._nodes], np.int64)
        return ids[self.src] * 10000 + ids[self.rec]

    def node_indices(self) -> np.ndarray:
        return np.union1d(self.src, self.rec)

    def write(self, tsv_path: Path, ids_path: Path, threads: int = 0) -> None:
        from ._lib import check

        n = len(self._nodes)
        ids = np.array([nd.gnsl_id for nd in self._nodes], np.int32)
        text = (ctypes.c_char_p * max(n, 1))(*[".".join(nd.gnsl).encode() for nd in self._nodes])
        check(self._lib.fxc_plan_paths_write(self._handle, n, ids.ctypes.data,
                                             ctypes.cast(text, ctypes.c_void_p),
                                             str(tsv_path).encode(), str(ids_path).encode(),
                                             int(threads)), "fxc_plan_paths_write")

    def __del__(self):
        if getattr(self, "_handle", None) is not None:
            self._lib.fxc_plan_paths_free(self._handle)
            self._handle = None


def build_valid_paths(nodes: list, config: PairFilterConfig, *, threads: int = 0):
    """(paths in (src, rec) id order, pairs examined) through fxc_plan_paths."""
    from ._lib import check, lib

    if config.group_pair_mode not in _GROUP_MODES:
        raise ValueError(f"unsupported group_pair_mode: {config.group_pair_mode!r}")
    n = len(nodes)
    has = np.array([nd.lat is not None and nd.lon is not None for nd in nodes], np.uint8)
    lat = np.array([nd.lat if h else 0.0 for nd, h in zip(nodes, has)], np.float64)
    lon = np.array([nd.lon if h else 0.0 for nd, h in zip(nodes, has)], np.float64)
    names: dict = {}
    grp = np.array([names.setdefault(nd.group, len(names)) for nd in nodes], np.int32)
    ranges = np.ascontiguousarray(np.asarray(config.azimuth_ranges_deg, np.float64).reshape(-1))
    handle, npaths, nchecks = ctypes.c_void_p(), ctypes.c_longlong(), ctypes.c_longlong()
    check(lib().fxc_plan_paths(n, lat.ctypes.data, lon.ctypes.data, has.ctypes.data,
                               grp.ctypes.data, int(config.allow_autocorr),
                               int(config.autocorr_only), _GROUP_MODES[config.group_pair_mode],
                               int(config.use_distance_filter), float(config.min_distance_km),
                               float(config.max_distance_km), int(config.use_azimuth_filter),
                               ranges.ctypes.data, len(ranges) // 2, int(threads), handle,
                               npaths, nchecks), "fxc_plan_paths")
    return PathList(nodes, handle, npaths.value), nchecks.value


def build_path_plan(*, files_groups: dict, distance_range: str = "-1/50000",
                    azimuth_range: str = "-1/360", double_array: bool = False,
                    allow_autocorr: bool = True, autocorr_mode=None, group_pair_mode=None,
                    external_geo_tsv_path=None, threads: int = 0) -> PathPlan:
    """Unique GNSL nodes and the retained canonical paths between them."""
    nodes = collect_geo_nodes_from_groups(files_groups=files_groups)
    enrich_nodes_from_sac_headers(nodes)
    stats = inject_external_geo(nodes, external_geo_tsv_path)
    config = parse_pair_filter_config(distance_range=distance_range, azimuth_range=azimuth_range,
                                      double_array=double_array, allow_autocorr=allow_autocorr,
                                      autocorr_mode=autocorr_mode,
                                      group_pair_mode=group_pair_mode)
    paths, checks = build_valid_paths(nodes, config, threads=threads)
    return PathPlan(nodes=nodes, paths=paths, pair_checks_total=checks,
                    missing_geo_node_count=sum(n.lat is None or n.lon is None for n in nodes),
                    sac_header_parse_error_count=sum(not n.sac_geo_parse_ok for n in nodes),
                    **stats, config=config)


def filter_group_by_path_plan(files_group: Dict, group_name: str, plan: PathPlan) -> Dict:
    keep = plan.retained_gnsls
    return {k: v for k, v in files_group.items() if _gnsl_of(group_name, k, v) in keep}


# ---------------------------------------------------------------- writing --
def _cell(v) -> str:
    return "" if v is None else f"{v:.7g}"


def write_path_plan(plan: PathPlan, output_dir) -> None:
    out = Path(output_dir).expanduser().resolve()
    out.mkdir(parents=True, exist_ok=True)
    with (out / "nsl_catalog.tsv").open("w", newline="", encoding="utf-8") as fh:
        w = csv.writer(fh, delimiter="\t")
        w.writerow(["nsl_id", "group", "network", "station", "location", "lat", "lon", "ele",
                    "components", "representative_file_path"])
        for n in plan.nodes:
            w.writerow([n.id_text, n.group, n.network, n.station, n.location, _cell(n.lat),
                        _cell(n.lon), _cell(n.ele), ",".join(sorted(n.components)),
                        Path(n.representative_file_path).as_posix()])
    if isinstance(plan.paths, PathList):
        plan.paths.write(out / "allowed_paths.tsv", out / "allowed_path_ids.txt")
        return
    with (out / "allowed_paths.tsv").open("w", newline="", encoding="utf-8") as fh:
        w = csv.writer(fh, delimiter="\t")
        w.writerow(["path_id", "src_nsl_id", "rec_nsl_id", "src_nsl", "rec_nsl",
                    "great_circle_deg", "distance_km", "azimuth_deg", "back_azimuth_deg"])
        for p in plan.paths:
            w.writerow([p.path_id_text, f"{p.src_gnsl_id:04d}", f"{p.rec_gnsl_id:04d}",
                        ".".join(p.src_gnsl), ".".join(p.rec_gnsl), _cell(p.great_circle_deg),
                        _cell(p.distance_km), _cell(p.azimuth_deg), _cell(p.back_azimuth_deg)])
    (out / "allowed_path_ids.txt").write_text("".join(f"{p.path_id_text}\n" for p in plan.paths),
                                              encoding="utf-8")


def _component_rank(c: str) -> int:
    return {"E": 0, "N": 1, "Z": 2}.get(c, 99)


def write_timestamp_manifests(files_group: Dict, group_name: str, plan: PathPlan,
                              output_dir) -> Path:
    """Per-timestamp SAC manifests of the retained nodes, their index and list."""
    root = Path(output_dir).expanduser().resolve()
    shard_root = root / "by_timestamp"
    shard_root.mkdir(parents=True, exist_ok=True)
    by_node, keep = plan.node_by_gnsl, plan.retained_gnsls
    rows_by_ts: OrderedDict = OrderedDict()
    for key, info in files_group.items():
        ts = key[1] if isinstance(key, tuple) and len(key) >= 2 else None
        if ts is None:
            continue
        gnsl = _gnsl_of(group_name, key, info)
        node = by_node.get(gnsl)
        if node is None or gnsl not in keep:
            continue
        text = ts.strftime("%Y.%j.%H%M") if hasattr(ts, "strftime") else str(ts)
        rows = rows_by_ts.setdefault(text, [])
        for comp, sac in zip(info.get("component", []), info.get("path", [])):
            rows.append([node.id_text, node.group, node.network, node.station, node.location,
                         str(comp), Path(str(sac)).expanduser().resolve().as_posix()])
    index_rows, shards = [], []
    for text, rows in rows_by_ts.items():
        rows.sort(key=lambda r: (r[0], _component_rank(r[5]), r[5], r[6]))
        shard = shard_root / f"{text}.tsv"
        with shard.open("w", newline="", encoding="utf-8") as fh:
            w = csv.writer(fh, delimiter="\t")
            w.writerow(["nsl_id", "group", "network", "station", "location", "component",
                        "sac_path"])
            w.writerows(rows)
        shards.append(shard)
        index_rows.append([text, shard.relative_to(root).as_posix(), str(len(rows))])
    index = root / "timestamp_index.tsv"
    with index.open("w", newline="", encoding="utf-8") as fh:
        w = csv.writer(fh, delimiter="\t")
        w.writerow(["timestamp", "shard_path", "row_count"])
        w.writerows(index_rows)
    (root / "timestamp_manifest_list.txt").write_text(
        "".join(f"{p.as_posix()}\n" for p in shards), encoding="utf-8")
    return index
