"""Station-pair planning restated as plain Python loops -- test
infrastructure only.

  valid_paths   build_valid_paths + _build_path_record_if_valid +
                azimuth_in_ranges (fastxc/inventory/planner.py:294-312,
                564-617, 988-1002): candidates src <= rec in id order (src only
                with autocorr_only, src < rec without autocorrelation), the
                group rule for non-self pairs, pairs without coordinates kept
                only when no distance or azimuth filter is on, distance inside
                [min, max], azimuth or back azimuth inside a range (wrapping
                ranges allowed).  The geometry function is passed in.

Pinning (tests/test_planner.py): the reference's own planner cases
(fastxc/inventory/tests/test_path_planner.py: ids, autocorrelation modes,
group modes, distance filter, external geometry), restated as tests of both
this loop and the native planner; hand-worked azimuth-range cases
(wrapping, boundaries, the full circle).
"""
from __future__ import annotations


def in_ranges(az, ranges):
    az = az % 360.0
    for lo, hi in ranges:
        if lo == 0.0 and hi == 360.0:
            return True
        a, b = lo % 360.0, hi % 360.0
        if a <= b:
            if a <= az <= b:
                return True
        elif az >= a or az <= b:
            return True
    return False


def valid_paths(nodes, cfg, geometry_fn):
    """nodes: [(lat|None, lon|None, group)] in id order; cfg: dict with
    allow_autocorr, autocorr_only, group_mode ('all' | 'same_group_only' |
    'cross_group_only'), use_distance, min_km, max_km, use_azimuth, ranges.
    Returns ([(src index, rec index, (gcarc, az, baz, dist) | None)], checks)."""
    out, checks = [], 0
    n = len(nodes)
    for i in range(n):
        if cfg["autocorr_only"]:
            js = [i]
        else:
            js = range(i if cfg["allow_autocorr"] else i + 1, n)
        for j in js:
            checks += 1
            self_pair = i == j
            if self_pair and not cfg["allow_autocorr"]:
                continue
            if cfg["autocorr_only"] and not self_pair:
                continue
            gi, gj = nodes[i][2], nodes[j][2]
            if not self_pair:
                if cfg["group_mode"] == "same_group_only" and gi != gj:
                    continue
                if cfg["group_mode"] == "cross_group_only" and gi == gj:
                    continue
            geo = None
            if None in (nodes[i][0], nodes[i][1], nodes[j][0], nodes[j][1]):
                if cfg["use_distance"] or cfg["use_azimuth"]:
                    continue
            else:
                geo = geometry_fn(nodes[i][1], nodes[i][0], nodes[j][1], nodes[j][0])
            if cfg["use_distance"]:
                if geo is None or not cfg["min_km"] <= geo[3] <= cfg["max_km"]:
                    continue
            if cfg["use_azimuth"]:
                if geo is None or not (in_ranges(geo[1], cfg["ranges"]) or
                                       in_ranges(geo[2], cfg["ranges"])):
                    continue
            out.append((i, j, geo))
    return out, checks
