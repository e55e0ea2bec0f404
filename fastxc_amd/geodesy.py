"""Epicentral geometry of a station pair on the WGS84 ellipsoid.

``distkm_az_baz_rudoe`` restates fastxc/inventory/planner.py:833-967 (itself
the geometry of the reference's xc C code, the SAC distaz method): great-
circle arc and azimuths from geocentric-latitude direction cosines, distance
from Rudoe's normal-section series.  The operations run in the same order
in float64, so the values written into SAC headers round to the same floats.
Host code: it runs once per pair, before the rotation kernel.
"""
from __future__ import annotations

import math

_A_KM = 6378.137                      # WGS84 equatorial radius
_F = 1.0 / 298.257223563              # WGS84 flattening
_EC2 = 2.0 * _F - _F * _F             # first eccentricity squared
_ONE_MINUS_EC2 = 1.0 - _EC2
_FLT_MAX = 3.4028234663852886e38      # C FLT_MAX, the tangent taken at the poles
# e^2 coefficient of the sin 2u term of the distance series as the reference
# has it (planner.py:950); the textbook value is 1/32.  With 13/32 the
# distance runs up to ~3e-5 relative above the geodesic; it is kept so that
# header and index values equal the reference's.
_C2_E2 = 13.0 / 32.0
_D2R = math.pi / 180.0
_R2D = 180.0 / math.pi


def _cosines(lat_deg: float, lon_deg: float):
    """(geodetic latitude in radians, longitude in radians, and the eight
    direction-cosine terms of the geocentric point used by the arc and azimuth
    formulas)."""
    lat_rad = _D2R * (lat_deg if lat_deg != 0.0 else 1.0e-10)
    lon_rad = _D2R * lon_deg
    if lat_deg in (90.0, -90.0):
        geoc = lat_deg * _D2R
    else:
        geoc = math.atan(_ONE_MINUS_EC2 * math.tan(lat_rad))
    sin_lon, mcos_lon = math.sin(lon_rad), -math.cos(lon_rad)
    mcos_geoc, sin_geoc = -math.cos(geoc), math.sin(geoc)
    # x, y, z of the point and the two horizontal unit directions
    return lat_rad, lon_rad, (mcos_geoc * mcos_lon, -mcos_geoc * sin_lon, sin_geoc,
                              sin_lon, mcos_lon, mcos_geoc, -sin_geoc * mcos_lon,
                              sin_geoc * sin_lon)


def _azimuth(p, q) -> float:
    """Azimuth (degrees, [0, 360)) at point p towards point q."""
    x1, y1, z1 = q[0], q[1], q[2]
    ss = (x1 - p[3]) ** 2 + (y1 - p[4]) ** 2 + z1 ** 2 - 2.0
    sc = (x1 - p[6]) ** 2 + (y1 - p[7]) ** 2 + (z1 - p[5]) ** 2 - 2.0
    az = math.atan2(ss, sc) * _R2D
    if az < 0.0 and abs(az) < 1.0e-8:
        az = 0.0
    if az < 0.0:
        az += 360.0
    return az


def _pole_trig(lat_deg: float, lat_rad: float):
    """(cos, sin, tan) of the latitude, with the poles' tangent at +-FLT_MAX."""
    if lat_deg == 90.0:
        return 0.0, 1.0, _FLT_MAX
    if lat_deg == -90.0:
        return 0.0, -1.0, -_FLT_MAX
    c, s = math.cos(lat_rad), math.sin(lat_rad)
    return c, s, s / c


def _rudoe_km(lat_k: float, rad_k: float, lon_k: float, lat_i: float, rad_i: float,
              lon_i: float) -> float:
    """Rudoe's normal-section length (km) from point k to point i."""
    cos_k, sin_k, tan_k = _pole_trig(lat_k, rad_k)
    cos_i, sin_i, tan_i = _pole_trig(lat_i, rad_i)
    e_prime = _EC2 / _ONE_MINUS_EC2
    one_e = 1.0 + e_prime
    al = tan_i / (one_e * tan_k) + _EC2 * math.sqrt((one_e + tan_i ** 2) / (one_e + tan_k ** 2))
    dl = lon_i - lon_k
    a12 = math.atan2(math.sin(dl), (al - math.cos(dl)) * sin_k)
    cos_a12, sin_a12 = math.cos(a12), math.sin(a12)
    e1 = e_prime * ((cos_k * cos_a12) ** 2 + sin_k ** 2)
    e2 = e1 * e1
    e3 = e1 * e2
    c0 = 1.0 + 0.25 * e1 - (3.0 / 64.0) * e2 + (5.0 / 256.0) * e3
    c2 = -(1.0 / 8.0) * e1 + _C2_E2 * e2 - (15.0 / 1024.0) * e3
    c4 = -(1.0 / 256.0) * e2 + (3.0 / 1024.0) * e3
    nu_k = _A_KM / math.sqrt(1.0 - _EC2 * sin_k ** 2)
    nu_i = _A_KM / math.sqrt(1.0 - _EC2 * sin_i ** 2)
    z_k = nu_k * _ONE_MINUS_EC2 * sin_k
    z_i = nu_i * _ONE_MINUS_EC2 * sin_i
    x_i = nu_i * cos_i * math.cos(dl)
    y_i = nu_i * cos_i * math.sin(dl)
    e1p = e1 + 1.0
    root = math.sqrt(e1p)
    u1 = math.atan2(tan_k, root * cos_a12)
    u2 = math.atan2(nu_k * sin_k + e1p * (z_i - z_k),
                    root * (x_i * cos_a12 - y_i * sin_k * sin_a12))
    b0 = nu_k * math.sqrt(1.0 + e_prime * (cos_k * cos_a12) ** 2) / e1p
    du = u2 - u1
    if abs(du) > math.pi:
        du = (math.pi * 2.0 - abs(du)) if du > 0.0 else (abs(du) - math.pi * 2.0)
    series = b0 * (c2 * (math.sin(2.0 * u2) - math.sin(2.0 * u1)) +
                   c4 * (math.sin(4.0 * u2) - math.sin(4.0 * u1)))
    return abs(b0 * c0 * du + series)


def distkm_az_baz_rudoe(evlo: float, evla: float, stlo: float,
                        stla: float) -> tuple[float, float, float, float]:
    """(great-circle arc in degrees, azimuth at the event towards the station,
    back azimuth at the station towards the event, distance in km)."""
    if evlo == stlo and evla == stla:
        return 0.0, 0.0, 0.0, 0.0
    ev_rad, ev_lon, ev = _cosines(evla, evlo)
    st_rad, st_lon, st = _cosines(stla, stlo)
    sc = ev[0] * st[0] + ev[1] * st[1] + ev[2] * st[2]
    sd = 0.5 * math.sqrt(((ev[0] - st[0]) ** 2 + (ev[1] - st[1]) ** 2 + (ev[2] - st[2]) ** 2) *
                         ((ev[0] + st[0]) ** 2 + (ev[1] + st[1]) ** 2 + (ev[2] + st[2]) ** 2))
    gcarc = math.atan2(sd, sc) * _R2D
    if gcarc < 0.0:
        gcarc += 360.0
    az = _azimuth(ev, st)
    baz = _azimuth(st, ev)
    # the series runs from the point of the positive (else the first) latitude
    if st_rad > 0.0:
        dist = _rudoe_km(evla, ev_rad, ev_lon, stla, st_rad, st_lon)
    else:
        dist = _rudoe_km(stla, st_rad, st_lon, evla, ev_rad, ev_lon)
    return gcarc, az, baz, dist
