"""Seeded synthetic station data in the shapes of the measured workloads.

One common source wavefield recorded by every station: white Gaussian noise
band-limited to [f_lo, f_hi] with an FFT mask, delayed per station by its
distance from a seeded source point at 3 km/s (the fractional delay applied as
a phase ramp, so the delay wraps around the record), plus independent white
noise at a signal-to-noise power ratio `snr`.  Stations sit uniformly in a
box_km x box_km square; the source point is drawn uniformly in the square
three times as wide around it.  With three components the wavefield is
projected onto E and N by a random back-azimuth per station (Z keeps it
whole); every component gets its own noise.  numpy's default generator
(PCG64) with a fixed seed makes every run identical.
"""
from __future__ import annotations

import os

import numpy as np

COMPONENTS = {1: ("BHZ",), 3: ("BHE", "BHN", "BHZ")}


def wavefield(nsta: int, npts: int, dt: float, seed: int, ncomp: int = 1, snr: float = 0.5,
              band=(0.1, 2.0), box_km: float = 100.0, v_kms: float = 3.0):
    """Traces [nsta, ncomp, npts] float32, station positions [nsta, 2] km and
    the delays [nsta] s."""
    rng = np.random.default_rng(seed)
    xy = rng.uniform(0.0, box_km, (nsta, 2))
    src = rng.uniform(-box_km, 2.0 * box_km, 2)
    delay = np.hypot(xy[:, 0] - src[0], xy[:, 1] - src[1]) / v_kms
    f = np.fft.rfftfreq(npts, dt)
    S = np.fft.rfft(rng.standard_normal(npts))
    S[(f < band[0]) | (f > band[1])] = 0.0
    baz = rng.uniform(0.0, 2.0 * np.pi, nsta)
    out = np.empty((nsta, ncomp, npts), np.float32)
    for s in range(nsta):
        sig = np.fft.irfft(S * np.exp(-2j * np.pi * f * delay[s]), n=npts)
        sigma = np.sqrt(np.mean(sig * sig) / snr)
        proj = (np.sin(baz[s]), np.cos(baz[s]), 1.0) if ncomp == 3 else (1.0,)
        for c in range(ncomp):
            out[s, c] = proj[c] * sig + sigma * rng.standard_normal(npts)
    return out, xy, delay


def write_day(tmp, traces, dt, timestamp, write_sac, first_id=1):
    """SAC files and index rows of one timestamp; traces [nsta, ncomp, npts].
    Returns the sac_index.tsv rows (timestamp, nsl_id, net, sta, loc, comp, path)."""
    nsta, ncomp, _ = traces.shape
    comps = COMPONENTS[ncomp]
    d = os.path.join(tmp, "sac", timestamp)
    os.makedirs(d, exist_ok=True)
    rows = []
    for s in range(nsta):
        sid = first_id + s
        for c, cn in enumerate(comps):
            p = os.path.join(d, "S%02d.%s.sac" % (sid, cn))
            write_sac(p, traces[s, c], dt, stla=35.0 + 0.01 * sid, stlo=-97.0 - 0.01 * sid,
                      kstnm="S%02d" % sid, knetwk="SY", kcmpnm=cn)
            rows.append((timestamp, str(sid), "SY", "S%02d" % sid, "--", cn, p))
    return rows


def write_index(path, rows):
    with open(path, "w") as f:
        f.write("timestamp\tnsl_id\tnetwork\tstation\tlocation\tcomponent\tsac_path\n")
        for r in rows:
            f.write("\t".join(r) + "\n")
