"""FP64 oracle of xc_fast (stage 2) -- test infrastructure only.

Follows, for one timestamp:
  finalize_shape            native/xc/io/input.cu:1033-1054
  build_jobs                native/xc/worker/scheduler.cu:43-63 (kRowTargetBlocks = 1,
                            native/xc/runtime.hpp:25)
  build_tasks_for_job       native/xc/worker/scheduler.cu:65-152
  find_allowed_path_canonical / make_allowed_path_id
                            native/xc/io/path_table.c:9-36, 179-191
  cross_spectrum            accumulateStepXc2DKernel native/xc/kernels/cuda.kernels.cu:16-36
                            launched per window at native/xc/worker/async_workers.cu:65-83
  lags                      finalize_xc_batch native/xc/worker/executor.cu:6-45 and
                            InvNormalize2DKernel native/xc/kernels/cuda.kernels.cu:51-62

Pinning (tests/test_oracle_xc.py):
  cross_spectrum + lags  pinned by brute-force time-domain correlation
                         dt * sum_t x[t] y[t+tau] of windows whose spectra are
                         built as sac2spec does (rfft * dt), and by the peak of a
                         delayed copy landing at +delay (conj on the source side);
  build_tasks_for_job    pinned by invariants: every allowed unordered pair of
                         the upper block triangle appears exactly once, oriented
                         src = lower nsl id, in the documented loop order, plus a
                         brute-force enumeration of the same rule on small inputs;
  finalize_shape         pinned by closed forms on hand-computed cases;
  cross_spectrum_valid   (the gap-excluding stack of the batched uneven form,
                         which has no reference counterpart) pinned by its
                         reduction to cross_spectrum when every window is
                         valid, and by pairs whose stations share exactly one
                         valid window (the stack is that window's product) or
                         none (zero).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Shape:
    nspec: int
    nstep: int
    nfft: int
    half_cc: int
    cc_size: int
    dt: float


def finalize_shape(nspec: int, nstep: int, dt: float, cc_length: float) -> Shape:
    """nfft = 2(nspec-1); H = llround(C/dt) clamped to nspec-1; cc = 2H+1."""
    if nspec < 2 or nstep <= 0 or dt <= 0 or cc_length <= 0:
        raise ValueError("invalid shape")
    nfft = 2 * (nspec - 1)
    # llround of a double quotient of two float32 values
    q = float(np.float64(np.float32(cc_length)) / np.float64(np.float32(dt)))
    half = int(np.floor(abs(q) + 0.5)) * (1 if q >= 0 else -1)
    if half >= nspec:
        half = nspec - 1
    cc = 2 * half + 1
    if cc <= 0 or cc > nfft:
        raise ValueError("cc_size out of range")
    return Shape(nspec, nstep, nfft, half, cc, float(np.float32(dt)))


def make_allowed_path_id(src: int, rec: int) -> int:
    if src <= 0 or src > 9999 or rec <= 0 or rec > 9999:
        return -1
    return src * 10000 + rec


@dataclass
class Job:
    anchor_block: int
    target_begin_block: int
    target_end_block: int
    block_size: int
    file_count: int
    anchor_begin: int
    anchor_end: int
    target_begin: int
    target_end: int


def _fill_job(file_count, block_size, anchor, tbeg, tend) -> Job:
    ab = anchor * block_size
    return Job(anchor, tbeg, tend, block_size, file_count, ab, min(file_count, ab + block_size),
               tbeg * block_size, min(file_count, tend * block_size))


def build_jobs(file_count: int, block_size: int, row_target_blocks: int = 1) -> list[Job]:
    jobs = []
    nblk = (file_count + block_size - 1) // block_size
    for anchor in range(nblk):
        jobs.append(_fill_job(file_count, block_size, anchor, anchor, anchor + 1))
        t = anchor + 1
        while t < nblk:
            jobs.append(_fill_job(file_count, block_size, anchor, t, min(nblk, t + row_target_blocks)))
            t += row_target_blocks
    return jobs


@dataclass
class Task:
    src_meta: int
    rec_meta: int
    src_local: int
    rec_local: int
    path_id: int
    is_autocorr: bool


def build_tasks_for_job(gnsl_ids, allowed: set[int], job: Job):
    """Returns (tasks, loaded_meta_indices) in the reference's order.

    gnsl_ids: nsl id of each NSLC row (rows sorted by id; several channels of
    one station share the id).  allowed: set of canonical path ids.
    """
    loaded = list(range(job.anchor_begin, job.anchor_end))
    if job.target_begin != job.anchor_begin:
        loaded += list(range(job.target_begin, job.target_end))

    def local_idx(meta):
        if job.anchor_begin <= meta < job.anchor_end:
            return meta - job.anchor_begin
        return (job.anchor_end - job.anchor_begin) + (meta - job.target_begin)

    tasks: list[Task] = []

    def push(s, r, pid, auto):
        tasks.append(Task(s, r, local_idx(s), local_idx(r), pid, auto))

    def try_pair(a, b, same_block):
        ga, gb = int(gnsl_ids[a]), int(gnsl_ids[b])
        auto = ga == gb
        if same_block and not auto and b <= a:
            return
        lo, hi = (ga, gb) if ga <= gb else (gb, ga)
        pid = make_allowed_path_id(lo, hi)
        if pid <= 0 or pid not in allowed:
            return
        if auto:
            push(a, b, pid, True)
            if not same_block:
                push(b, a, pid, True)
            return
        if ga < gb:
            push(a, b, pid, False)
        else:
            push(b, a, pid, False)

    for tb in range(job.target_begin_block, job.target_end_block):
        begin = tb * job.block_size
        end = min(job.file_count, begin + job.block_size)
        same = tb == job.anchor_block
        if same:
            begin, end = job.anchor_begin, job.anchor_end
        for a in range(job.anchor_begin, job.anchor_end):
            for b in range(begin, end):
                try_pair(a, b, same)
    return tasks, loaded


def cross_spectrum(spec: np.ndarray, src, rec) -> np.ndarray:
    """(1/nstep) sum_w conj(S[w,src,k]) S[w,rec,k] in FP64; spec [nstep,nsta,nspec]."""
    s = np.asarray(spec, dtype=np.complex128)
    src = np.asarray(src)
    rec = np.asarray(rec)
    nstep = s.shape[0]
    acc = np.zeros((len(src), s.shape[2]), dtype=np.complex128)
    for w in range(nstep):
        acc += np.conj(s[w, src, :]) * s[w, rec, :]
    return acc / nstep


def cross_spectrum_valid(spec: np.ndarray, valid: np.ndarray, src, rec) -> np.ndarray:
    """Gap-excluding stack: sum over the windows w in which both valid[w, src]
    and valid[w, rec] are set of conj(S[w,src,k]) S[w,rec,k], divided by the
    number of such windows (0 windows: a zero row); FP64.  valid [nstep, nsta]."""
    s = np.asarray(spec, dtype=np.complex128)
    v = np.asarray(valid).astype(bool)
    src = np.asarray(src)
    rec = np.asarray(rec)
    acc = np.zeros((len(src), s.shape[2]), dtype=np.complex128)
    cnt = np.zeros(len(src))
    for w in range(s.shape[0]):
        both = v[w, src] & v[w, rec]
        acc[both] += np.conj(s[w, src[both], :]) * s[w, rec[both], :]
        cnt += both
    out = np.zeros_like(acc)
    nz = cnt > 0
    out[nz] = acc[nz] / cnt[nz, None]
    return out


def lags(xspec: np.ndarray, nfft: int, half_cc: int, dt: float) -> np.ndarray:
    """(-1)^k shift, unnormalised C2R(nfft), x 1/(nfft*dt), lags [nfft/2-H, nfft/2+H].

    The reference scale is 1.0/((float)nfft*dt) evaluated in double.
    """
    X = np.asarray(xspec, dtype=np.complex128)
    nspec = X.shape[-1]
    sign = np.where(np.arange(nspec) % 2 == 0, 1.0, -1.0)
    t = np.fft.irfft(X * sign, n=nfft, axis=-1) * nfft  # unnormalised C2R
    weight = 1.0 / float(np.float32(np.float32(nfft) * np.float32(dt)))
    t = t * weight
    return t[..., nfft // 2 - half_cc: nfft // 2 + half_cc + 1]


def correlate(spec: np.ndarray, src, rec, nfft: int, half_cc: int, dt: float) -> np.ndarray:
    return lags(cross_spectrum(spec, src, rec), nfft, half_cc, dt)


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.float64) if not np.iscomplexobj(a) else np.asarray(a, np.complex128)
    b = np.asarray(b, dtype=np.float64) if not np.iscomplexobj(b) else np.asarray(b, np.complex128)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))
