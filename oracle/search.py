"""Dedispersion + detection oracle (test infrastructure only).

Pinned by (``tests/test_oracle_search.py``): the closed-form shift table and
the dispersion-delay known answer ``delta_tau(nu_bot, nu_top, 3000) ~ 1697.78 ms``
(``tests/test_numerical_conventions.py:85-102``), brute-force shift-and-add
with zero fill (``tests/test_fine_dm_combiner.py:336-458``), explicit
sliding-window boxcar sums (``tests/test_detector_protocol.py:371-448``),
Welford/Chan merge == two-pass float64 mean/variance, an injected boxcar of
``snr / sqrt(w)`` on unit noise giving ``snr`` (``inject/cube_injection.py:285-329``),
and the top-k + merge decoder against an exhaustive greedy search
(``tests/x2/test_search2x_ref.py:151-175``).  The C inner loop is pinned
against ``search_block_np`` (pure numpy) on small inputs.
"""
from __future__ import annotations

import math
import ctypes
import os
import subprocess
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

K_DM = 4.148808
NATIVE_US = 32.768
_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "libsearch_oracle.so")
_SRC = os.path.join(_HERE, "csrc", "search_oracle.c")


def build_c(force: bool = False) -> str:
    """Compile the C inner loop with gcc (plain C, no GPU)."""
    if force or not os.path.exists(_SO) or os.path.getmtime(_SO) < os.path.getmtime(_SRC):
        os.makedirs(os.path.dirname(_SO), exist_ok=True)
        tmp = _SO + f".tmp{os.getpid()}"
        subprocess.check_call(["gcc", "-O2", "-fPIC", "-shared", "-std=c99", "-fopenmp", "-o", tmp, _SRC, "-lm"])
        os.replace(tmp, _SO)
    return _SO


_lib = None


def _c():
    global _lib
    if _lib is None:
        _lib = ctypes.CDLL(build_c())
        _lib.oracle_search_block.restype = ctypes.c_int
    return _lib


def shift_table(dm: np.ndarray, nu_ref_GHz: np.ndarray, nu_bot_GHz: float, t_step_us: float) -> np.ndarray:
    """``rint(1000 K dm (1/nu_bot^2 - 1/nu_ref^2) / t_step)``, half-to-even
    (dispersion.py:40-67, combiner.py:422-451)."""
    out = np.zeros((len(dm), len(nu_ref_GHz)), dtype=np.int32)
    for d, dmv in enumerate(dm):
        for s, nu in enumerate(nu_ref_GHz):
            d_ms = K_DM * float(dmv) * (1.0 / (nu_bot_GHz * nu_bot_GHz) - 1.0 / (float(nu) * float(nu)))
            out[d, s] = int(np.rint(1000.0 * d_ms / t_step_us))
    return out


def dedisperse_np(img: np.ndarray, tau0: int, shifts: np.ndarray, t_lo: int, t_hi: int) -> np.ndarray:
    """``D[d, t, p] = sum_s img[s, t - shift[d, s], p]`` for t in [t_lo, t_hi);
    zero before stream start.  ``img`` is ``(n_sub, n_tau, n_pix)`` for absolute
    steps ``[tau0, tau0 + n_tau)``."""
    n_sub, n_tau, n_pix = img.shape
    out = np.zeros((shifts.shape[0], t_hi - t_lo, n_pix))
    t = np.arange(t_lo, t_hi)
    for d in range(shifts.shape[0]):
        for s in range(n_sub):
            tau = t - int(shifts[d, s])
            ok = tau >= 0
            j = tau[ok] - tau0
            if j.size and (j.min() < 0 or j.max() >= n_tau):
                raise ValueError("image history does not cover the requested window")
            out[d, ok] += img[s, j]
        out[d, t < 0] = 0.0
    return out


def boxcar_np(x: np.ndarray, w: int, axis: int = 0) -> np.ndarray:
    """Centred sum over ``[i - w//2, i - w//2 + w)`` with zeros outside
    (forward.py:163-247); explicit sliding window."""
    x = np.moveaxis(np.asarray(x, dtype=np.float64), axis, 0)
    n = x.shape[0]
    out = np.zeros_like(x)
    for i in range(n):
        lo = max(0, i - w // 2)
        hi = min(n, i - w // 2 + w)
        if hi > lo:
            out[i] = x[lo:hi].sum(axis=0)
    return np.moveaxis(out, 0, axis)


def search_block_np(img, tau0, shifts, c_lo, n_centres, chunk, widths, st_n, st_mean, st_m2,
                    min_count=512):
    """Pure-numpy twin of ``oracle_search_block`` (slow; for pinning the C file)."""
    n_sub, n_tau, n_pix = img.shape
    wmax = max(widths)
    half = wmax // 2
    D = dedisperse_np(img, tau0, shifts, c_lo - half, c_lo + n_centres + half + 1)
    n_dm = shifts.shape[0]
    n_chunks = n_centres // chunk
    best_snr = np.full((n_dm, n_chunks, n_pix), -np.inf)
    best_t = np.full((n_dm, n_chunks, n_pix), -1, dtype=np.int64)
    best_k = np.full((n_dm, n_chunks, n_pix), -1, dtype=np.int32)
    cent = D[:, half:half + n_centres]
    valid_from = shifts.max(axis=1).astype(np.int64)
    tc = c_lo + np.arange(n_centres)
    blk_sum = np.zeros((n_dm, n_pix))
    blk_sumsq = np.zeros((n_dm, n_pix))
    blk_count = np.zeros(n_dm)
    for d in range(n_dm):
        ok = tc >= valid_from[d]
        blk_sum[d] = cent[d, ok].sum(axis=0)
        blk_sumsq[d] = (cent[d, ok] ** 2).sum(axis=0)
        blk_count[d] = ok.sum()
    for d in range(n_dm):
        if st_n[d] <= 0 or st_n[d] < min_count:
            continue
        mu = st_mean[d]
        sig = np.sqrt(st_m2[d] / st_n[d])
        for k, w in enumerate(widths):
            B = boxcar_np(D[d], w, axis=0)[half:half + n_centres]          # (n_centres, n_pix)
            for ch in range(n_chunks):
                if c_lo + ch * chunk - half < valid_from[d]:
                    continue
                blk = B[ch * chunk:(ch + 1) * chunk]
                mb = blk.max(axis=0)
                mt = blk.argmax(axis=0)
                snr = (mb - w * mu) / (sig * np.sqrt(w))
                better = (snr > best_snr[d, ch]) & (sig > 0)
                best_snr[d, ch] = np.where(better, snr, best_snr[d, ch])
                best_t[d, ch] = np.where(better, c_lo + ch * chunk + mt, best_t[d, ch])
                best_k[d, ch] = np.where(better, k, best_k[d, ch])
    return best_snr, best_t, best_k, blk_sum, blk_sumsq, blk_count


def search_block_c(img, tau0, shifts, c_lo, n_centres, chunk, widths, st_n, st_mean, st_m2,
                   min_count=512):
    """C inner loop; ``img`` is ``(n_sub, n_tau, n_pix)``, transposed to time-contiguous."""
    im = np.ascontiguousarray(np.transpose(img, (0, 2, 1)), dtype=np.float64)
    return search_block_c_tmajor(im, tau0, shifts, c_lo, n_centres, chunk, widths, st_n, st_mean, st_m2,
                                 min_count)


def search_block_c_tmajor(im, tau0, shifts, c_lo, n_centres, chunk, widths, st_n, st_mean, st_m2,
                          min_count=512):
    """:func:`search_block_c` on a time-contiguous history ``im`` ``(n_sub, n_pix, n_tau)``."""
    n_sub, n_pix, n_tau = im.shape
    im = np.ascontiguousarray(im, dtype=np.float64)
    n_dm = shifts.shape[0]
    n_chunks = n_centres // chunk
    sh = np.ascontiguousarray(shifts, dtype=np.int32)
    wd = np.ascontiguousarray(widths, dtype=np.int32)
    stn = np.ascontiguousarray(st_n, dtype=np.float64)
    stm = np.ascontiguousarray(st_mean, dtype=np.float64)
    st2 = np.ascontiguousarray(st_m2, dtype=np.float64)
    bs = np.empty((n_dm, n_chunks, n_pix))
    bt = np.empty((n_dm, n_chunks, n_pix), dtype=np.int64)
    bk = np.empty((n_dm, n_chunks, n_pix), dtype=np.int32)
    s1 = np.empty((n_dm, n_pix))
    s2 = np.empty((n_dm, n_pix))
    cnt = np.empty(n_dm)
    P = ctypes.c_void_p
    rc = _c().oracle_search_block(
        im.ctypes.data_as(P), ctypes.c_int(n_sub), ctypes.c_int(n_pix), ctypes.c_int64(tau0),
        ctypes.c_int(n_tau), sh.ctypes.data_as(P), ctypes.c_int(n_dm), ctypes.c_int64(c_lo),
        ctypes.c_int(n_centres), ctypes.c_int(chunk), wd.ctypes.data_as(P), ctypes.c_int(len(widths)),
        stn.ctypes.data_as(P), stm.ctypes.data_as(P), st2.ctypes.data_as(P),
        ctypes.c_double(float(min_count)), bs.ctypes.data_as(P), bt.ctypes.data_as(P), bk.ctypes.data_as(P),
        s1.ctypes.data_as(P), s2.ctypes.data_as(P), cnt.ctypes.data_as(P))
    if rc != 0:
        raise RuntimeError(f"oracle_search_block failed ({rc})")
    return bs, bt, bk, s1, s2, cnt


def merge_stats(n_a, mean_a, m2_a, n_b, sum_b, sumsq_b):
    """Chan et al. parallel merge of running (n, mean, M2) per DM with a block
    given as (count per DM, sum, sum of squares), float64."""
    n_a = np.asarray(n_a, dtype=np.float64)
    n_b = np.asarray(n_b, dtype=np.float64)
    nb = n_b[:, None]
    safe_b = np.where(nb > 0, nb, 1.0)
    mean_b = sum_b / safe_b
    m2_b = sumsq_b - sum_b * sum_b / safe_b
    n = n_a[:, None] + nb
    safe_n = np.where(n > 0, n, 1.0)
    delta = mean_b - mean_a
    mean = np.where(nb > 0, mean_a + delta * (nb / safe_n), mean_a)
    m2 = np.where(nb > 0, m2_a + m2_b + delta * delta * (n_a[:, None] * nb / safe_n), m2_a)
    return n_a + n_b, mean, m2


def block_decay(n_steps: int, t_int: int, tau_s: float) -> float:
    """Weight the running statistics keep after a block of ``n_steps`` steps
    of ``t_int`` native samples: ``exp(-duration / tau_s)``, the forgetting
    factor of an EMA with time constant ``tau_s`` (the reference's Layer-2
    ``gamma = 1 - exp(-cadence / tau)``, ``src/dsart/noise_norm/layer2.py:190-230``)."""
    return math.exp(-n_steps * t_int * NATIVE_US * 1e-6 / float(tau_s))


def decay_stats(n, m2, lam):
    """Forget after a merge: weight and M2 scaled by ``lam`` (the mean is a
    ratio and stays), so each earlier block weighs ``lam**age``."""
    return np.asarray(n, dtype=np.float64) * lam, np.asarray(m2, dtype=np.float64) * lam


@dataclass
class Cand:
    t: int
    dm_idx: int
    row: int
    col: int
    width: int
    snr: float


def decode(best_snr, best_t, best_k, widths, n_grid, *, threshold, n_top, radius_t, radius_dm,
           radius_lm) -> List[Cand]:
    """Top-k over chunk maxima above threshold, ordered by (-snr, t, dm, pixel),
    then greedy merge-radius suppression with AND semantics
    (decoder.py:436-533, search2x_ref.py:285-313)."""
    d_i, c_i, p_i = np.nonzero(best_snr > threshold)
    if d_i.size == 0:
        return []
    snr = best_snr[d_i, c_i, p_i]
    t = best_t[d_i, c_i, p_i].astype(np.int64)
    order = np.lexsort((p_i, d_i, t, -snr))[: int(n_top)]
    acc: List[Tuple[int, int, int, int]] = []
    out: List[Cand] = []
    for o in order:
        tt, dd, pp = int(t[o]), int(d_i[o]), int(p_i[o])
        r, c = divmod(pp, n_grid)
        if any(abs(tt - a) <= radius_t and abs(dd - b) <= radius_dm and abs(r - x) <= radius_lm
               and abs(c - y) <= radius_lm for a, b, x, y in acc):
            continue
        acc.append((tt, dd, r, c))
        out.append(Cand(t=tt, dm_idx=dd, row=r, col=c,
                        width=int(widths[int(best_k[d_i[o], c_i[o], p_i[o]])]), snr=float(snr[o])))
    return out
