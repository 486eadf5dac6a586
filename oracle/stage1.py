"""Stage 1 (per-channel coarse-DM shifts before gridding) and static-sky
subtraction, restated on the CPU (test infrastructure only).

* :func:`stage1_bins` -- ``DMPlan.delay_bins_per_subband``
  (``src/dsart/coarse_dm/dm_plan.py:479-515``): channels summed in groups of
  ``group`` (the summed-channel table, ``dm_plan.py:160-176``), each group
  referenced to its subband's first group, two roundings (native sample,
  then step).
* :func:`stage1_visibilities` -- ``Stage1MultiDMCoarseDM`` on a 2-block
  window (``services/corr_fast_integration.py:917-1040``): output step ``t``
  of a channel is its visibility at ``t + bins[c]``; per-channel exact
  int64 visibilities of the concatenated (previous tail, block) samples,
  summed over each output channel's ``csf`` channels after the shift.
* :class:`StaticSkyMean` -- ``corr_fast_integration.py:1049-1246`` on complex
  uv grids: subtract the mean of the previous ``window`` blocks' per-cell
  means (after ``warmup`` blocks), then push this block's per-cell mean.

Pinned by ``tests/test_oracle_stage1.py``: every subband's first group has
shift 0, shifts never exceed the whole-chgroup ones and grow with DM and
with distance from the subband top (``tests/test_subband_n_sub.py:185-202``);
a per-channel impulse at ``t0 + bins[c]`` lands on ``t0`` in every channel
(``tests/test_coarse_dm_stage1.py:120-200``); a burst dispersed at the coarse
DM is aligned within each subband; the static-sky window equals a brute-force
mean of the last ``window`` block means, a constant sky is removed exactly
after warm-up and a block never subtracts itself.
"""
from __future__ import annotations

import numpy as np

from . import corr as ocorr

K_DM_MS_GHZ2_PC = 4.148808
NATIVE_US = 32.768


def group_freqs(freqs_GHz, group: int) -> np.ndarray:
    f = np.asarray(freqs_GHz, dtype=np.float64)
    return f if group == 1 else f.reshape(-1, group).mean(axis=1)


def stage1_bins(freqs_GHz, *, n_sub: int, group: int, t_int: int, dm: float) -> np.ndarray:
    """``(n_chan,)`` int64 shift of every channel (its group's shift)."""
    f = group_freqs(freqs_GHz, group)
    per = f.size // n_sub
    ref = np.repeat(f[::per][:n_sub], per)
    us = K_DM_MS_GHZ2_PC * float(dm) * ((1.0 / f ** 2) - (1.0 / ref ** 2)) * 1e3
    bins = np.rint(np.rint(us / NATIVE_US) / t_int).astype(np.int64)
    return np.repeat(bins, group)


def subband_refs(freqs_GHz, *, n_sub: int, group: int) -> np.ndarray:
    """Frequency each subband is aligned to by stage 1: its first group's."""
    f = group_freqs(freqs_GHz, group)
    return f[:: f.size // n_sub][:n_sub]


def stage1_visibilities(raw_tail: np.ndarray, raw: np.ndarray, *, n_ant: int, n_chan: int, t_int: int,
                        csf: int, bins: np.ndarray, delay: int) -> np.ndarray:
    """int64 ``(L, n_chan // csf, nbase, 2)`` Stokes-I visibilities of the
    stage-1 stream for one block: output step ``t`` (absolute ``T0 - delay + t``)
    sums channel ``c`` of step ``t + bins[c]`` of the concatenated
    (``raw_tail``: the previous block's last ``delay`` steps, ``raw``) samples."""
    ppf = t_int // 2
    both = np.concatenate([np.asarray(raw_tail, np.uint8).reshape(-1, n_ant, n_chan, 2, 2),
                           np.asarray(raw, np.uint8).reshape(-1, n_ant, n_chan, 2, 2)])
    n_pkt_all = both.shape[0]
    L = n_pkt_all // ppf - delay
    vc = ocorr.correlate_blas(both, n_pkt=n_pkt_all, n_ant=n_ant, n_chan=n_chan, t_int=t_int, csf=1)
    # vc: (steps, n_chan, nbase, 2)
    nco = n_chan // csf
    out = np.zeros((L, nco, vc.shape[2], 2), dtype=np.int64)
    t = np.arange(L)
    for c in range(n_chan):
        out[:, c // csf] += vc[t + int(bins[c]), c]
    return out


class StaticSkyMean:
    """Per-cell window mean of the previous blocks, subtracted from each block
    (complex; causal: a block never subtracts itself)."""

    def __init__(self, window: int = 8, warmup: int = 8):
        if window < 1 or warmup < 0:
            raise ValueError("window >= 1 and warmup >= 0")
        self.window = int(window)
        self.warmup = int(warmup)
        self.means = []          # the last `window` block means, oldest first
        self.seen = 0

    def apply(self, grids: np.ndarray) -> np.ndarray:
        """``grids`` (n_steps, ...) -> the block with the window mean removed."""
        g = np.asarray(grids)
        mean_now = g.mean(axis=0)
        if self.means and self.seen >= self.warmup:
            out = g - np.mean(np.stack(self.means), axis=0)[None]
        else:
            out = g.copy()
        self.means.append(mean_now)
        if len(self.means) > self.window:
            self.means.pop(0)
        self.seen += 1
        return out
