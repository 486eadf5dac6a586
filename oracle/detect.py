"""Detector-over-a-cube oracle (test infrastructure only).

Restates, in float64 numpy:

* ``sigma_clipped_std`` -- ``src/dsart/noise_norm/layer1.py:53-122``: 3
  iterations of a 3-sigma clip about the *lower* median (``torch.nanmedian``),
  sigma = rms about that median over the kept samples;
* ``Layer2`` -- ``src/dsart/noise_norm/layer2.py:190-452``: per-kernel
  sigma_k, burn-in running mean for the first ``n_burnin`` cubes then an EMA
  with ``gamma = 1 - exp(-cadence / tau)``; a zero estimate keeps the old
  value; optional ceiling ``sigma_max_ratio * s_k`` with the clamp-escape
  hatch; optional floor; invalid cubes leave the state unchanged;
* ``cube_scores`` -- centred time boxcars of a ``[T, F, H, W]`` cube
  (``src/dsart/detector/forward.py:163-247``) scored as ``box_k / s_k``, the
  maximum over kernels with its winner (strict ``>``, ties to the lowest k;
  ``src/dsart/detector/triton_boxcar.py:193-302``) over the interior
  centres ``[w_max/2, T - w_max/2)`` (``forward.py:1370-1381``);
* ``decode_topk`` -- ``src/dsart/detector/decoder.py:388-533``: the global
  top ``n_top`` cells (ties in flat-index order), stop at ``snr <=
  threshold``, greedy suppression within the merge radii.

Pinned by ``tests/test_oracle_detect.py``: the lower-median known answer
(``[1, 2, 3, 100]`` -> sigma = sqrt(2401.5), which a mid-point median would
miss), the truncated-normal fixed point of the clip on 10^6 normal samples,
the burn-in mean and the EMA closed form of ``Layer2``, box sums against
explicit sliding windows, and planted, well-separated peaks surviving the
decode with the merge-radius invariants.
"""
from __future__ import annotations

import math
from typing import List, Optional, Sequence, Tuple

import numpy as np


def lower_median(x: np.ndarray) -> float:
    """Median of the finite values; the lower middle value for an even count
    (what ``torch.nanmedian`` returns)."""
    v = np.sort(x[np.isfinite(x)])
    if v.size == 0:
        return float("nan")
    return float(v[(v.size - 1) // 2])


def sigma_clipped_std(samples: np.ndarray, n_sigma: float = 3.0, n_iterations: int = 3) -> float:
    s = np.asarray(samples, dtype=np.float64).reshape(-1)
    active = np.isfinite(s)
    if not active.any():
        return 0.0
    med = lower_median(np.where(active, s, np.nan))
    sig = math.sqrt(np.sum((s[active] - med) ** 2) / max(1, active.sum()))
    for _ in range(int(n_iterations)):
        lo, hi = med - n_sigma * sig, med + n_sigma * sig
        active = active & (s >= lo) & (s <= hi)
        if not active.any():
            return 0.0
        med = lower_median(np.where(active, s, np.nan))
        sig = math.sqrt(np.sum((s[active] - med) ** 2) / max(1, active.sum()))
    return 0.0 if math.isnan(sig) else sig


class Layer2:
    def __init__(self, n_kernels: int, *, cadence_s: float, tau_s: float = 30.0, n_burnin: int = 30,
                 seed: Optional[Sequence[float]] = None, sigma_floor: float = 0.0,
                 sigma_max_ratio: float = 0.0, clamp_escape_cubes: int = 0):
        self.gamma = 1.0 - math.exp(-cadence_s / tau_s)
        self.n_burnin = int(n_burnin)
        self.s_k = np.ones(n_kernels) if seed is None else np.asarray(seed, dtype=np.float64).copy()
        self.cube_count = 0
        self.floor = float(sigma_floor)
        self.ratio = float(sigma_max_ratio)
        self.escape = int(clamp_escape_cubes)
        self.streak = np.zeros(n_kernels, dtype=np.int64)

    @property
    def is_warming_up(self) -> bool:
        return self.cube_count < self.n_burnin

    def update(self, sigma_this: np.ndarray, valid: bool = True) -> np.ndarray:
        if not valid:
            return self.s_k.copy()
        x = np.asarray(sigma_this, dtype=np.float64).copy()
        x = np.where(x == 0, self.s_k, x)
        if self.ratio > 0 and self.cube_count > 0:
            high = x > self.ratio * self.s_k
            if high.any():
                esc = high & (self.streak >= self.escape) if self.escape > 0 else np.zeros_like(high)
                clamp = high & ~esc
                self.s_k = np.where(esc, x, self.s_k)
                x = np.where(clamp, self.s_k, x)
                self.streak = np.where(clamp, self.streak + 1, 0)
            else:
                self.streak[:] = 0
        if self.cube_count < self.n_burnin:
            c = float(self.cube_count)
            self.s_k = (c * self.s_k + x) / (c + 1.0)
        else:
            self.s_k = self.gamma * x + (1.0 - self.gamma) * self.s_k
        if self.floor > 0:
            self.s_k = np.maximum(self.s_k, self.floor)
        self.cube_count += 1
        return self.s_k.copy()


def box_sums(cube: np.ndarray, widths: Sequence[int]) -> np.ndarray:
    """``[K, T, ...]``: sum of ``cube[c - w//2 : c - w//2 + w]`` along axis 0,
    zeros outside."""
    x = np.asarray(cube, dtype=np.float64)
    T = x.shape[0]
    P = np.concatenate([np.zeros((1,) + x.shape[1:]), np.cumsum(x, axis=0)])
    out = np.zeros((len(widths),) + x.shape)
    for k, w in enumerate(widths):
        for c in range(T):
            lo = min(max(c - w // 2, 0), T)
            hi = min(max(c - w // 2 + w, 0), T)
            out[k, c] = P[hi] - P[lo]
    return out


def cube_scores(cube: np.ndarray, widths: Sequence[int], s_k: np.ndarray) -> Tuple[np.ndarray, np.ndarray, int]:
    """Maximum S/N over kernels and the winner, over the interior centres;
    returns ``(snr [n_out, F, H, W], winner, t_base)``."""
    T = cube.shape[0]
    h = max(widths) // 2
    t_base = h if 2 * h < T else 0
    n_out = T - 2 * t_base
    B = box_sums(cube, widths)[:, t_base:t_base + n_out]
    snr = B / np.asarray(s_k, dtype=np.float64).reshape((-1,) + (1,) * (B.ndim - 1))
    best = np.full(snr.shape[1:], -np.inf)
    win = np.zeros(snr.shape[1:], dtype=np.int64)
    for k in range(len(widths)):
        better = snr[k] > best
        best = np.where(better, snr[k], best)
        win = np.where(better, k, win)
    return best, win, t_base


def decode_topk(snr: np.ndarray, winner: np.ndarray, *, threshold: float, n_top: int, radius_t: int,
                radius_fdm: int, radius_lm: int) -> List[Tuple[int, int, int, int, int, float]]:
    """Candidates ``(t, f, l, m, winner k, snr)`` in emission order."""
    n_out, F, H, W = snr.shape
    flat = snr.reshape(-1)
    order = np.lexsort((np.arange(flat.size), -flat))[: int(n_top)]
    out, acc = [], []
    for ix in order:
        v = float(flat[ix])
        if v <= threshold:
            break
        t, rem = divmod(int(ix), F * H * W)
        f, rem = divmod(rem, H * W)
        l, m = divmod(rem, W)
        if any(abs(t - a) <= radius_t and abs(f - b) <= radius_fdm and abs(l - c) <= radius_lm and
               abs(m - d) <= radius_lm for a, b, c, d in acc):
            continue
        acc.append((t, f, l, m))
        out.append((t, f, l, m, int(winner.reshape(-1)[ix]), v))
    return out
