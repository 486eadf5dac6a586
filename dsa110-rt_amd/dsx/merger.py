"""Cross-kernel candidate merger: SNR-ordered greedy suppression.

Two rules, both from ``src/dsart/detector/merger.py``:

* ``merge_across_kernels_c1`` (``merger.py:123-200``) -- candidate ``c`` is
  suppressed by an earlier survivor ``s`` iff

  - ``|c.event_specnum - s.event_specnum| <= t_frac * 0.5 * (c.w + s.w) * sample_period_specnum``,
  - ``|c.dm_idx - s.dm_idx| <= dm_max_trials``, and
  - ``|c.l - s.l| <= lm_max_cells`` OR ``|c.m - s.m| <= lm_max_cells``
    (the cross-shaped core PSF leaves sidelobes along one arm at a time).

  The ``sample_period_specnum`` factor widens the window beyond one
  search sample per width unit; the reference keeps it on purpose as a
  tuned radius (``merger.py:26-40``), and so does this module.
* ``merge_across_kernels`` (``merger.py:229-268``) -- the legacy box rule:
  suppressed iff within ``merge_radius_t``, ``merge_radius_fdm`` and
  ``merge_radius_lm`` (on both l and m) of a survivor.

Candidates are visited in ``(-snr, event_specnum, dm_idx, l, m, kernel_id)``
order, so the survivors do not depend on the input order.  Each candidate
is tested against all survivors at once with numpy (one vector comparison
per candidate instead of the reference's per-pair Python loop); the
predicates and their float64 arithmetic are the reference's.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Sequence

import numpy as np

DEFAULT_MERGE_RADIUS_LM: int = 3
DEFAULT_MERGE_RADIUS_FDM: int = 5
DEFAULT_MERGE_RADIUS_T: int = 128


@dataclass(frozen=True)
class MergerConfig:
    """C1 merger geometry (``merger.py:78-120``): OR-mode (l, m)
    half-window in cells, fine-DM half-window in trials, scale on the
    width-aware time window, and raw spec-num units per search sample."""

    lm_max_cells: int = 3
    dm_max_trials: int = 2
    t_frac: float = 1.0
    sample_period_specnum: int = 16

    def __post_init__(self) -> None:
        if self.lm_max_cells < 0:
            raise ValueError(f"lm_max_cells={self.lm_max_cells}, expected ≥ 0")
        if self.dm_max_trials < 0:
            raise ValueError(f"dm_max_trials={self.dm_max_trials}, expected ≥ 0")
        if self.t_frac < 0.0:
            raise ValueError(f"t_frac={self.t_frac}, expected ≥ 0")
        if self.sample_period_specnum <= 0:
            raise ValueError(f"sample_period_specnum={self.sample_period_specnum}, expected > 0")


def _sorted(candidates: Sequence) -> list:
    return sorted(candidates, key=lambda c: (-c.snr, c.event_specnum, c.dm_idx, c.l, c.m, c.kernel_id))


def _columns(cands: list):
    t = np.fromiter((int(c.event_specnum) for c in cands), dtype=np.int64, count=len(cands))
    d = np.fromiter((int(c.dm_idx) for c in cands), dtype=np.int64, count=len(cands))
    l = np.fromiter((float(c.l) for c in cands), dtype=np.float64, count=len(cands))
    m = np.fromiter((float(c.m) for c in cands), dtype=np.float64, count=len(cands))
    return t, d, l, m


def merge_across_kernels_c1(candidates: Sequence, cfg: MergerConfig) -> List:
    """Survivors of the C1 rule, in visiting order (brightest first)."""
    if not candidates:
        return []
    cands = _sorted(candidates)
    n = len(cands)
    t, d, l, m = _columns(cands)
    w = np.fromiter((int(c.width_samples) for c in cands), dtype=np.int64, count=n)
    lm = int(cfg.lm_max_cells)
    dm = int(cfg.dm_max_trials)
    tf = float(cfg.t_frac)
    sps = float(cfg.sample_period_specnum)
    surv = np.empty(n, dtype=np.int64)
    ns = 0
    for i in range(n):
        if ns:
            s = surv[:ns]
            win = tf * 0.5 * (w[i] + w[s]).astype(np.float64) * sps
            hit = ((np.abs(t[i] - t[s]) <= win) & (np.abs(d[i] - d[s]) <= dm)
                   & ((np.abs(l[i] - l[s]) <= lm) | (np.abs(m[i] - m[s]) <= lm)))
            if hit.any():
                continue
        surv[ns] = i
        ns += 1
    return [cands[i] for i in surv[:ns]]


def merge_across_kernels(candidates: Sequence, *, merge_radius_lm: int = DEFAULT_MERGE_RADIUS_LM,
                         merge_radius_fdm: int = DEFAULT_MERGE_RADIUS_FDM,
                         merge_radius_t: int = DEFAULT_MERGE_RADIUS_T) -> List:
    """Survivors of the legacy axis-AND box rule, brightest first."""
    if merge_radius_lm < 0 or merge_radius_fdm < 0 or merge_radius_t < 0:
        raise ValueError(f"merge_radius_lm/fdm/t must be ≥ 0; got "
                         f"({merge_radius_lm}, {merge_radius_fdm}, {merge_radius_t})")
    if not candidates:
        return []
    cands = _sorted(candidates)
    n = len(cands)
    t, d, l, m = _columns(cands)
    surv = np.empty(n, dtype=np.int64)
    ns = 0
    for i in range(n):
        if ns:
            s = surv[:ns]
            hit = ((np.abs(t[i] - t[s]) <= merge_radius_t) & (np.abs(d[i] - d[s]) <= merge_radius_fdm)
                   & (np.abs(l[i] - l[s]) <= merge_radius_lm) & (np.abs(m[i] - m[s]) <= merge_radius_lm))
            if hit.any():
                continue
        surv[ns] = i
        ns += 1
    return [cands[i] for i in surv[:ns]]
