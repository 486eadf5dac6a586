"""Cross-kernel merger oracle (test infrastructure only).

Pairwise pure-Python restatement of ``src/dsart/detector/merger.py``:
``merge_c1`` follows ``merge_across_kernels_c1`` (``merger.py:123-200``) and
``merge_box`` follows ``merge_across_kernels`` (``merger.py:229-268``).
Candidates are any objects with ``snr, event_specnum, dm_idx, l, m,
kernel_id, width_samples``.

Pinned by ``tests/test_merger.py``: the reference's own known-answer cases
(``tests/test_c1_merger.py``, ``tests/test_detector_merger.py`` in the
reference, restated with their inputs and expected survivor counts), plus
the invariants that fix a greedy suppression scan -- no survivor lies in
the window of an earlier survivor, every dropped candidate lies in the
window of an earlier survivor, and the result does not depend on the
input order.
"""
from __future__ import annotations

from typing import List, Sequence


def _order(cands: Sequence) -> list:
    return sorted(cands, key=lambda c: (-c.snr, c.event_specnum, c.dm_idx, c.l, c.m, c.kernel_id))


def c1_pair(c, s, *, lm_max: int, dm_max: int, t_frac: float, sps: float) -> bool:
    """True iff ``s`` suppresses ``c`` under the C1 rule."""
    win = t_frac * 0.5 * float(int(c.width_samples) + int(s.width_samples)) * sps
    if abs(int(c.event_specnum) - int(s.event_specnum)) > win:
        return False
    if abs(int(c.dm_idx) - int(s.dm_idx)) > dm_max:
        return False
    return abs(float(c.l) - float(s.l)) <= lm_max or abs(float(c.m) - float(s.m)) <= lm_max


def box_pair(c, s, *, r_lm: int, r_fdm: int, r_t: int) -> bool:
    """True iff ``s`` suppresses ``c`` under the legacy box rule."""
    return (abs(c.event_specnum - s.event_specnum) <= r_t and abs(c.dm_idx - s.dm_idx) <= r_fdm
            and abs(c.l - s.l) <= r_lm and abs(c.m - s.m) <= r_lm)


def merge_c1(cands: Sequence, *, lm_max: int = 3, dm_max: int = 2, t_frac: float = 1.0,
             sps: float = 16.0) -> List:
    out: List = []
    for c in _order(cands):
        if not any(c1_pair(c, s, lm_max=lm_max, dm_max=dm_max, t_frac=t_frac, sps=sps) for s in out):
            out.append(c)
    return out


def merge_box(cands: Sequence, *, r_lm: int = 3, r_fdm: int = 5, r_t: int = 128) -> List:
    out: List = []
    for c in _order(cands):
        if not any(box_pair(c, s, r_lm=r_lm, r_fdm=r_fdm, r_t=r_t) for s in out):
            out.append(c)
    return out
