"""Dispersion delays and integer shift tables.

``delta_tau_ms`` / ``delta_tau_us`` keep the operation order of
``src/dsart/common/dispersion.py:40-67`` so that ``np.rint`` of a delay lands
on the same integer as the reference at rounding boundaries.  The per-subband
search shift follows ``src/dsart/fine_dm/combiner.py:422-451`` (one rounding
of the fine delay between the band bottom and each subband's reference
frequency, ``np.rint`` = half-to-even).
"""
from __future__ import annotations

import numpy as np

from .constants import K_DM_MS_GHZ2_PC


def delta_tau_ms(nu_low_GHz: float, nu_high_GHz: float, dm_pc_cm3: float) -> float:
    """``K * DM * (1/nu_low^2 - 1/nu_high^2)`` in ms (dispersion.py:40-58)."""
    if nu_low_GHz <= 0.0 or nu_high_GHz <= 0.0:
        raise ValueError(
            f"frequencies must be > 0; got nu_low={nu_low_GHz!r}, nu_high={nu_high_GHz!r}"
        )
    return K_DM_MS_GHZ2_PC * dm_pc_cm3 * (
        1.0 / (nu_low_GHz * nu_low_GHz) - 1.0 / (nu_high_GHz * nu_high_GHz)
    )


def delta_tau_us(nu_low_GHz: float, nu_high_GHz: float, dm_pc_cm3: float) -> float:
    """Pair-wise delay in us (dispersion.py:61-67)."""
    return 1000.0 * delta_tau_ms(nu_low_GHz, nu_high_GHz, dm_pc_cm3)


def subband_shift_table(
    dm_trials: np.ndarray,
    nu_ref_GHz: np.ndarray,
    nu_bot_GHz: float,
    t_step_us: float,
) -> np.ndarray:
    """``[n_dm, n_sub] int32`` read-back shifts of each subband stream.

    ``shift[d, s] = rint(delta_tau_us(nu_bot, nu_ref[s], dm[d]) / t_step_us)``.
    A dedispersed series reads subband ``s`` at ``t - shift[d, s]``, i.e. the
    output time is the arrival time at ``nu_bot`` (combiner.py:625-653).
    """
    dm = np.asarray(dm_trials, dtype=np.float64)
    refs = np.asarray(nu_ref_GHz, dtype=np.float64)
    if dm.ndim != 1 or refs.ndim != 1:
        raise ValueError("dm_trials and nu_ref_GHz must be 1-D")
    if t_step_us <= 0.0:
        raise ValueError(f"t_step_us must be > 0; got {t_step_us!r}")
    out = np.zeros((dm.size, refs.size), dtype=np.int32)
    for d in range(dm.size):
        for s in range(refs.size):
            out[d, s] = int(np.rint(
                delta_tau_us(float(nu_bot_GHz), float(refs[s]), float(dm[d])) / t_step_us
            ))
    return out
