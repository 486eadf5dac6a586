"""Subband dedispersion tables and the stand-alone combine stage.

* ``compute_time_shift_search`` -- the per-(DM trial, subband) read-back
  shift table of ``src/dsart/fine_dm/combiner.py:150-491`` in the form a
  single node uses: one rounding of the fine delay from each subband's
  reference frequency to the band bottom (the ``n_sub > 1`` branch,
  ``combiner.py:422-451``, with no corr-side coarse term).
* ``combine_chgroups`` -- ``combiner.py:548-655``: shift-and-add of dense
  per-stream grids with zero fill, on the GPU (``dsx_combine``).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Mapping, Optional, Tuple

import numpy as np
import torch

from . import _lib
from .dispersion import subband_shift_table


@dataclass(frozen=True)
class TimeShiftSearchTable:
    shifts: np.ndarray            # [n_dm, n_stream] int32
    dm_pc_cm3: np.ndarray         # [n_dm] float64
    t_int_search_us: float

    def __post_init__(self):
        if self.shifts.ndim != 2 or self.shifts.dtype != np.int32:
            raise TypeError("shifts must be a 2-D int32 array")
        if self.dm_pc_cm3.shape[0] != self.shifts.shape[0]:
            raise ValueError("one DM per shift row")
        if self.t_int_search_us <= 0:
            raise ValueError("t_int_search_us must be > 0")


def compute_time_shift_search(*, dm_pc_cm3: np.ndarray, nu_subband_ref_GHz: np.ndarray,
                              nu_bot_GHz: float, t_int_search_us: float) -> TimeShiftSearchTable:
    dm = np.asarray(dm_pc_cm3, dtype=np.float64)
    if dm.ndim != 1:
        raise ValueError("dm_pc_cm3 must be 1-D")
    return TimeShiftSearchTable(shifts=subband_shift_table(dm, nu_subband_ref_GHz, nu_bot_GHz, t_int_search_us),
                                dm_pc_cm3=dm, t_int_search_us=float(t_int_search_us))


def combine_chgroups(*, per_chgroup_streams: Mapping[int, torch.Tensor], time_shift_per_chgroup,
                     t_window: Tuple[int, int], n_grid: Optional[int] = None,
                     n_chgroup: Optional[int] = None) -> torch.Tensor:
    """``out[t] = sum_g stream_g[t - shift[g]]`` for ``t`` in ``t_window``,
    zero-filled outside each stream; missing streams contribute nothing.
    Streams are CUDA tensors ``[T_stream, N, N]`` (complex64 or float32);
    returns a CUDA tensor ``[t_hi - t_lo, N, N]`` of the streams' dtype."""
    shifts = np.asarray(time_shift_per_chgroup)
    if not np.issubdtype(shifts.dtype, np.integer):
        raise TypeError(f"time_shift_per_chgroup.dtype={shifts.dtype}, expected integer")
    n_g = int(n_chgroup if n_chgroup is not None else shifts.shape[0])
    if shifts.shape != (n_g,):
        raise ValueError(f"time_shift_per_chgroup must have shape ({n_g},); got {shifts.shape}")
    t_lo, t_hi = int(t_window[0]), int(t_window[1])
    if t_hi <= t_lo:
        raise ValueError(f"t_window={t_window!r} has non-positive width")
    present = sorted(g for g in per_chgroup_streams if 0 <= g < n_g)
    if not present:
        if n_grid is None:
            raise ValueError("could not infer n_grid; pass n_grid= when no stream is present")
        return torch.zeros((t_hi - t_lo, n_grid, n_grid), dtype=torch.complex64, device="cuda")
    first = per_chgroup_streams[present[0]]
    if first.dim() != 3 or first.shape[1] != first.shape[2]:
        raise ValueError(f"streams must be [T_stream, N, N]; got {tuple(first.shape)}")
    N = int(first.shape[1]) if n_grid is None else int(n_grid)
    T = int(first.shape[0])
    for g in present:
        s = per_chgroup_streams[g]
        if tuple(s.shape) != (T, N, N) or s.dtype != first.dtype:
            raise ValueError(f"per_chgroup_streams[{g}] shape/dtype differs from the first stream")
    stacked = torch.stack([per_chgroup_streams[g] for g in present]).contiguous()
    is_c = stacked.is_complex()
    flt = torch.view_as_real(stacked).contiguous() if is_c else stacked.to(torch.float32).contiguous()
    inner = N * N * (2 if is_c else 1)
    sh = torch.from_numpy(shifts[present].astype(np.int32)).to(stacked.device)
    out = torch.empty((t_hi - t_lo,) + tuple(flt.shape[2:]), dtype=torch.float32, device=stacked.device)
    _lib.require_gpu(flt, sh, out)
    _lib.call("dsx_combine", _lib.ptr(flt), _lib.ptr(sh), _lib.ptr(out), len(present), T, inner, t_lo,
              t_hi - t_lo, _lib.stream())
    return torch.view_as_complex(out) if is_c else out.to(first.dtype)
