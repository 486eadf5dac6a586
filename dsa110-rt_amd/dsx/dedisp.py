"""Subband dedispersion tables and the stand-alone combine stage.

* ``compute_time_shift_search`` / ``TimeShiftSearchTable`` --
  ``src/dsart/fine_dm/combiner.py:66-491``: the per-(fine DM, stream) signed
  read-back shift table, with every mode of the reference (residual after the
  corr-side coarse alignment, ``include_coarse_offset``,
  ``merge_coarse_rounding`` and ``n_sub > 1`` subband streams).
* ``compute_delay_native_samples_table`` -- ``coarse_dm/dm_plan.py:194-243``.
* ``build_time_shifts`` -- the three integer tables stored in a DM plan file
  (``tools/build_dm_plan.py:322-387``: per-channel stage 1, per-chgroup
  stage 2, search residual).
* ``combine_chgroups`` -- ``combiner.py:548-655``: shift-and-add of dense
  per-stream grids with zero fill, on the GPU (``dsx_combine``).

All delays go through ``dispersion.delta_tau_us`` and ``np.rint``
(half-to-even), in the reference's operation order, so integer tables land on
the same side of every rounding boundary.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Mapping, Optional, Tuple

import numpy as np
import torch

from . import _lib
from .constants import N_CHGROUP, NCHAN_PER_CHGROUP, NATIVE_SAMPLE_US, freq_GHz
from .dispersion import delta_tau_us, subband_shift_table

#: Band edges of the processed band (``common/constants.py:85-156``).
NU_CHGROUP_TOP_GHZ: Tuple[float, ...] = tuple(freq_GHz(g, 0) for g in range(N_CHGROUP))
NU_CHGROUP_BOT_GHZ: Tuple[float, ...] = tuple(freq_GHz(g, NCHAN_PER_CHGROUP - 1) for g in range(N_CHGROUP))
NU_BOT_PROC_GHZ: float = freq_GHz(N_CHGROUP - 1, NCHAN_PER_CHGROUP - 1)
#: Search-side sample period at the default operating point (16 native samples).
T_INT_SEARCH_US_DEFAULT: float = 16 * NATIVE_SAMPLE_US


@dataclass(frozen=True)
class TimeShiftSearchTable:
    """``shifts[f, stream]`` int32 (signed): stream ``g`` (or ``g * n_sub + s``)
    is read at ``t - shift`` to align it to the band bottom at fine DM ``f``
    (combiner.py:66-147)."""
    shifts: np.ndarray
    fine_to_coarse: np.ndarray
    t_int_search_us: float
    coarse_offset_baked: bool = False

    def __post_init__(self):
        if self.shifts.ndim != 2:
            raise ValueError(f"shifts must be 2D [N_fine, N_chgroup]; got shape {self.shifts.shape}")
        if self.shifts.shape[1] % N_CHGROUP != 0:
            raise ValueError(f"shifts.shape[1]={self.shifts.shape[1]} is not a multiple of N_CHGROUP={N_CHGROUP}")
        if self.shifts.dtype != np.int32:
            raise TypeError(f"shifts.dtype={self.shifts.dtype}, expected int32")
        if self.fine_to_coarse.ndim != 1:
            raise ValueError(f"fine_to_coarse must be 1D; got shape {self.fine_to_coarse.shape}")
        if self.fine_to_coarse.shape[0] != self.shifts.shape[0]:
            raise ValueError(f"fine_to_coarse.shape[0]={self.fine_to_coarse.shape[0]} != "
                             f"shifts.shape[0]={self.shifts.shape[0]} (N_fine)")
        if self.t_int_search_us <= 0:
            raise ValueError(f"t_int_search_us must be > 0; got {self.t_int_search_us!r}")


def _rint_samples(nu_low: float, nu_high: float, dm: float, t_us: float) -> float:
    return float(np.rint(delta_tau_us(float(nu_low), float(nu_high), float(dm)) / t_us))


def compute_time_shift_search(*, coarse_dm_pc_cm3: np.ndarray, fine_dm_pc_cm3: np.ndarray,
                              fine_to_coarse: np.ndarray, t_int_search_us: float = T_INT_SEARCH_US_DEFAULT,
                              nu_chgroup_bot_GHz: Optional[Tuple[float, ...]] = None,
                              nu_chgroup_top_GHz: Optional[Tuple[float, ...]] = None,
                              nu_bot_proc_GHz: float = NU_BOT_PROC_GHZ, include_coarse_offset: bool = False,
                              merge_coarse_rounding: bool = False, t_int_corr_us: Optional[float] = None,
                              n_sub: int = 1, nu_subband_ref_GHz: Optional[np.ndarray] = None
                              ) -> TimeShiftSearchTable:
    """Per-(fine DM, stream) shift table (combiner.py:150-491).  Each chgroup
    is referenced to its TOP channel; ``shift = rint(dtau(nu_bot_proc, ref, ddm) / T)``
    with ``ddm = fine - coarse`` (or ``fine`` with ``include_coarse_offset``);
    ``merge_coarse_rounding`` hands over ``rint(dtau(fine)/T) - rint(dtau(coarse)/T)``;
    ``n_sub > 1`` uses each subband's own reference for the fine term."""
    coarse = np.asarray(coarse_dm_pc_cm3, dtype=np.float64)
    fine = np.asarray(fine_dm_pc_cm3, dtype=np.float64)
    f2c = np.asarray(fine_to_coarse)
    if coarse.ndim != 1 or fine.ndim != 1:
        raise ValueError("coarse_dm and fine_dm must both be 1D")
    if f2c.ndim != 1:
        raise ValueError("fine_to_coarse must be 1D")
    if f2c.shape[0] != fine.shape[0]:
        raise ValueError(f"fine_to_coarse.shape[0]={f2c.shape[0]} != len(fine_dm)={fine.shape[0]}")
    if int(f2c.max()) >= len(coarse):
        raise ValueError(f"fine_to_coarse.max()={int(f2c.max())} >= len(coarse_dm)={len(coarse)}")
    bot = np.asarray(nu_chgroup_bot_GHz if nu_chgroup_bot_GHz is not None else NU_CHGROUP_BOT_GHZ, dtype=np.float64)
    if bot.shape != (N_CHGROUP,):
        raise ValueError(f"nu_chgroup_bot_GHz must be a {N_CHGROUP}-tuple; got shape {bot.shape}")
    ref = np.asarray(nu_chgroup_top_GHz if nu_chgroup_top_GHz is not None else NU_CHGROUP_TOP_GHZ, dtype=np.float64)
    if ref.shape != (N_CHGROUP,):
        raise ValueError(f"nu_chgroup_top_GHz must be a {N_CHGROUP}-tuple; got shape {ref.shape}")
    T = float(t_int_search_us)
    if merge_coarse_rounding:
        if include_coarse_offset:
            raise ValueError("merge_coarse_rounding is meaningless with include_coarse_offset=True")
        if t_int_corr_us is None:
            raise ValueError("merge_coarse_rounding=True requires t_int_corr_us")
        if abs(float(t_int_corr_us) - T) > 1e-9:
            raise ValueError(f"merge_coarse_rounding needs t_int_corr_us ({t_int_corr_us}) == "
                             f"t_int_search_us ({t_int_search_us})")
    n_fine = fine.shape[0]
    ns = int(n_sub)
    if ns > 1:
        if not merge_coarse_rounding:
            raise ValueError("n_sub > 1 requires merge_coarse_rounding=True")
        refs = np.asarray(nu_subband_ref_GHz, dtype=np.float64)
        if refs.shape != (N_CHGROUP, ns):
            raise ValueError(f"nu_subband_ref_GHz must be ({N_CHGROUP}, {ns}); got {refs.shape}")
        shifts = np.zeros((n_fine, N_CHGROUP * ns), dtype=np.int32)
        for f in range(n_fine):
            c = int(f2c[f])
            for g in range(N_CHGROUP):
                s_coarse = _rint_samples(nu_bot_proc_GHz, ref[g], coarse[c], T)
                for s in range(ns):
                    shifts[f, g * ns + s] = int(_rint_samples(nu_bot_proc_GHz, refs[g, s], fine[f], T) - s_coarse)
        return TimeShiftSearchTable(shifts=shifts, fine_to_coarse=f2c.astype(np.int64, copy=False),
                                    t_int_search_us=T, coarse_offset_baked=False)
    shifts = np.zeros((n_fine, N_CHGROUP), dtype=np.int32)
    for f in range(n_fine):
        c = int(f2c[f])
        ddm = float(fine[f]) if include_coarse_offset else float(fine[f] - coarse[c])
        for g in range(N_CHGROUP):
            if merge_coarse_rounding:
                shifts[f, g] = int(_rint_samples(nu_bot_proc_GHz, ref[g], fine[f], T)
                                   - _rint_samples(nu_bot_proc_GHz, ref[g], coarse[c], T))
            else:
                shifts[f, g] = int(_rint_samples(nu_bot_proc_GHz, ref[g], ddm, T))
    return TimeShiftSearchTable(shifts=shifts, fine_to_coarse=f2c.astype(np.int64, copy=False),
                                t_int_search_us=T, coarse_offset_baked=bool(include_coarse_offset))


def compute_delay_native_samples_table(coarse_dm: np.ndarray, chgroup_freqs_GHz: np.ndarray) -> np.ndarray:
    """``(G, C, K)`` int64 ``rint(dtau(nu[g, c], nu[g, 0], dm[k]) / 32.768 us)``,
    each chgroup referenced to its top channel (dm_plan.py:194-243)."""
    if coarse_dm.ndim != 1:
        raise ValueError(f"coarse_dm must be 1-D; got shape {coarse_dm.shape}")
    if chgroup_freqs_GHz.ndim != 2:
        raise ValueError(f"chgroup_freqs_GHz must be 2-D; got shape {chgroup_freqs_GHz.shape}")
    nu = chgroup_freqs_GHz.astype(np.float64)
    inv_diff = (1.0 / (nu ** 2)) - (1.0 / (nu[:, 0:1] ** 2))
    from .constants import K_DM_MS_GHZ2_PC
    delay_us = K_DM_MS_GHZ2_PC * coarse_dm.astype(np.float64)[None, None, :] * inv_diff[:, :, None] * 1e3
    return np.rint(delay_us / NATIVE_SAMPLE_US).astype(np.int64)


def build_time_shifts(fine_dm: np.ndarray, coarse_dm: np.ndarray, fine_to_coarse: np.ndarray,
                      t_int_fast_us: float, t_int_search_us: float):
    """The three tables of a DM plan file (tools/build_dm_plan.py:322-387):
    stage 1 ``[G, C, K]`` = rint(dtau(nu_bot[g], nu[g, c], coarse) / t_fast),
    stage 2 ``[G, K]`` = rint(dtau(nu_bot_proc, nu_bot[g], coarse) / t_fast),
    search ``[F, G]`` = rint(dtau(nu_bot_proc, nu_bot[g], fine - coarse) / t_search)."""
    fine = np.asarray(fine_dm, dtype=np.float64)
    coarse = np.asarray(coarse_dm, dtype=np.float64)
    f2c = np.asarray(fine_to_coarse)
    nu = np.array([[freq_GHz(g, c) for c in range(NCHAN_PER_CHGROUP)] for g in range(N_CHGROUP)])
    bot = np.asarray(NU_CHGROUP_BOT_GHZ, dtype=np.float64)
    s1 = np.empty((N_CHGROUP, NCHAN_PER_CHGROUP, coarse.size), dtype=np.int32)
    s2 = np.empty((N_CHGROUP, coarse.size), dtype=np.int32)
    for g in range(N_CHGROUP):
        for k in range(coarse.size):
            s2[g, k] = int(_rint_samples(NU_BOT_PROC_GHZ, bot[g], coarse[k], t_int_fast_us))
            for c in range(NCHAN_PER_CHGROUP):
                s1[g, c, k] = int(_rint_samples(bot[g], nu[g, c], coarse[k], t_int_fast_us))
    ddm = fine - coarse[f2c]
    ss = np.empty((fine.size, N_CHGROUP), dtype=np.int32)
    for f in range(fine.size):
        for g in range(N_CHGROUP):
            ss[f, g] = int(_rint_samples(NU_BOT_PROC_GHZ, bot[g], ddm[f], t_int_search_us))
    return s1, s2, ss


def node_shift_table(dm_pc_cm3: np.ndarray, nu_subband_ref_GHz: np.ndarray, nu_bot_GHz: float,
                     t_int_search_us: float) -> np.ndarray:
    """Single-node form used by the fused search: ``[n_dm, n_sub]`` int32
    ``rint(dtau(nu_bot, nu_ref[s], dm) / T)`` -- the ``n_sub > 1`` fine term of
    ``compute_time_shift_search`` with no corr-side coarse alignment."""
    return subband_shift_table(dm_pc_cm3, nu_subband_ref_GHz, nu_bot_GHz, t_int_search_us)


@dataclass(frozen=True)
class DedispPlan:
    """LDS layout of ``dsx_dedisp_detect`` for one DM plan and block length
    (``dsx_dedisp_plan``): host arrays, copied to the device by the caller."""
    tile: int
    n_rows: int
    n_terms: int
    r_total: int
    q_total: int
    paired: bool
    lds_bytes: int            # LDS per work-group (windows, pair rows, lane term bases, table)
    sub_info: np.ndarray      # [n_sub, 3] int32
    rows: np.ndarray          # [n_rows, 6] int32
    tbase: np.ndarray         # [n_terms, n_dm] int32
    dma: bool = False         # windows copied HBM -> LDS by LDS-direct loads during compute
    rows_bf16: bool = False   # subband-pair rows stored as BF16

    @property
    def flags(self) -> int:
        """``DSX_DD_*`` layout flags of ``dsx_dedisp_detect``."""
        return (1 if self.dma else 0) | (2 if self.rows_bf16 else 0)


def plan_dedisp(shift_sd: np.ndarray, *, n_widths: int, n_centres: int, ring_dtype: torch.dtype,
                lds_budget: int, mode: int = 0) -> DedispPlan:
    """Run ``dsx_dedisp_plan`` (host code, no GPU) for ``shift_sd`` = [n_sub, n_dm].
    ``mode`` 0 = pair rows when they fit, 1 = pair rows only, 2 = unpaired,
    3 = unpaired with two window buffers, the next tile's windows loaded from
    HBM straight into LDS while the current tile computes (float32 history;
    ``q_total`` = ``r_total`` marks the second buffer), 4 = pair rows whose
    offsets spread a wave's 64 lanes over the LDS banks, built from one of
    two window buffers while the next tile's windows load into the other
    (float32 history), 5 = pair rows as mode 1 stored as BF16 (the
    reduced-precision partial sums)."""
    sh = np.ascontiguousarray(shift_sd, dtype=np.int32)
    if sh.ndim != 2:
        raise ValueError("shift_sd must be [n_sub, n_dm]")
    n_sub, n_dm = sh.shape
    code = _lib.DSX_F16 if ring_dtype == torch.float16 else _lib.DSX_F32
    n_pair = (n_sub + 1) // 2
    rows_cap = int(sum(len(np.unique(sh[min(2 * g + 1, n_sub - 1)] - sh[2 * g])) for g in range(n_pair)))
    sub_info = np.zeros((n_sub, 3), np.int32)
    rows = np.zeros((max(rows_cap, 1), 6), np.int32)
    tbase = np.zeros((max(n_sub, n_pair), n_dm), np.int32)
    info = np.zeros(8, np.int32)
    P = _lib._P
    _lib.call("dsx_dedisp_plan", sh.ctypes.data_as(P), n_sub, n_dm, n_widths, n_centres, code, int(lds_budget),
              int(mode), sub_info.ctypes.data_as(P), rows.ctypes.data_as(P), rows_cap, tbase.ctypes.data_as(P),
              info.ctypes.data_as(P))
    tile, n_rows, n_terms, r_total, q_total, paired, lds, flags = (int(x) for x in info)
    return DedispPlan(tile=tile, n_rows=n_rows, n_terms=n_terms, r_total=r_total, q_total=q_total,
                      paired=bool(paired), lds_bytes=lds, sub_info=sub_info, rows=rows[:n_rows].copy(),
                      tbase=tbase[:n_terms].copy(), dma=bool(flags & 1), rows_bf16=bool(flags & 2))


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
