"""Sparsity pattern + sparse-cell gridder with the reference's stage APIs.

* ``build_pattern`` -- ``src/dsart/grid/sparsity_pattern.py:893-1218``:
  sorted unique ``(row, col)`` cells any (cross baseline, channel) lands in.
  Generalised beyond the 96-antenna / chgroup table: ``chan_freqs_GHz``
  overrides the chgroup's channel frequencies and ``grid_freq_GHz`` scales
  every channel at one frequency.
* ``FastVisGridder.from_pattern`` / ``.compute`` / ``.cell_weights`` --
  ``src/dsart/grid/kernel.py:234-611``: ``(n_fv, NBASE, NCHAN)`` visibilities
  -> ``(n_fv, n_filled)`` cell sums, run by the CSR gather kernel
  ``dsx_grid_gather`` (exact for int32 visibilities, CSR-order float sums for
  complex64).
"""
from __future__ import annotations

import hashlib
from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np
import torch

from . import _lib
from .constants import NCHAN_PER_CHGROUP, N_GRID_DEFAULT, freq_GHz
from .pattern import build_cell_map

SUPPORTED_KERNEL_SUPPORTS = (1,)
PATTERN_DEC_QUANT_DEG = 0.25


def quantise_dec_deg(dec_deg: float) -> float:
    return float(np.round(float(dec_deg) / PATTERN_DEC_QUANT_DEG) * PATTERN_DEC_QUANT_DEG)


def compute_antpos_hash(antpos_e: np.ndarray, antpos_n: np.ndarray) -> int:
    h = hashlib.blake2b(digest_size=8, person=b"dsx-antpos")
    h.update(np.ascontiguousarray(antpos_e, dtype=np.float64).tobytes())
    h.update(np.ascontiguousarray(antpos_n, dtype=np.float64).tobytes())
    return int.from_bytes(h.digest(), "little")


@dataclass(frozen=True)
class SparsityPattern:
    ix_row: np.ndarray          # (n_filled,) uint16, lexicographically sorted with ix_col
    ix_col: np.ndarray
    pattern_id: int
    n_grid: int
    n_filled: int
    chgroup: int
    dec_deg_quant: float
    kernel_support: int
    antpos_hash: int
    chan_sum_factor: int
    cell_lambda: float
    sub_band: Optional[Tuple[int, int]]
    freqs_GHz: np.ndarray       # gridding frequency of each summed channel of the pattern


def _pattern_freqs(chgroup, chan_sum_factor, chan_freqs_GHz, grid_freq_GHz):
    f = (np.asarray(chan_freqs_GHz, dtype=np.float64) if chan_freqs_GHz is not None
         else np.asarray([freq_GHz(chgroup, c) for c in range(NCHAN_PER_CHGROUP)], dtype=np.float64))
    if f.size % chan_sum_factor != 0:
        raise ValueError(f"chan_sum_factor={chan_sum_factor} does not divide nchan={f.size}")
    f = f if chan_sum_factor == 1 else f.reshape(-1, chan_sum_factor).mean(axis=1)
    if grid_freq_GHz is not None:
        f = np.full(f.size, float(grid_freq_GHz))
    return f


def build_pattern(antpos_e: np.ndarray, antpos_n: np.ndarray, *, chgroup: int = 0, dec_deg: float = 0.0,
                  n_grid: int = N_GRID_DEFAULT, kernel_support: int = 1, chan_sum_factor: int = 1,
                  cell_lambda: Optional[float] = None, is_core_baseline_mask: Optional[np.ndarray] = None,
                  antpos_hash: Optional[int] = None, sub_band: Optional[Tuple[int, int]] = None,
                  chan_freqs_GHz: Optional[np.ndarray] = None,
                  grid_freq_GHz: Optional[float] = None) -> SparsityPattern:
    """Deterministic sparsity pattern of one channel group (sparsity_pattern.py:893-1218)."""
    e = np.asarray(antpos_e)
    n = np.asarray(antpos_n)
    if e.ndim != 1 or e.shape != n.shape:
        raise ValueError(f"antpos shapes must be equal 1-D; got E={e.shape}, N={n.shape}")
    if n_grid <= 0 or n_grid > 65535:
        raise ValueError(f"n_grid={n_grid} out of range")
    if kernel_support not in SUPPORTED_KERNEL_SUPPORTS:
        raise ValueError(f"kernel_support={kernel_support} not in {SUPPORTED_KERNEL_SUPPORTS}")
    if chan_sum_factor < 1:
        raise ValueError(f"chan_sum_factor={chan_sum_factor}, expected >= 1")
    if cell_lambda is not None and not (np.isfinite(cell_lambda) and cell_lambda > 0.0):
        raise ValueError(f"cell_lambda={cell_lambda} must be finite and > 0")
    f = _pattern_freqs(chgroup, chan_sum_factor, chan_freqs_GHz, grid_freq_GHz)
    if sub_band is not None:
        s_idx, n_sub = int(sub_band[0]), int(sub_band[1])
        if n_sub < 1 or f.size % n_sub != 0 or not 0 <= s_idx < n_sub:
            raise ValueError(f"sub_band={sub_band}: need 0 <= s < n_sub and n_sub dividing nchan_eff={f.size}")
        if cell_lambda is None:
            raise ValueError("sub_band requires an explicit cell_lambda")
        per = f.size // n_sub
        f = f[s_idx * per:(s_idx + 1) * per]
    cm = build_cell_map(e, n, f, n_grid, cell_lambda=cell_lambda, is_core_baseline_mask=is_core_baseline_mask)
    ok = cm.row >= 0
    keys = (cm.row[ok].astype(np.uint32) << 16) | cm.col[ok].astype(np.uint32)
    uk = np.unique(keys)
    ah = compute_antpos_hash(e, n) if antpos_hash is None else int(antpos_hash)
    h = hashlib.blake2b(digest_size=8, person=b"dsx-pattern")
    h.update(repr((chgroup, quantise_dec_deg(dec_deg), n_grid, kernel_support, chan_sum_factor, ah,
                   float(cm.cell_lambda), sub_band, f.tobytes())).encode())
    return SparsityPattern(ix_row=((uk >> 16) & 0xFFFF).astype(np.uint16), ix_col=(uk & 0xFFFF).astype(np.uint16),
                           pattern_id=int.from_bytes(h.digest(), "little"), n_grid=n_grid, n_filled=int(uk.size),
                           chgroup=chgroup, dec_deg_quant=quantise_dec_deg(dec_deg), kernel_support=kernel_support,
                           antpos_hash=ah, chan_sum_factor=chan_sum_factor, cell_lambda=float(cm.cell_lambda),
                           sub_band=sub_band, freqs_GHz=f)


class FastVisGridder:
    """Sparse-cell gridder of one pattern (grid/kernel.py:108-611)."""

    def __init__(self, pattern: SparsityPattern, cell_index_map: np.ndarray, cell_weights_cpu: torch.Tensor,
                 device="cuda"):
        self.pattern = pattern
        self.device = torch.device(device)
        if self.device.type == "cuda" and self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device() if torch.cuda.is_available() else 0)
        cim = np.asarray(cell_index_map, dtype=np.int64)
        nf = pattern.n_filled
        if cim.min() < 0 or cim.max() > nf:
            raise ValueError("cell_index_map values must lie in [0, n_filled] (n_filled = dropped)")
        self.cell_index_map = torch.from_numpy(cim)
        self.cell_weights_cpu = cell_weights_cpu
        # CSR over cells: sources (bls * nchan + ch) sorted by cell then source
        src = np.nonzero(cim < nf)[0]
        order = np.lexsort((src, cim[src]))
        src = src[order]
        counts = np.bincount(cim[src], minlength=nf)
        self._ptr = torch.from_numpy(np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)).to(self.device)
        self._src = torch.from_numpy(src.astype(np.int32)).to(self.device)
        self._n_src = int(cim.size)
        self._cell_weights_device = cell_weights_cpu.to(self.device)

    @classmethod
    def from_pattern(cls, pattern: SparsityPattern, antpos_e: np.ndarray, antpos_n: np.ndarray, *,
                     is_core_baseline_mask: Optional[np.ndarray] = None, device="cuda") -> "FastVisGridder":
        if compute_antpos_hash(antpos_e, antpos_n) != pattern.antpos_hash:
            raise ValueError("the antpos arrays passed to FastVisGridder.from_pattern must match build_pattern")
        cm = build_cell_map(antpos_e, antpos_n, pattern.freqs_GHz, pattern.n_grid,
                            cell_lambda=pattern.cell_lambda, is_core_baseline_mask=is_core_baseline_mask)
        keys_pat = (pattern.ix_row.astype(np.uint32) << 16) | pattern.ix_col.astype(np.uint32)
        ok = cm.row >= 0
        keys = (cm.row.astype(np.int64).clip(0) << 16) | cm.col.astype(np.int64).clip(0)
        pos = np.searchsorted(keys_pat, keys.astype(np.uint32))
        pos = np.where(ok, pos, pattern.n_filled)
        if ok.any() and (pos[ok].max() >= pattern.n_filled or
                         not np.array_equal(keys_pat[pos[ok]], keys[ok].astype(np.uint32))):
            raise RuntimeError("pattern / gridder geometry drift")
        cim = pos.reshape(-1)                                    # (NBASE * NCHAN_eff,)
        w = np.bincount(cim[cim < pattern.n_filled], minlength=pattern.n_filled).astype(np.float32)
        return cls(pattern, cim, torch.from_numpy(w), device=device)

    @property
    def cell_weights(self) -> torch.Tensor:
        """Per-cell sample count (sum of squared unit tap weights)."""
        return self._cell_weights_device

    def compute(self, vis_stokes_i: torch.Tensor) -> torch.Tensor:
        """``(n_fv, NBASE, NCHAN)`` complex (or ``(..., 2)`` int32) -> ``(n_fv, n_filled)``
        complex64 (or int32 pairs, exact)."""
        is_int = vis_stokes_i.dtype == torch.int32
        if not is_int and not vis_stokes_i.is_complex():
            raise TypeError(f"vis_stokes_i must be complex or int32 pairs; got {vis_stokes_i.dtype}")
        nd = 4 if is_int else 3
        if vis_stokes_i.ndim != nd or (is_int and vis_stokes_i.shape[-1] != 2):
            raise ValueError(f"vis_stokes_i must be (n_fast_vis, NBASE, NCHAN){' + (2,)' if is_int else ''}; "
                             f"got {tuple(vis_stokes_i.shape)}")
        n_fv = int(vis_stokes_i.shape[0])
        if int(vis_stokes_i.shape[1]) * int(vis_stokes_i.shape[2]) != self._n_src:
            raise ValueError(f"vis_stokes_i.shape {tuple(vis_stokes_i.shape)} does not match the pattern")
        if vis_stokes_i.device != self.device:
            raise ValueError(f"vis_stokes_i on {vis_stokes_i.device}, gridder on {self.device}")
        nf = self.pattern.n_filled
        if is_int:
            src = vis_stokes_i.contiguous()
            out = torch.empty((n_fv, nf, 2), dtype=torch.int32, device=self.device)
            code = _lib.DSX_I32
        else:
            src = torch.view_as_real(vis_stokes_i.to(torch.complex64).contiguous())
            out = torch.empty((n_fv, nf, 2), dtype=torch.float32, device=self.device)
            code = _lib.DSX_F32
        _lib.require_gpu(src, out)
        _lib.call("dsx_grid_gather", _lib.ptr(src), _lib.ptr(self._ptr), _lib.ptr(self._src), _lib.ptr(out),
                  n_fv, self._n_src, nf, code, _lib.stream())
        return out if is_int else torch.view_as_complex(out)
