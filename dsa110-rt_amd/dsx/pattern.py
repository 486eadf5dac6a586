"""uv-cell geometry of the nearest-cell Stokes-I gridder (host side, once per pattern).

Follows the arithmetic of ``src/dsart/grid/sparsity_pattern.py``:
baselines in upper-triangle order ``bls = a (a + 1) / 2 + b`` with ``b <= a``
(``_per_baseline_uv_meters`` ``:657-714``), ``du = E[a] - E[b]``,
``u = -(du / lambda)`` (the F20 negation, ``:1086-1087``), ``cell_lambda``
auto-fit to ``2 max(|u|, |v|) / N`` (``:1118-1121``), and cell indices
``col = rint(u / cell) + N/2``, ``row = rint(v / cell) + N/2`` with
out-of-grid cells dropped (``:1168-1191``).  The same float64 operation order
is kept so ``np.rint`` lands on the same cell at rounding boundaries.

Everything here is host-side table building; the per-step gridding runs on
the GPU (``dsx.imager``, ``dsx.grid``).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np

from .constants import SPEED_OF_LIGHT_M_S, nbase

#: Flag bits of an imager tap's destination word.
TAP_CONJ_BIT = 1 << 20
TAP_NEG_BIT = 1 << 21
TAP_CELL_MASK = (1 << 20) - 1


def upper_tri_indices(nants: int) -> Tuple[np.ndarray, np.ndarray]:
    """``(a_idx, b_idx)`` with ``b <= a`` in ``bls = a(a+1)/2 + b`` order.

    Same order as ``src/dsart/services/slow_corr_kernel.py:161-178``.
    """
    a_idx = np.repeat(np.arange(nants, dtype=np.int64), np.arange(1, nants + 1))
    b_idx = np.concatenate([np.arange(a + 1, dtype=np.int64) for a in range(nants)])
    return a_idx, b_idx


def baseline_uv_lambda(
    antpos_e: np.ndarray,
    antpos_n: np.ndarray,
    freqs_GHz: np.ndarray,
    *,
    is_core_baseline_mask: Optional[np.ndarray] = None,
) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """``(kept_bls, u_lam, v_lam)`` for kept cross baselines x channels.

    ``u_lam``/``v_lam`` are ``(n_kept, n_chan)`` float64 already negated (F20).
    """
    e = np.asarray(antpos_e)
    n = np.asarray(antpos_n)
    nants = e.shape[0]
    a_idx, b_idx = upper_tri_indices(nants)
    keep = a_idx != b_idx
    if is_core_baseline_mask is not None:
        m = np.asarray(is_core_baseline_mask, dtype=bool)
        if m.shape != (nbase(nants),):
            raise ValueError(f"is_core_baseline_mask shape {m.shape} != ({nbase(nants)},)")
        keep = keep & m
    kept = np.nonzero(keep)[0]
    a_k, b_k = a_idx[kept], b_idx[kept]
    du_m = e[a_k].astype(np.float64) - e[b_k].astype(np.float64)
    dv_m = n[a_k].astype(np.float64) - n[b_k].astype(np.float64)
    nu = np.asarray(freqs_GHz, dtype=np.float64)
    wavelength_m = SPEED_OF_LIGHT_M_S / (nu * 1e9)
    u_lam = du_m[:, None] / wavelength_m[None, :]
    v_lam = dv_m[:, None] / wavelength_m[None, :]
    u_lam = -u_lam
    v_lam = -v_lam
    return kept, u_lam, v_lam


def auto_cell_lambda(u_lam: np.ndarray, v_lam: np.ndarray, n_grid: int) -> float:
    """``2 max(|u|, |v|) / N`` (sparsity_pattern.py:1099-1121)."""
    max_baseline_lambda = float(np.max(np.maximum(np.abs(u_lam), np.abs(v_lam))))
    if max_baseline_lambda == 0.0:
        raise ValueError("max_baseline_lambda == 0 (degenerate antpos / mask)")
    return max_baseline_lambda * 2.0 / n_grid


@dataclass(frozen=True)
class CellMap:
    """Per ``(bls, output channel)`` reference-layout cell, or -1 when dropped."""

    row: np.ndarray          # (nbase, n_chan_out) int32
    col: np.ndarray          # (nbase, n_chan_out) int32
    cell_lambda: float
    n_grid: int

    @property
    def valid(self) -> np.ndarray:
        return self.row >= 0


def build_cell_map(
    antpos_e: np.ndarray,
    antpos_n: np.ndarray,
    freqs_GHz: np.ndarray,
    n_grid: int,
    *,
    cell_lambda: Optional[float] = None,
    is_core_baseline_mask: Optional[np.ndarray] = None,
) -> CellMap:
    """Nearest cell of every (baseline, channel); autos/outriggers/out-of-grid -> -1."""
    nants = np.asarray(antpos_e).shape[0]
    kept, u_lam, v_lam = baseline_uv_lambda(
        antpos_e, antpos_n, freqs_GHz, is_core_baseline_mask=is_core_baseline_mask)
    if cell_lambda is None:
        cell = auto_cell_lambda(u_lam, v_lam, n_grid)
    else:
        cell = float(cell_lambda)
        if not (np.isfinite(cell) and cell > 0.0):
            raise ValueError(f"cell_lambda={cell_lambda} must be finite and > 0")
    half = n_grid // 2
    col = np.rint(u_lam / cell).astype(np.int64) + half
    row = np.rint(v_lam / cell).astype(np.int64) + half
    in_grid = (row >= 0) & (row < n_grid) & (col >= 0) & (col < n_grid)
    nb = nbase(nants)
    nch = u_lam.shape[1]
    row_full = np.full((nb, nch), -1, dtype=np.int32)
    col_full = np.full((nb, nch), -1, dtype=np.int32)
    r = np.where(in_grid, row, -1).astype(np.int32)
    c = np.where(in_grid, col, -1).astype(np.int32)
    row_full[kept] = r
    col_full[kept] = c
    return CellMap(row=row_full, col=col_full, cell_lambda=cell, n_grid=n_grid)


@dataclass(frozen=True)
class ImagerTaps:
    """Half-plane scatter taps of every subband, CSR over subbands.

    For a visibility ``v`` at reference cell ``(row, col)``, physical
    ``U = row - N/2``, ``W = col - N/2`` and sign ``(-1)^(row + col)``:
    tap 1 adds ``sign * v`` to ``H[U mod N][W mod N]`` when ``U mod N <= N/2``;
    tap 2 adds ``sign * conj(v)`` to ``H[-U mod N][-W mod N]`` when
    ``-U mod N <= N/2``.  ``H`` is then the Hermitian fold of the
    sign-modulated grid, so ``Re(ifft2)`` of the reference grid after
    ``fftshift`` and the signed edge mask equals
    ``mask01 * C2R(H) / (2 N^2)``.
    """

    src: np.ndarray       # (n_taps,) int32 index into one (step, subband) slab of (per_sub, nbase)
    dst: np.ndarray       # (n_taps,) int32: cell = r * N + w, | TAP_CONJ_BIT | TAP_NEG_BIT
    offsets: np.ndarray   # (n_sub + 1,) int32 CSR over subbands
    n_rows: int           # rows r in [0, n_rows) carry taps
    n_grid: int
    per_sub: int


def build_imager_taps(cmap: CellMap, n_sub: int) -> ImagerTaps:
    n = cmap.n_grid
    nb, nch = cmap.row.shape
    if nch % n_sub != 0:
        raise ValueError("n_sub must divide the number of output channels")
    per = nch // n_sub
    srcs, dsts, offs = [], [], [0]
    max_r = 0
    for s in range(n_sub):
        s_src, s_dst = [], []
        for cl in range(per):
            c = s * per + cl
            rows = cmap.row[:, c].astype(np.int64)
            cols = cmap.col[:, c].astype(np.int64)
            ok = np.nonzero(rows >= 0)[0]
            if ok.size == 0:
                continue
            rr, cc = rows[ok], cols[ok]
            u = rr - n // 2
            w = cc - n // 2
            neg = ((rr + cc) & 1).astype(np.int64) * TAP_NEG_BIT
            src = cl * nb + ok
            r1 = np.mod(u, n)
            w1 = np.mod(w, n)
            m1 = r1 <= n // 2
            r2 = np.mod(-u, n)
            w2 = np.mod(-w, n)
            m2 = r2 <= n // 2
            s_src.append(src[m1])
            s_dst.append(r1[m1] * n + w1[m1] + neg[m1])
            s_src.append(src[m2])
            s_dst.append(r2[m2] * n + w2[m2] + neg[m2] + TAP_CONJ_BIT)
            if m1.any():
                max_r = max(max_r, int(r1[m1].max()))
            if m2.any():
                max_r = max(max_r, int(r2[m2].max()))
        if s_src:
            src_s = np.concatenate(s_src)
            dst_s = np.concatenate(s_dst)
            order = np.argsort(src_s, kind="stable")
            srcs.append(src_s[order])
            dsts.append(dst_s[order])
            offs.append(offs[-1] + src_s.size)
        else:
            offs.append(offs[-1])
    src = np.concatenate(srcs).astype(np.int32) if srcs else np.zeros(0, np.int32)
    dst = np.concatenate(dsts).astype(np.int32) if dsts else np.zeros(0, np.int32)
    return ImagerTaps(src=src, dst=dst, offsets=np.asarray(offs, dtype=np.int32),
                      n_rows=max_r + 1, n_grid=n, per_sub=per)
