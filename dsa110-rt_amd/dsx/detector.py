"""Detector contracts and stages: ``Candidate``, boxcars, multi-width argmax, decode.

* ``Candidate`` / ``CandidateFlags`` keep the fields of
  ``src/dsart/common/contracts.py:73-83,256-301``; ``time_bin`` (the
  dedispersed step) is added with a default, which the contract allows.
* ``boxcar_via_cumsum`` and ``multi_boxcar_argmax`` keep the signatures of
  ``src/dsart/detector/forward.py:163-169`` and
  ``src/dsart/detector/triton_boxcar.py:305-314`` and run HIP kernels.
* ``decode_chunk_maxima`` is the top-k + merge-radius decoder of
  ``src/dsart/detector/decoder.py:388-533`` applied to the per-(DM, chunk,
  pixel) maxima the fused dedispersion kernel emits.
"""
from __future__ import annotations

import enum
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib

DETECTOR_VERSION = "dsx.v1"


class CandidateFlags(enum.IntFlag):
    """``Candidate.flags`` bits (contracts.py:73-83)."""

    NONE = 0
    NOISE_WARMUP = 1 << 3
    RFI_WARMING_UP = 1 << 4
    HALO_DROPPED = 1 << 5
    TIME_EDGE_DROPPED = 1 << 6


@dataclass(frozen=True)
class Candidate:
    """Detector output (contracts.py:256-301) plus the dedispersed step."""

    l: float
    m: float
    dm_fine: float
    dm_idx: int
    event_specnum: int
    width_samples: int
    kernel_id: str
    snr: float
    detector_version: str
    flags: int
    search_node_id: int
    gpu_half: int
    time_bin: int = -1

    @property
    def key(self) -> Tuple[int, int, int, int, int]:
        """Candidate identity used to compare candidate lists: (DM index, time bin, width, row, col)."""
        return (self.dm_idx, self.time_bin, self.width_samples, int(self.l), int(self.m))


#: 16-byte device record of ``dsx_dedisp_detect``: time relative to the block's first centre.
CAND_DTYPE = np.dtype([("t", "<i4"), ("p", "<i4"), ("d", "<i2"), ("k", "<i2"), ("snr", "<f4")])
#: host record: absolute dedispersed step (int64, no wrap however long the stream runs).
REC_DTYPE = np.dtype([("t", "<i8"), ("p", "<i4"), ("d", "<i2"), ("k", "<i2"), ("snr", "<f4")])

#: thresholds an overflowed block may be detected again at
N_LEVELS = 16


def record_levels(threshold: float) -> np.ndarray:
    """Overflow ladder: ``threshold * sqrt(2)**j``, ``j < 16``, in float32
    (the kernel compares float32 S/N against a float32 threshold)."""
    thr = np.float32(threshold)
    odd = np.float32(thr * np.float32(1.41421356))
    return np.array([np.ldexp(odd if j & 1 else thr, j >> 1) for j in range(N_LEVELS)], dtype=np.float32)


def device_records(rec_dev: np.ndarray, c_lo: int) -> np.ndarray:
    """Device records (``CAND_DTYPE``, time relative to ``c_lo``) -> host
    records (``REC_DTYPE``, absolute int64 time)."""
    out = np.empty(rec_dev.size, dtype=REC_DTYPE)
    out["t"] = rec_dev["t"].astype(np.int64) + int(c_lo)
    for f in ("p", "d", "k", "snr"):
        out[f] = rec_dev[f]
    return out


def decode_chunk_maxima(rec: np.ndarray, *, widths: Sequence[int], n_grid: int,
                        dm_trials: np.ndarray, threshold: float, n_top: int, radius_t: int,
                        radius_dm: int, radius_lm: int, t_int: int,
                        search_node_id: int = 0, gpu_half: int = 0) -> List[Candidate]:
    """Top-k over above-threshold chunk maxima, ordered by (-snr, t, dm, pixel)
    (the flat-index order of decoder.py:441-443 on a [T, F, H, W] cube), then
    greedy suppression of anything within ``|dt| <= radius_t``,
    ``|d dm| <= radius_dm``, ``|d row|, |d col| <= radius_lm`` of an accepted
    candidate (decoder.py:471-499)."""
    if rec.size == 0:
        return []
    rec = rec[rec["snr"] > threshold]
    if rec.size == 0:
        return []
    order = np.lexsort((rec["p"], rec["d"], rec["t"], -rec["snr"].astype(np.float64)))[: int(n_top)]
    acc: List[Tuple[int, int, int, int]] = []
    out: List[Candidate] = []
    for o in order:
        t, d, p = int(rec["t"][o]), int(rec["d"][o]), int(rec["p"][o])
        r, c = divmod(p, n_grid)
        if any(abs(t - a) <= radius_t and abs(d - b) <= radius_dm and abs(r - x) <= radius_lm
               and abs(c - y) <= radius_lm for a, b, x, y in acc):
            continue
        acc.append((t, d, r, c))
        w = int(widths[int(rec["k"][o])])
        out.append(Candidate(
            l=float(r), m=float(c), dm_fine=float(dm_trials[d]), dm_idx=d,
            event_specnum=t * (t_int // 2), width_samples=w, kernel_id=f"unit:d1:b{w}",
            snr=float(rec["snr"][o]), detector_version=DETECTOR_VERSION,
            flags=int(CandidateFlags.NONE), search_node_id=search_node_id, gpu_half=gpu_half,
            time_bin=t))
    return out


# ---------------------------------------------------------------------------
# reference detector primitives on the GPU
# ---------------------------------------------------------------------------


def boxcar_via_cumsum(x: torch.Tensor, *, axis: int, width: int,
                      tile_size: Optional[int] = None) -> torch.Tensor:
    """Centred unweighted boxcar ``out[i] = sum x[i - w//2 : i - w//2 + w]``
    along ``axis`` with zeros outside (forward.py:163-247).  ``width == 1``
    returns ``x`` itself.  float32 / float16 in, same dtype out; summed in
    float32 on the GPU.  ``tile_size`` is accepted for signature
    compatibility (the kernel streams columns and needs no tiling)."""
    if width < 1:
        raise ValueError(f"width={width}, expected >= 1")
    if width == 1:
        return x
    n = x.shape[axis]
    if width > n:
        raise ValueError(f"width={width} exceeds axis-{axis} length {n}; cannot apply centred boxcar")
    _lib.require_gpu(x.contiguous())
    xm = torch.movedim(x, axis, 0).contiguous()
    xf = xm.to(torch.float32).reshape(n, -1).contiguous()
    out = torch.empty_like(xf)
    _lib.call("dsx_boxcar", _lib.ptr(xf), _lib.ptr(out), n, xf.shape[1], int(width), _lib.stream())
    out = out.reshape(xm.shape).to(x.dtype)
    return torch.movedim(out, 0, axis).contiguous()


def multi_boxcar_argmax(cs: torch.Tensor, *, widths: torch.Tensor, offsets: torch.Tensor,
                        sigma_inv: torch.Tensor, n_out: int, t_base: int = 0,
                        out_max_dtype: Optional[torch.dtype] = None):
    """Fused multi-width boxcar argmax over a padded cumsum
    (triton_boxcar.py:305-403): for every output cell keep
    ``max_k (cs[t+o_k+w_k] - cs[t+o_k]) * sigma_inv[k]`` (strict ``>``, ties
    to the lowest k) and its winner.  Returns ``(out_max, out_winner int16)``."""
    if cs.dim() != 4:
        raise ValueError(f"cs must be 4-D [Tpad, F, H, W]; got {cs.dim()}-D")
    K = int(widths.numel())
    if not 1 <= K <= 32:
        raise ValueError(f"K={K} widths; supported 1..32")
    if int(offsets.numel()) != K or int(sigma_inv.numel()) != K:
        raise ValueError("widths, offsets and sigma_inv must have the same length")
    if widths.dtype != torch.int32 or offsets.dtype != torch.int32 or sigma_inv.dtype != torch.float32:
        raise TypeError("widths/offsets must be int32 and sigma_inv float32")
    if cs.dtype not in (torch.float16, torch.float32):
        raise TypeError(f"cs dtype {cs.dtype} not supported")
    eff = out_max_dtype if out_max_dtype is not None else cs.dtype
    if eff not in (torch.float16, torch.float32):
        raise TypeError(f"out_max_dtype {eff} not supported")
    _lib.require_gpu(cs, widths, offsets, sigma_inv)
    tpad = int(cs.shape[0])
    w_h = widths.cpu().numpy()
    o_h = offsets.cpu().numpy()
    if (t_base + n_out - 1 + (o_h + w_h).max()) >= tpad or (t_base + o_h.min()) < 0:
        raise ValueError("Requested (t_base, n_out, widths) window exceeds padded cumsum bounds")
    inner = int(np.prod(cs.shape[1:]))
    out_max = torch.empty((n_out,) + tuple(cs.shape[1:]), dtype=eff, device=cs.device)
    out_win = torch.empty((n_out,) + tuple(cs.shape[1:]), dtype=torch.int16, device=cs.device)
    _lib.call("dsx_multi_boxcar_argmax", _lib.ptr(cs),
              _lib.DSX_F16 if cs.dtype == torch.float16 else _lib.DSX_F32,
              _lib.ptr(widths), _lib.ptr(offsets), _lib.ptr(sigma_inv), K, _lib.ptr(out_max),
              _lib.DSX_F16 if eff == torch.float16 else _lib.DSX_F32, _lib.ptr(out_win), tpad,
              int(n_out), inner, int(t_base), _lib.stream())
    return out_max, out_win
