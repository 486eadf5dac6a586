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
import math
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from .merger import MergerConfig, merge_across_kernels_c1

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
                        radius_dm: int, radius_lm: int, specnum_start: int = 0,
                        search_node_id: int = 0, gpu_half: int = 0) -> List[Candidate]:
    """Top-k over above-threshold chunk maxima, ordered by (-snr, t, dm, pixel)
    (the flat-index order of decoder.py:441-443 on a [T, F, H, W] cube), then
    greedy suppression of anything within ``|dt| <= radius_t``,
    ``|d dm| <= radius_dm``, ``|d row|, |d col| <= radius_lm`` of an accepted
    candidate (decoder.py:471-499).  ``event_specnum`` counts search samples
    (dedispersed steps) from ``specnum_start``, as ``decoder.py:216`` does."""
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
            event_specnum=int(specnum_start) + t, width_samples=w, kernel_id=f"unit:d1:b{w}",
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


# ---------------------------------------------------------------------------
# Detector over a materialised cube (the Detector protocol drop-in)
# ---------------------------------------------------------------------------

DETECTOR_K_TIME_WIDTHS = (1, 2, 4, 8, 16, 32, 64, 128)
DEFAULT_MERGE_RADIUS_LM = 3
DEFAULT_MERGE_RADIUS_FDM = 5
DEFAULT_MERGE_RADIUS_T = 128
CUBE_CADENCE_S_DEFAULT = 256 * 524.288e-6

#: 16-byte cube candidate record of ``dsx_cube_candidates``
CUBE_REC_DTYPE = np.dtype([("t", "<i4"), ("cell", "<i4"), ("k", "<i4"), ("snr", "<f4")])


@dataclass(frozen=True)
class Kernel:
    """One detector kernel (the v1 bank: unit image kernel, unit DM width,
    a time boxcar), ``src/dsart/detector/forward.py`` kernel records."""
    kernel_id: str
    k_time_width: int
    k_dm_width: int = 1
    image_kernel_size: int = 1


class Layer2State:
    """Per-kernel noise sigma_k: running mean over the first ``n_burnin``
    cubes, then an EMA with ``gamma = 1 - exp(-cadence / tau)``; zero
    estimates keep the old value; optional ceiling ``sigma_max_ratio * s_k``
    with the clamp-escape hatch and a floor; invalid cubes leave the state
    alone (``src/dsart/noise_norm/layer2.py:190-452``).  Host float64."""

    def __init__(self, n_kernels: int, *, cube_cadence_s: float = CUBE_CADENCE_S_DEFAULT, tau_s: float = 30.0,
                 n_burnin: int = 30, sigma_floor: float = 0.0, sigma_max_ratio: float = 0.0,
                 clamp_escape_cubes: int = 0):
        if n_kernels < 1:
            raise ValueError(f"n_kernels={n_kernels}, expected >= 1")
        if tau_s <= 0:
            raise ValueError(f"tau_s={tau_s}, expected > 0")
        if cube_cadence_s <= 0:
            raise ValueError(f"cube_cadence_s={cube_cadence_s}, expected > 0")
        if n_burnin < 0:
            raise ValueError(f"n_burnin={n_burnin}, expected >= 0")
        self.n_kernels = int(n_kernels)
        self.gamma = 1.0 - math.exp(-float(cube_cadence_s) / float(tau_s))
        self.n_burnin = int(n_burnin)
        self.sigma_floor = float(sigma_floor)
        self.sigma_max_ratio = float(sigma_max_ratio)
        self.clamp_escape_cubes = int(clamp_escape_cubes)
        self._s_k = np.ones(self.n_kernels)
        self._streak = np.zeros(self.n_kernels, dtype=np.int64)
        self._cube_count = 0

    @property
    def s_k(self) -> np.ndarray:
        return self._s_k.copy()

    @property
    def cube_count(self) -> int:
        return self._cube_count

    @property
    def is_warming_up(self) -> bool:
        return self._cube_count < self.n_burnin

    def seed(self, s_k) -> None:
        self._s_k = np.asarray(s_k, dtype=np.float64).reshape(self.n_kernels).copy()

    def reset(self) -> None:
        self._s_k = np.ones(self.n_kernels)
        self._streak[:] = 0
        self._cube_count = 0

    def update_and_query(self, *, per_kernel_sigma, valid: bool = True) -> Tuple[np.ndarray, bool]:
        if not valid:
            return self.s_k, self.is_warming_up
        x = np.asarray(per_kernel_sigma, dtype=np.float64).reshape(-1).copy()
        if x.shape != (self.n_kernels,):
            raise ValueError(f"per-kernel sigma shape {x.shape} != (n_kernels={self.n_kernels},)")
        x = np.where(x == 0, self._s_k, x)
        if self.sigma_max_ratio > 0 and self._cube_count > 0:
            high = x > self.sigma_max_ratio * self._s_k
            if high.any():
                esc = (high & (self._streak >= self.clamp_escape_cubes)) if self.clamp_escape_cubes > 0 \
                    else np.zeros_like(high)
                clamp = high & ~esc
                self._s_k = np.where(esc, x, self._s_k)
                x = np.where(clamp, self._s_k, x)
                self._streak = np.where(clamp, self._streak + 1, 0)
            else:
                self._streak[:] = 0
        if self._cube_count < self.n_burnin:
            c = float(self._cube_count)
            self._s_k = (c * self._s_k + x) / (c + 1.0)
        else:
            self._s_k = self.gamma * x + (1.0 - self.gamma) * self._s_k
        if self.sigma_floor > 0:
            self._s_k = np.maximum(self._s_k, self.sigma_floor)
        self._cube_count += 1
        return self.s_k, self.is_warming_up


def sigma_clipped_std_host(samples: np.ndarray, n_sigma: float = 3.0, n_iterations: int = 3) -> float:
    """sigma-clipped standard deviation of one sample (layer1.py:53-122):
    clip about the lower median (``torch.nanmedian``), ``n_iterations``
    passes at ``n_sigma``; NaN samples are ignored.  Host code on the
    gathered sample (at most ``sigma_max_samples`` values per kernel)."""
    s = np.asarray(samples, dtype=np.float64).reshape(-1)
    active = np.isfinite(s)

    def stats(act):
        v = np.sort(s[act])
        if v.size == 0:
            return float("nan"), float("nan")
        med = float(v[(v.size - 1) // 2])
        return med, math.sqrt(float(np.sum((v - med) ** 2)) / v.size)

    med, sig = stats(active)
    for _ in range(int(n_iterations)):
        if not np.isfinite(sig):
            break
        active = active & (s >= med - n_sigma * sig) & (s <= med + n_sigma * sig)
        med, sig = stats(active)
    return 0.0 if not np.isfinite(sig) else sig


class DeterministicDetector:
    """Detector protocol drop-in (``src/dsart/detector/forward.py:117-155``,
    ``DeterministicDetector.forward`` ``:891-1009`` in its streaming
    top-k form ``:1264-1854``) over a materialised ``[T_det, N_fdm, H, W]``
    cube on the GPU.

    Per cube: the padded time prefix of the cube (``dsx_cube_cumsum``);
    centred boxcars of the bank's time widths scored against the Layer-2
    sigma_k of *earlier* cubes, the maximum over kernels per cell over the
    interior centres ``[w_max/2, T - w_max/2)``, and the ``n_top`` strongest
    cells above threshold (``dsx_cube_candidates``: queue + radix select on
    the GPU); then on the host the merge-radius suppression of
    ``decode_topk_argmax_lowmem`` (``decoder.py:388-533``), the
    ``NOISE_WARMUP`` flag during burn-in and the canonical fine-DM range
    gate.  sigma_k is updated at the end of the cube (when at least
    ``layer2_valid_min_fraction`` of the validity mask is True) from the
    sigma-clipped std of each kernel's raw box sums at a fixed seeded sample
    of interior cells (``dsx_cube_box_samples``), the reference's
    ``layer2_sigma_max_samples`` path.  ``sigma_layer1`` is accepted and,
    as in v1, unused."""

    def __init__(self, *, time_widths: Sequence[int] = DETECTOR_K_TIME_WIDTHS, threshold_sigma: float = 8.0,
                 detector_version: str = "v1.M5", device=None, merge_radius_lm: int = DEFAULT_MERGE_RADIUS_LM,
                 merge_radius_fdm: int = DEFAULT_MERGE_RADIUS_FDM, merge_radius_t: int = DEFAULT_MERGE_RADIUS_T,
                 search_node_id: int = 0, gpu_half: int = 0, cube_cadence_s: float = CUBE_CADENCE_S_DEFAULT,
                 layer2_tau_s: float = 30.0, layer2_n_burnin: int = 30, n_kernel_max_t: Optional[int] = None,
                 layer2_seed_unit: bool = True, layer2_sigma_floor: float = 0.0,
                 layer2_sigma_max_ratio: float = 0.0, layer2_clamp_escape_cubes: int = 0,
                 layer2_valid_min_fraction: float = 1.0, layer2_sigma_max_samples: int = 1 << 16,
                 streaming_decoder_n_top: int = 64, spatial_active: Optional[np.ndarray] = None,
                 queue_records: int = 1 << 20, merger_config: Optional[MergerConfig] = None,
                 c1_snr_min: Optional[float] = None):
        ws = [int(w) for w in time_widths]
        if c1_snr_min is not None:                       # the C1 floor is the decode threshold
            threshold_sigma = float(c1_snr_min)
        self._merger_config = merger_config
        self._c1_snr_min = None if c1_snr_min is None else float(c1_snr_min)
        if not ws or any(w < 1 for w in ws) or len(ws) > 16:
            raise ValueError(f"time_widths={tuple(ws)}: 1-16 positive widths")
        self._kernel_bank = tuple(Kernel(kernel_id=f"unit:d1:b{w}", k_time_width=w) for w in ws)
        self._widths = np.asarray(ws, dtype=np.int32)
        self._threshold = float(threshold_sigma)
        self._detector_version = detector_version
        self._device = torch.device(device) if device is not None else torch.device("cuda")
        self._radii = (int(merge_radius_t), int(merge_radius_fdm), int(merge_radius_lm))
        if min(self._radii) < 0:
            raise ValueError(f"merge radii must be >= 0; got {self._radii}")
        self._search_node_id = int(search_node_id)
        self._gpu_half = int(gpu_half)
        self._n_kernel_max_t = int(n_kernel_max_t) if n_kernel_max_t is not None else max(ws)
        self._layer2 = Layer2State(len(ws), cube_cadence_s=cube_cadence_s, tau_s=layer2_tau_s,
                                   n_burnin=layer2_n_burnin, sigma_floor=layer2_sigma_floor,
                                   sigma_max_ratio=layer2_sigma_max_ratio,
                                   clamp_escape_cubes=layer2_clamp_escape_cubes)
        if layer2_seed_unit:
            self._layer2.seed([math.sqrt(k.k_dm_width * k.k_time_width) for k in self._kernel_bank])
        self._valid_min_fraction = float(layer2_valid_min_fraction)
        self._n_samples = int(layer2_sigma_max_samples)
        self._n_top = int(streaming_decoder_n_top)
        self._spatial_active = None if spatial_active is None else np.asarray(spatial_active, dtype=bool).reshape(-1)
        self._queue_records = int(queue_records)
        self._ws = None
        self._cs = None
        self._samples_key = None
        self._dev_small = None

    # -- protocol -------------------------------------------------------------
    def kernels(self) -> Tuple[str, ...]:
        return tuple(k.kernel_id for k in self._kernel_bank)

    @property
    def kernel_bank(self) -> Tuple[Kernel, ...]:
        return self._kernel_bank

    @property
    def layer2_state(self) -> Layer2State:
        return self._layer2

    def _validate(self, cube, validity_mask):
        if cube.dim() != 4:
            raise ValueError(f"cube.dim()={cube.dim()}, expected 4 [T_det, N_fdm, H, W]; got shape "
                             f"{tuple(cube.shape)}")
        if cube.shape[2] != cube.shape[3]:
            raise ValueError(f"cube spatial axes must be square (H == W); got H={cube.shape[2]}, W={cube.shape[3]}")
        if validity_mask.dim() != 2:
            raise ValueError(f"validity_mask.dim()={validity_mask.dim()}, expected 2 [T_det, N_fdm]; got shape "
                             f"{tuple(validity_mask.shape)}")
        if validity_mask.shape[0] != cube.shape[0]:
            raise ValueError(f"validity_mask.shape[0]={validity_mask.shape[0]} != cube.shape[0]={cube.shape[0]} (T_det)")
        if validity_mask.shape[1] != cube.shape[1]:
            raise ValueError(f"validity_mask.shape[1]={validity_mask.shape[1]} != cube.shape[1]={cube.shape[1]} (N_fdm)")
        if validity_mask.dtype != torch.bool:
            raise TypeError(f"validity_mask.dtype={validity_mask.dtype}, expected torch.bool")
        if cube.dtype not in (torch.float16, torch.float32):
            raise TypeError(f"cube.dtype={cube.dtype}, expected float16 or float32")

    def _sample_cells(self, n_out: int, F: int, HW: int) -> torch.Tensor:
        """Fixed seeded sample of interior cells ((t, f, pixel) flat indices),
        restricted to active pixels when ``spatial_active`` is given."""
        key = (n_out, F, HW)
        if self._samples_key != key:
            total = n_out * F * HW
            rng = np.random.default_rng(0x5A3C)
            if self._spatial_active is not None and self._spatial_active.size == HW:
                pix = np.flatnonzero(self._spatial_active)
                m = min(self._n_samples, n_out * F * pix.size)
                t = rng.integers(0, n_out, m)
                f = rng.integers(0, F, m)
                p = pix[rng.integers(0, pix.size, m)]
                cells = (t * F + f) * HW + p
            elif total <= self._n_samples:
                cells = np.arange(total)
            else:
                cells = rng.integers(0, total, self._n_samples)
            self._samples = torch.from_numpy(cells.astype(np.int64)).to(self._device)
            self._samples_out = torch.empty((len(self._widths), cells.size), dtype=torch.float32,
                                            device=self._device)
            self._samples_key = key
        return self._samples

    def forward(self, cube: torch.Tensor, validity_mask: torch.Tensor, sigma_layer1: torch.Tensor, *,
                dm_idx_canonical_lo: Optional[int] = None, dm_idx_canonical_hi: Optional[int] = None,
                n_kernel_max_t: Optional[int] = None, event_specnum: int = 0,
                fine_to_coarse: Optional[torch.Tensor] = None,
                fine_dm_pc_cm3: Optional[torch.Tensor] = None) -> List[Candidate]:
        self._validate(cube, validity_mask)
        _lib.require_gpu(cube.contiguous())
        T, F, H, W = (int(x) for x in cube.shape)
        n_cell = F * H * W
        w_max = int(self._widths.max())
        h = w_max // 2
        t_edge = max(0, int(self._n_kernel_max_t if n_kernel_max_t is None else n_kernel_max_t) // 2)
        t_base = t_edge if (t_edge > 0 and 2 * t_edge < T) else 0
        n_out = T - 2 * t_base
        x = cube.contiguous()
        if self._cs is None or tuple(self._cs.shape) != (T + 2 * h + 1, n_cell):
            self._cs = torch.empty((T + 2 * h + 1, n_cell), dtype=torch.float32, device=x.device)
        if self._ws is None:
            self._ws = torch.empty(256 + 16 * self._queue_records, dtype=torch.uint8, device=x.device)
            self._top = torch.empty((4096, 4), dtype=torch.int32, device=x.device)
            self._counts = torch.zeros(2, dtype=torch.int32, device=x.device)
        s_k = self._layer2.s_k                           # earlier cubes' sigma_k
        inv = (1.0 / np.maximum(s_k, 1e-30)).astype(np.float32)
        P = _lib._P
        stream = _lib.stream()
        _lib.call("dsx_cube_cumsum", _lib.ptr(x), _lib.DSX_F16 if x.dtype == torch.float16 else _lib.DSX_F32,
                  _lib.ptr(self._cs), T, n_cell, h, stream)
        _lib.call("dsx_cube_candidates", _lib.ptr(self._cs), T, n_cell, h, self._widths.ctypes.data_as(P),
                  inv.ctypes.data_as(P), len(self._widths), float(self._threshold), t_base, n_out,
                  min(self._n_top, 4096), _lib.ptr(self._top), _lib.ptr(self._counts[:1]),
                  _lib.ptr(self._counts[1:]), _lib.ptr(self._ws), int(self._ws.numel()), stream)
        cells = self._sample_cells(n_out, F, H * W)
        _lib.call("dsx_cube_box_samples", _lib.ptr(self._cs), T, n_cell, h, self._widths.ctypes.data_as(P),
                  len(self._widths), t_base, n_out, _lib.ptr(cells), int(cells.numel()),
                  _lib.ptr(self._samples_out), stream)
        counts = self._counts.cpu().numpy()
        m = int(counts[0])
        self.last_n_above = int(counts[1])
        self.last_overflow = self.last_n_above > self._queue_records
        rec = self._top[:m].cpu().numpy().copy().view(CUBE_REC_DTYPE).reshape(-1)
        samples = self._samples_out.cpu().numpy()
        # Layer-2 update from this cube (used from the next cube on)
        if self._valid_min_fraction >= 1.0:
            cube_valid = bool(torch.all(validity_mask).item())
        else:
            cube_valid = float(validity_mask.float().mean().item()) >= self._valid_min_fraction
        warming = self._layer2.is_warming_up
        sig_this = np.array([sigma_clipped_std_host(samples[k]) for k in range(len(self._widths))])
        self._layer2.update_and_query(per_kernel_sigma=sig_this, valid=cube_valid)
        # merge-radius decode of the strongest cells (already in (snr, flat index) order)
        rt, rf, rl = self._radii
        f2c = None if fine_to_coarse is None else fine_to_coarse.detach().cpu().numpy()
        fdm = None if fine_dm_pc_cm3 is None else fine_dm_pc_cm3.detach().cpu().numpy()
        acc, out = [], []
        flags = int(CandidateFlags.NOISE_WARMUP) if warming else int(CandidateFlags.NONE)
        for r in rec:
            v = float(r["snr"])
            if v <= self._threshold:
                break
            t = int(r["t"])
            f, rem = divmod(int(r["cell"]), H * W)
            l_, m_ = divmod(rem, W)
            if any(abs(t - a) <= rt and abs(f - b) <= rf and abs(l_ - c) <= rl and abs(m_ - d) <= rl
                   for a, b, c, d in acc):
                continue
            acc.append((t, f, l_, m_))
            k = int(r["k"])
            w = int(self._widths[k])
            out.append(Candidate(l=float(l_), m=float(m_),
                                 dm_fine=float(fdm[f]) if fdm is not None else float(f),
                                 dm_idx=int(f2c[f]) if f2c is not None else int(f),
                                 event_specnum=int(event_specnum) + t_base + t, width_samples=w,
                                 kernel_id=self._kernel_bank[k].kernel_id, snr=v,
                                 detector_version=self._detector_version, flags=flags,
                                 search_node_id=self._search_node_id, gpu_half=self._gpu_half,
                                 time_bin=t_base + t))
        if self._merger_config is not None:              # C1 geometry over the decoded survivors
            out = merge_across_kernels_c1(out, self._merger_config)
        if dm_idx_canonical_lo is not None and dm_idx_canonical_hi is not None:
            out = [c for c in out if dm_idx_canonical_lo <= c.dm_idx < dm_idx_canonical_hi]
        if self._c1_snr_min is not None:
            out = [c for c in out if float(c.snr) >= self._c1_snr_min]
        return out

    @property
    def merger_config(self) -> Optional[MergerConfig]:
        return self._merger_config

    __call__ = forward
