"""Imager stage: edge mask helpers and the fused gridder + C2R FFT imager.

``image_mask_npad`` / ``compute_edge_mask`` / ``apply_edge_mask`` keep the
signatures and results of ``src/dsart/image/imager.py:57-175,304-324``.
``SubbandImager`` wraps ``dsx_image_subbands``, which produces, per
(subband, step), ``compute_edge_mask(...) * fftshift(Re(ifft2(grid)))`` of the
reference's dense grid without ever materialising that grid in HBM.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np
import torch

from . import _lib
from .pattern import CellMap, build_imager_taps


def image_mask_npad(*, n_grid: int, kernel_support: int = 5, sigma_l_pix: Optional[float] = None,
                    envelope_threshold: float = 0.5) -> int:
    """``max(K//2 + 2, ceil(N/2 - sigma_l sqrt(2 ln(1/thr))))`` (imager.py:57-109)."""
    if n_grid <= 0 or n_grid & (n_grid - 1):
        raise ValueError(f"n_grid={n_grid}, expected positive power of two")
    if kernel_support < 1:
        raise ValueError(f"kernel_support={kernel_support}, expected >= 1")
    if not (0.0 < envelope_threshold <= 1.0):
        raise ValueError(f"envelope_threshold={envelope_threshold}, expected in (0, 1]")
    pad_kernel = (kernel_support // 2) + 2
    if sigma_l_pix is not None and sigma_l_pix > 0:
        rad = float(sigma_l_pix) * math.sqrt(2.0 * math.log(1.0 / envelope_threshold))
        pad_env = max(0, int(math.ceil((n_grid / 2.0) - rad)))
    else:
        pad_env = 0
    return max(pad_kernel, pad_env)


def mask_npad(n_grid: int, kernel_support: int = 1) -> int:
    return image_mask_npad(n_grid=n_grid, kernel_support=kernel_support)


def compute_edge_mask(*, n_grid: int, kernel_support: int = 5, sigma_l_pix: Optional[float] = None,
                      envelope_threshold: float = 0.5, drop_dc: bool = True, dtype=np.float32,
                      apply_ifft2_dc_correction: bool = True) -> np.ndarray:
    """Multiplicative edge mask with the ``(-1)^(i+j)`` DC-at-centre correction
    (imager.py:112-175)."""
    npad = image_mask_npad(n_grid=n_grid, kernel_support=kernel_support, sigma_l_pix=sigma_l_pix,
                           envelope_threshold=envelope_threshold)
    mask = np.zeros((n_grid, n_grid), dtype=dtype)
    if 2 * npad < n_grid:
        mask[npad:n_grid - npad, npad:n_grid - npad] = 1.0
    if drop_dc:
        dc = n_grid // 2
        mask[dc, dc] = 0.0
    if apply_ifft2_dc_correction:
        ii, jj = np.indices((n_grid, n_grid))
        mask = mask * (1.0 - 2.0 * ((ii + jj) & 1)).astype(dtype)
    return mask


def apply_edge_mask(image, mask):
    """``image * mask`` broadcast over leading axes (imager.py:304-324)."""
    if isinstance(image, torch.Tensor):
        m = mask if isinstance(mask, torch.Tensor) else torch.from_numpy(np.asarray(mask))
        return image * m.to(dtype=image.dtype, device=image.device)
    arr = np.asarray(image)
    return arr * np.asarray(mask).astype(arr.dtype, copy=False)


class SubbandImager:
    """Per-(subband, step) masked dirty images from exact int32 visibilities."""

    def __init__(self, cmap: CellMap, n_sub: int, *, device: str = "cuda", kernel_support: int = 1,
                 out_dtype: torch.dtype = torch.float16, scale: float = 1.0):
        taps = build_imager_taps(cmap, n_sub)
        self.n_grid = cmap.n_grid
        self.n_sub = n_sub
        self.per_sub = taps.per_sub
        self.n_rows = taps.n_rows
        self.npad = mask_npad(cmap.n_grid, kernel_support)
        self.scale = float(scale)
        self.out_dtype = out_dtype
        self.nbase = cmap.row.shape[0]
        dev = torch.device(device)
        self.src = torch.from_numpy(taps.src).to(dev)
        self.dst = torch.from_numpy(taps.dst).to(dev)
        self.off = torch.from_numpy(taps.offsets).to(dev)

    def __call__(self, vis: torch.Tensor, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        """``vis`` (n_steps, n_sub * per_sub, nbase, 2) int32 -> (n_sub, n_steps, N, N)."""
        if vis.dtype != torch.int32 or vis.dim() != 4 or vis.shape[-1] != 2:
            raise ValueError("vis must be int32 (n_steps, n_chan_out, nbase, 2)")
        n_steps = int(vis.shape[0])
        if vis.shape[1] != self.n_sub * self.per_sub or vis.shape[2] != self.nbase:
            raise ValueError(f"vis shape {tuple(vis.shape)} does not match the gridder")
        N = self.n_grid
        if out is None:
            out = torch.empty((self.n_sub, n_steps, N, N), dtype=self.out_dtype, device=vis.device)
        _lib.require_gpu(vis, out)
        _lib.call("dsx_image_subbands", _lib.ptr(vis), _lib.ptr(self.src), _lib.ptr(self.dst),
                  _lib.ptr(self.off), _lib.ptr(out), n_steps, self.n_sub, self.per_sub, self.nbase,
                  N, self.n_rows, self.npad, self.scale,
                  _lib.DSX_F16 if out.dtype == torch.float16 else _lib.DSX_F32, _lib.stream())
        return out


def dirty_image_from_uv_grid(uv_grid, *, out_dtype=None, prescale: float = 1.0):
    """``fftshift(Re(ifft2(prescale * uv_grid)))`` on the last two axes
    (imager.py:248-301), on the HIP imager (``dsx_dirty_image``); leading axes
    pass through.  A numpy grid is copied to the GPU and the image comes back
    as numpy (float32 unless ``out_dtype``), as the reference's numpy path
    returns numpy; a CUDA tensor stays on the device (float32 unless
    ``out_dtype``)."""
    if not isinstance(uv_grid, torch.Tensor):
        arr = np.asarray(uv_grid)
        if arr.ndim < 2 or arr.shape[-1] != arr.shape[-2]:
            raise ValueError(f"uv_grid must be [..., N, N]; got {arr.shape}")
        if not torch.cuda.is_available():
            raise RuntimeError("dirty_image_from_uv_grid runs on the GPU; no ROCm device is visible")
        out = dirty_image_from_uv_grid(torch.from_numpy(np.ascontiguousarray(arr, dtype=np.complex64)).cuda(),
                                       prescale=prescale).cpu().numpy()
        return out if out_dtype is None else out.astype(out_dtype, copy=False)
    if uv_grid.dim() < 2 or uv_grid.shape[-1] != uv_grid.shape[-2]:
        raise ValueError(f"uv_grid must be [..., N, N]; got {tuple(uv_grid.shape)}")
    N = int(uv_grid.shape[-1])
    lead = tuple(uv_grid.shape[:-2])
    g = uv_grid.to(torch.complex64).reshape(-1, N, N).contiguous()
    out = torch.empty(g.shape, dtype=torch.float32, device=g.device)
    _lib.require_gpu(g, out)
    _lib.call("dsx_dirty_image", _lib.ptr(torch.view_as_real(g)), _lib.ptr(out), int(g.shape[0]), N,
              float(prescale), _lib.stream())
    out = out.reshape(lead + (N, N))
    return out if out_dtype is None else out.to(out_dtype)


# ---------------------------------------------------------------------------
# search-cube imager (GpuImager.build(cfg).process_cube drop-in)
# ---------------------------------------------------------------------------


@dataclass
class GpuImagerConfig:
    """Cube imager configuration (fields and defaults of
    ``src/dsart/image/imager_gpu.py:76-117``).  ``complex_dtype`` is kept for
    signature compatibility: no uv slab is materialised here."""
    n_grid: int = 256
    t_det: int = 256
    n_fdm: int = 32
    n_chgroup: int = 16
    kernel_support: int = 5
    sigma_l_pix: Optional[float] = None
    envelope_threshold: float = 0.5
    drop_dc: bool = True
    cube_dtype: torch.dtype = torch.float16
    complex_dtype: torch.dtype = torch.complex32
    device: Optional[torch.device] = None
    imager_uv_prescale: float = 1.0

    def resolved_device(self) -> torch.device:
        if self.device is not None:
            return torch.device(self.device)
        if not torch.cuda.is_available():
            raise RuntimeError("GpuImager runs on a ROCm GPU; no device is visible")
        return torch.device("cuda")


@dataclass
class GpuImager:
    """Search-cube imager: per fine-DM trial, the shift-and-add of the
    chgroups' cint8 uv streams, the C2R FFT and the edge mask in one HIP
    kernel (``dsx_image_cube``), written into alternating output cubes
    ``[T_det, N_fdm, N, N]`` (``imager_gpu.py:123-229``)."""
    config: GpuImagerConfig
    device: torch.device
    edge_mask_real: torch.Tensor      # [N, N] cube_dtype
    output_cube: torch.Tensor         # ping-pong buffer A [T_det, N_fdm, N, N]
    output_cube_alt: torch.Tensor     # ping-pong buffer B
    edge_mask_per_fdm: Optional[torch.Tensor] = None
    _mask_f32: Optional[torch.Tensor] = None
    _mask_per_fdm_f32: Optional[torch.Tensor] = None
    _acc: Optional[torch.Tensor] = None      # float32 partial images for a float16 cube
    _output_index: int = 0
    _timings: list = field(default_factory=list)

    @classmethod
    def build(cls, config: GpuImagerConfig) -> "GpuImager":
        """Allocate the output cubes and the edge mask (no compute)."""
        if config.cube_dtype not in (torch.float16, torch.float32):
            raise ValueError(f"cube_dtype={config.cube_dtype}, expected float16 or float32")
        device = config.resolved_device()
        edge = compute_edge_mask(n_grid=config.n_grid, kernel_support=config.kernel_support,
                                 sigma_l_pix=config.sigma_l_pix, envelope_threshold=config.envelope_threshold,
                                 drop_dc=config.drop_dc)
        shape = (config.t_det, config.n_fdm, config.n_grid, config.n_grid)
        out = cls(config=config, device=device,
                  edge_mask_real=torch.from_numpy(edge).to(device=device, dtype=config.cube_dtype),
                  output_cube=torch.empty(shape, dtype=config.cube_dtype, device=device),
                  output_cube_alt=torch.empty(shape, dtype=config.cube_dtype, device=device))
        out._mask_f32 = torch.from_numpy(edge.astype(np.float32)).to(device)
        if config.cube_dtype != torch.float32:
            out._acc = torch.empty(shape, dtype=torch.float32, device=device)
        return out

    def set_edge_mask_per_fdm(self, mask: Optional[torch.Tensor]) -> None:
        """Use a per-trial mask ``[N_fdm, N, N]`` instead of ``edge_mask_real``
        (e.g. with a per-trial scale folded in); ``None`` reverts."""
        cfg = self.config
        if mask is None:
            self.edge_mask_per_fdm = self._mask_per_fdm_f32 = None
            return
        if tuple(mask.shape) != (cfg.n_fdm, cfg.n_grid, cfg.n_grid):
            raise ValueError(f"edge_mask_per_fdm.shape={tuple(mask.shape)}, expected "
                             f"({cfg.n_fdm}, {cfg.n_grid}, {cfg.n_grid})")
        self.edge_mask_per_fdm = mask
        self._mask_per_fdm_f32 = mask.to(device=self.device, dtype=torch.float32).contiguous()

    def _cal(self, name, cal, t_stream):
        cfg = self.config
        if cal is None:
            return None
        t = cal if isinstance(cal, torch.Tensor) else torch.as_tensor(np.asarray(cal, dtype=np.float32))
        valid = ((cfg.n_chgroup,), (cfg.n_chgroup, t_stream))
        if tuple(t.shape) not in valid:
            raise ValueError(f"{name}.shape={tuple(t.shape)}, expected {valid[0]} or {valid[1]}")
        return t.to(device=self.device, dtype=torch.float32).contiguous()

    def process_cube(self, *, streams_cint8: torch.Tensor, time_shifts_gpu: torch.Tensor,
                     chgroup_scales=None, chgroup_offsets_re=None, chgroup_offsets_im=None, t_lo: int = 0,
                     compact_inv_lut=None, compact_t_stream: int = 0, compact_shift_bounds=None) -> torch.Tensor:
        """One cube: for trial ``f`` and cube time ``t`` in ``[t_lo, T_det)``,
        ``cube[t, f] = mask * fftshift(Re(ifft2(prescale * sum_g z_g[t - shifts[f, g]])))``
        with ``z_g`` the dequantised (optionally calibrated,
        ``scale * z + offset``) rows of ``streams_cint8`` ``[N_chg, T_stream, 2,
        N, N]`` int8 (zero outside ``[0, T_stream)``).  Writes the next output
        cube in place (rows ``[0, t_lo)`` untouched) and returns it.
        Raises ``ValueError`` on shape / dtype mismatch, like
        ``imager_gpu.py:295-405``."""
        cfg = self.config
        if compact_inv_lut is not None:
            raise ValueError("the compact (sparse-cell) stream layout is not supported by this build; "
                             "pass the dense [N_chg, T_stream, 2, N_grid, N_grid] cint8 streams")
        if streams_cint8.dtype != torch.int8:
            raise ValueError(f"streams_cint8.dtype={streams_cint8.dtype}, expected int8")
        if streams_cint8.ndim != 5:
            raise ValueError(f"streams_cint8 ndim={streams_cint8.ndim}, expected 5 "
                             "([N_chg, T_stream, 2, N_grid, N_grid])")
        n_chg, t_stream, two, n_g, n_g2 = streams_cint8.shape
        if two != 2:
            raise ValueError(f"streams_cint8 inner-2-axis={two}, expected 2 (real / imaginary split planes)")
        if n_g != cfg.n_grid or n_g2 != cfg.n_grid:
            raise ValueError(f"streams_cint8 grid {n_g}x{n_g2} != config N_grid={cfg.n_grid}")
        if n_chg != cfg.n_chgroup:
            raise ValueError(f"streams_cint8 N_chgroup={n_chg} != config N_chgroup={cfg.n_chgroup}")
        if time_shifts_gpu.dtype != torch.int32:
            raise ValueError(f"time_shifts_gpu.dtype={time_shifts_gpu.dtype}, expected int32")
        if tuple(time_shifts_gpu.shape) != (cfg.n_fdm, cfg.n_chgroup):
            raise ValueError(f"time_shifts_gpu.shape={tuple(time_shifts_gpu.shape)}, expected "
                             f"({cfg.n_fdm}, {cfg.n_chgroup})")
        if t_stream < cfg.t_det:
            raise ValueError(f"streams_cint8 T_stream={t_stream} < T_det={cfg.t_det}; no fdm trial can fit")
        if t_lo < 0 or t_lo >= cfg.t_det:
            raise ValueError(f"t_lo={t_lo}, must be in [0, T_det={cfg.t_det})")
        cals = [self._cal(n, c, t_stream) for n, c in (("chgroup_scales", chgroup_scales),
                                                       ("chgroup_offsets_re", chgroup_offsets_re),
                                                       ("chgroup_offsets_im", chgroup_offsets_im))]
        cal_2d = any(c is not None and c.dim() == 2 for c in cals)
        if cal_2d:   # one layout for all three: broadcast per-chgroup values over the stream rows
            cals = [None if c is None else (c if c.dim() == 2 else c[:, None].expand(n_chg, t_stream).contiguous())
                    for c in cals]
        streams = streams_cint8.contiguous()
        shifts = time_shifts_gpu.contiguous()
        _lib.require_gpu(streams, shifts)
        out = self.output_cube if self._output_index == 0 else self.output_cube_alt
        mask = self._mask_per_fdm_f32 if self._mask_per_fdm_f32 is not None else self._mask_f32
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        _lib.call("dsx_image_cube", _lib.ptr(streams), _lib.ptr(shifts), n_chg, t_stream, cfg.n_fdm, cfg.t_det,
                  int(t_lo), cfg.n_grid, float(cfg.imager_uv_prescale), *(_lib.ptr(c) for c in cals),
                  int(cal_2d), _lib.ptr(mask), int(self._mask_per_fdm_f32 is not None), _lib.ptr(out),
                  _lib.DSX_F16 if cfg.cube_dtype == torch.float16 else _lib.DSX_F32, _lib.ptr(self._acc),
                  _lib.stream())
        e1.record()
        self._timings.append((e0, e1))
        if len(self._timings) > 64:
            self._timings.pop(0)
        self._output_index = 1 - self._output_index
        return out

    def pop_substage_timings(self) -> dict:
        """Mean GPU ms per cube since the last call.  Combine, FFT and mask run
        in one kernel here, so the whole time is reported as ``fft_ms``
        (``imager_gpu.py:659-720`` keys)."""
        ring, self._timings = self._timings, []
        if not ring:
            return {"combine_ms": 0.0, "fft_ms": 0.0, "mask_ms": 0.0, "total_ms": 0.0, "n": 0}
        ring[-1][1].synchronize()
        ms = float(np.mean([a.elapsed_time(b) for a, b in ring]))
        return {"combine_ms": 0.0, "fft_ms": ms, "mask_ms": 0.0, "total_ms": ms, "n": len(ring)}


def build_default_gpu_imager(*, n_grid: int = 256, t_det: int = 256, n_fdm: int = 32, n_chgroup: int = 16,
                             cube_dtype: torch.dtype = torch.float16, complex_dtype: torch.dtype = torch.complex32,
                             device: Optional[torch.device] = None) -> GpuImager:
    """A :class:`GpuImager` at the v1 geometry (imager_gpu.py:792-830)."""
    return GpuImager.build(GpuImagerConfig(n_grid=n_grid, t_det=t_det, n_fdm=n_fdm, n_chgroup=n_chgroup,
                                           cube_dtype=cube_dtype, complex_dtype=complex_dtype, device=device))
