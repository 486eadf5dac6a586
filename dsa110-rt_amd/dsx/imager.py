"""Imager stage: edge mask helpers and the fused gridder + C2R FFT imager.

``image_mask_npad`` / ``compute_edge_mask`` / ``apply_edge_mask`` keep the
signatures and results of ``src/dsart/image/imager.py:57-175,304-324``.
``SubbandImager`` wraps ``dsx_image_subbands``, which produces, per
(subband, step), ``compute_edge_mask(...) * fftshift(Re(ifft2(grid)))`` of the
reference's dense grid without ever materialising that grid in HBM.
"""
from __future__ import annotations

import math
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
    (imager.py:248-301).  CUDA tensors run the HIP imager (``dsx_dirty_image``);
    leading axes pass through.  ``out_dtype`` defaults to float32."""
    if not isinstance(uv_grid, torch.Tensor):
        raise TypeError("dirty_image_from_uv_grid runs on the GPU; pass a CUDA tensor")
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
