"""Unpack and correlator stages with the reference's stage APIs.

* ``unpack_int4_split`` -- ``src/dsart/services/slow_corr_kernel.py:220-403``:
  on-wire ``(NPKT, NANT, NCH, 2t, 2p)`` bytes -> ``(real, imag)`` planes in GEMM
  layout ``(NCH, 2t, 2p, NPKT, NANT)``; ``out_dtype=torch.int8`` keeps the
  exact 4-bit integers (``scale`` ignored).
* ``FastCorrKernel.compute_split`` -- ``src/dsart/services/corr_fast_kernel.py:289-499``:
  per-``t_int_fast_native`` visibilities with optional Stokes-I + channel sum.
  The HIP correlator runs int8 MFMA with int32 accumulation; fed int8 planes
  (or raw bytes through :meth:`FastCorrKernel.compute_raw`) it is exact, and
  ``accum_dtype=torch.int32`` returns the integers as ``(..., 2)`` int32.
  Float planes must hold ``scale * k`` with integer ``k`` in [-8, 7] (what
  ``unpack_int4_split`` produces); they are mapped back to ``k`` and the
  result is scaled by ``scale^2``.
* ``stokes_i_pol_sum`` -- ``corr_fast_kernel.py:764-792``.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional, Tuple

import numpy as np
import torch

from . import _lib
from .constants import LEGACY_FLUFF_SCALE, NPOL, NTIMES_PER_PACKET
from .pattern import upper_tri_indices

__all__ = ["unpack_int4_split", "FastCorrKernel", "stokes_i_pol_sum", "upper_tri_indices"]


def unpack_int4_split(raw, *, device="cuda", scale: float = LEGACY_FLUFF_SCALE,
                      out_dtype: torch.dtype = torch.float16,
                      shape: Optional[Tuple[int, int, int]] = None):
    """Fluff one on-wire page into ``(real, imag)`` in GEMM layout.

    ``shape = (n_pkt, n_ant, n_chan)`` (default: the production page
    2048 x 96 x 384).  ``raw`` may be bytes, a numpy array or a uint8 tensor;
    host data is copied to ``device`` first.  Raises ``ValueError`` on a size
    mismatch and ``TypeError`` on an unsupported input type.
    """
    n_pkt, n_ant, n_chan = shape if shape is not None else (2048, 96, 384)
    expected = n_pkt * n_ant * n_chan * NTIMES_PER_PACKET * NPOL
    if isinstance(raw, (bytes, memoryview)):
        raw = np.frombuffer(raw, dtype=np.uint8)
    if isinstance(raw, np.ndarray):
        if raw.dtype != np.uint8:
            raw = raw.view(np.uint8)
        if raw.size != expected:
            raise ValueError(f"fada raw size {raw.size} != expected {expected} "
                             f"(= prod{(n_pkt, n_ant, n_chan, 2, 2)})")
        raw_t = torch.from_numpy(np.ascontiguousarray(raw).reshape(-1)).to(device)
    elif isinstance(raw, torch.Tensor):
        if raw.numel() != expected:
            raise ValueError(f"fada raw size {raw.numel()} != expected {expected}")
        raw_t = raw.reshape(-1).to(device).contiguous()
        if raw_t.dtype != torch.uint8:
            raw_t = raw_t.view(torch.uint8)
    else:
        raise TypeError(f"unsupported raw type {type(raw)!r}")
    codes = {torch.int8: _lib.DSX_I8, torch.float16: _lib.DSX_F16, torch.float32: _lib.DSX_F32}
    if out_dtype not in codes:
        raise TypeError(f"out_dtype {out_dtype} not supported (int8, float16, float32)")
    out_shape = (n_chan, NTIMES_PER_PACKET, NPOL, n_pkt, n_ant)
    real = torch.empty(out_shape, dtype=out_dtype, device=raw_t.device)
    imag = torch.empty_like(real)
    _lib.require_gpu(raw_t, real, imag)
    _lib.call("dsx_unpack_i4", _lib.ptr(raw_t), _lib.ptr(real), _lib.ptr(imag), n_pkt, n_ant, n_chan,
              codes[out_dtype], float(scale), _lib.stream())
    return real, imag


def stokes_i_pol_sum(vis: torch.Tensor) -> torch.Tensor:
    """``V_xx + V_yy`` over the trailing pol axis (no 1/2)."""
    if vis.shape[-1] != 2:
        raise ValueError(f"expected last dim = 2 (V_xx, V_yy), got shape {tuple(vis.shape)}")
    if not vis.is_complex():
        raise ValueError(f"expected complex visibility, got dtype {vis.dtype}")
    return vis[..., 0] + vis[..., 1]


@dataclass
class FastCorrKernel:
    """Per-device fast correlator (``corr_fast_kernel.py:289-363``)."""

    device: torch.device
    t_int_fast_native: int = 8
    nants: int = 96
    nchan: int = 384
    nvolt_pol: int = NPOL
    nbada_pol: int = NPOL
    accum_dtype: torch.dtype = torch.float32
    _a_idx: torch.Tensor = field(init=False, repr=False)
    _b_idx: torch.Tensor = field(init=False, repr=False)

    def __post_init__(self) -> None:
        dev = torch.device(self.device)
        if dev.type == "cuda" and dev.index is None:
            dev = torch.device(f"cuda:{torch.cuda.current_device()}")
        self.device = dev
        if self.nbada_pol > self.nvolt_pol:
            raise ValueError(f"nbada_pol ({self.nbada_pol}) > nvolt_pol ({self.nvolt_pol})")
        if self.t_int_fast_native <= 0:
            raise ValueError(f"t_int_fast_native must be > 0, got {self.t_int_fast_native}")
        if self.t_int_fast_native % NTIMES_PER_PACKET != 0:
            raise ValueError(f"t_int_fast_native ({self.t_int_fast_native}) must be a multiple of "
                             f"NTIMES_PER_PACKET ({NTIMES_PER_PACKET})")
        if self.nvolt_pol != NPOL or self.nbada_pol != NPOL:
            raise ValueError("the HIP correlator handles two voltage and two output polarisations")
        a, b = upper_tri_indices(self.nants)
        self._a_idx = torch.from_numpy(a)
        self._b_idx = torch.from_numpy(b)

    @property
    def nbase(self) -> int:
        return self.nants * (self.nants + 1) // 2

    def _validate(self, real_v, imag_v):
        if real_v.shape != imag_v.shape:
            raise ValueError(f"real_v shape {tuple(real_v.shape)} != imag_v shape {tuple(imag_v.shape)}")
        if real_v.dtype != imag_v.dtype:
            raise ValueError(f"real_v dtype {real_v.dtype} != imag_v dtype {imag_v.dtype}")
        if real_v.ndim != 5:
            raise ValueError(f"voltage tensor must be 5D (nchan, NTIMES_PER_PACKET, nvolt_pol, "
                             f"n_packets_in, nants), got {real_v.ndim}D")
        nch, nt, npol, n_pk, na = real_v.shape
        if nch != self.nchan or nt != NTIMES_PER_PACKET or npol != self.nvolt_pol or na != self.nants:
            raise ValueError(f"voltage shape {tuple(real_v.shape)} does not match the kernel")
        if n_pk <= 0:
            raise ValueError(f"voltage n_packets_in must be > 0, got {n_pk}")
        ppfv = self.t_int_fast_native // NTIMES_PER_PACKET
        if n_pk % ppfv != 0:
            raise ValueError(f"n_packets_in ({n_pk}) not a multiple of packets_per_fast_vis ({ppfv})")
        for name, t in (("real_v", real_v), ("imag_v", imag_v)):
            if t.device != self.device:
                raise ValueError(f"{name} on {t.device}, kernel on {self.device}")
        return n_pk

    def _run(self, src, im, n_pk, input_layout, fuse_stokes_i, chan_sum_factor):
        if chan_sum_factor < 1 or self.nchan % chan_sum_factor != 0:
            raise ValueError(f"chan_sum_factor={chan_sum_factor} must be >= 1 and divide nchan={self.nchan}")
        if not fuse_stokes_i and chan_sum_factor != 1:
            raise ValueError("chan_sum_factor > 1 requires fuse_stokes_i")
        n_fv = n_pk * NTIMES_PER_PACKET // self.t_int_fast_native
        nco = self.nchan // chan_sum_factor
        npol = 1 if fuse_stokes_i else 2
        vis = torch.empty((n_fv, self.nbase, nco, npol, 2), dtype=torch.int32, device=self.device)
        _lib.require_gpu(src, im, vis)
        _lib.call("dsx_corr_i8", _lib.ptr(src), _lib.ptr(im), _lib.ptr(vis), n_pk, self.nants, self.nchan,
                  self.t_int_fast_native, chan_sum_factor, int(fuse_stokes_i), input_layout, 0,
                  _lib.stream())
        if fuse_stokes_i:
            vis = vis[:, :, :, 0]
        return vis

    def _finish(self, vis_i32, scale2):
        if self.accum_dtype == torch.int32:
            return vis_i32
        v = torch.complex(vis_i32[..., 0].to(torch.float32), vis_i32[..., 1].to(torch.float32))
        if scale2 != 1.0:
            v = v * scale2
        return v.to(torch.complex64)

    def compute_split(self, real_v: torch.Tensor, imag_v: torch.Tensor, *, n_fv_chunk: Optional[int] = None,
                      fuse_stokes_i: bool = False, chan_sum_factor: int = 1) -> torch.Tensor:
        """Visibilities ``(n_fv, NBASE, NCHAN[_eff][, 2])`` (complex64, or int32
        pairs when ``accum_dtype=torch.int32``); ``vis[bls] = sum conj(E_b) E_a``.
        ``n_fv_chunk`` is accepted for compatibility; the HIP kernel streams
        the block in one launch."""
        n_pk = self._validate(real_v, imag_v)
        n_fv = n_pk * NTIMES_PER_PACKET // self.t_int_fast_native
        if n_fv_chunk is not None and not (1 <= n_fv_chunk <= n_fv):
            raise ValueError(f"n_fv_chunk={n_fv_chunk} must be in [1, n_fv_total={n_fv}]")
        if real_v.dtype == torch.int8:
            re8, im8, scale = real_v.contiguous(), imag_v.contiguous(), 1.0
        elif real_v.dtype in (torch.float16, torch.float32):
            nz = torch.cat([real_v.reshape(-1), imag_v.reshape(-1)]).abs()
            scale = float(nz[nz > 0].min()) if bool((nz > 0).any()) else 1.0
            k_re = torch.round(real_v.float() / scale)
            k_im = torch.round(imag_v.float() / scale)
            if bool((k_re.abs() > 8).any()) or bool((k_im.abs() > 8).any()) or \
                    not torch.allclose(k_re * scale, real_v.float(), rtol=1e-3, atol=0) or \
                    not torch.allclose(k_im * scale, imag_v.float(), rtol=1e-3, atol=0):
                raise ValueError("float voltages must be 4-bit integers times one scale "
                                 "(as produced by unpack_int4_split)")
            re8, im8 = k_re.to(torch.int8).contiguous(), k_im.to(torch.int8).contiguous()
        else:
            raise TypeError(f"voltage dtype {real_v.dtype} not supported")
        vis = self._run(re8, im8, n_pk, 1, fuse_stokes_i, chan_sum_factor)
        return self._finish(vis, scale * scale)

    def compute_raw(self, raw: torch.Tensor, *, fuse_stokes_i: bool = True, chan_sum_factor: int = 1,
                    scale: float = 1.0) -> torch.Tensor:
        """Fused unpack + correlate straight from on-wire bytes
        ``(n_pkt, nants, nchan, 2, 2)`` uint8 (no voltage planes in HBM)."""
        if raw.dtype != torch.uint8 or raw.dim() != 5 or tuple(raw.shape[1:]) != (self.nants, self.nchan, 2, 2):
            raise ValueError(f"raw must be uint8 (n_pkt, {self.nants}, {self.nchan}, 2, 2)")
        n_pk = int(raw.shape[0])
        if n_pk % (self.t_int_fast_native // NTIMES_PER_PACKET):
            raise ValueError("n_pkt must be a multiple of t_int_fast_native / 2")
        vis = self._run(raw.contiguous(), None, n_pk, 0, fuse_stokes_i, chan_sum_factor)
        return self._finish(vis, scale * scale)
