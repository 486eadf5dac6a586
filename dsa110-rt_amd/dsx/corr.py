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
  Float planes (``unpack_int4_split``'s fp16 output, calibrated or not) run
  the fp16-MFMA correlator with float32 accumulation (``dsx_corr_cal``):
  the reference's fp16 tensor-core contract; float32 planes are rounded to
  fp16 operands first.
* ``apply_cal_split`` / ``make_cal_broadcast_tensors`` / ``compact_cal`` --
  ``slow_corr_kernel.py:459-569``, ``x2/corr/gemm2x.py:158-168``: complex
  gains (RFI mask folded in as zeros) on fp16 planes; the hot path applies
  them on the correlator's operand load instead
  (:meth:`FastCorrKernel.compute_raw` with ``cal=``).
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

__all__ = ["unpack_int4_split", "FastCorrKernel", "stokes_i_pol_sum", "upper_tri_indices",
           "apply_cal_split", "make_cal_broadcast_tensors", "compact_cal"]


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
            vis = self._run(real_v.contiguous(), imag_v.contiguous(), n_pk, 1, fuse_stokes_i,
                            chan_sum_factor)
            return self._finish(vis, 1.0)
        if real_v.dtype not in (torch.float16, torch.float32):
            raise TypeError(f"voltage dtype {real_v.dtype} not supported")
        re16 = real_v.to(torch.float16).contiguous()
        im16 = imag_v.to(torch.float16).contiguous()
        return self._run_cal(re16, im16, n_pk, 1, fuse_stokes_i, chan_sum_factor, None, None, 1.0)

    def _run_cal(self, src, im, n_pk, input_layout, fuse_stokes_i, chan_sum_factor, cal, time_keep,
                 scale):
        if chan_sum_factor < 1 or self.nchan % chan_sum_factor != 0:
            raise ValueError(f"chan_sum_factor={chan_sum_factor} must be >= 1 and divide nchan={self.nchan}")
        if not fuse_stokes_i and chan_sum_factor != 1:
            raise ValueError("chan_sum_factor > 1 requires fuse_stokes_i")
        cal_re = cal_im = tk = None
        n_acc = 1
        if cal is not None:
            cal_re, cal_im = _compact_gains(cal, self.nchan, self.nants, self.device)
        if time_keep is not None:
            if cal is None:
                raise ValueError("time_keep needs cal (the keep factors multiply the gains)")
            tk = torch.as_tensor(time_keep, device=self.device).to(torch.float32).contiguous()
            if tk.dim() != 3 or tk.shape[1:] != (self.nchan, NPOL) or n_pk % tk.shape[0] != 0:
                raise ValueError(f"time_keep must be (n_acc, {self.nchan}, {NPOL}) with n_acc "
                                 f"dividing n_packets ({n_pk}), got {tuple(tk.shape)}")
            n_acc = int(tk.shape[0])
        n_fv = n_pk * NTIMES_PER_PACKET // self.t_int_fast_native
        nco = self.nchan // chan_sum_factor
        npol = 1 if fuse_stokes_i else 2
        vis = torch.empty((n_fv, self.nbase, nco, npol, 2), dtype=torch.float32, device=self.device)
        _lib.require_gpu(src, im, vis, cal_re, cal_im, tk)
        _lib.call("dsx_corr_cal", _lib.ptr(src), _lib.ptr(im), _lib.ptr(cal_re), _lib.ptr(cal_im),
                  _lib.ptr(tk), n_acc, _lib.ptr(vis), n_pk, self.nants, self.nchan,
                  self.t_int_fast_native, chan_sum_factor, int(fuse_stokes_i), input_layout, 0,
                  float(scale), _lib.stream())
        v = torch.view_as_complex(vis)
        return v[:, :, :, 0] if fuse_stokes_i else v

    def compute_raw(self, raw: torch.Tensor, *, fuse_stokes_i: bool = True, chan_sum_factor: int = 1,
                    scale: float = 1.0, cal=None, time_keep=None) -> torch.Tensor:
        """Fused unpack + correlate straight from on-wire bytes
        ``(n_pkt, nants, nchan, 2, 2)`` uint8 (no voltage planes in HBM).

        Without ``cal``: exact int8 MFMA (result times ``scale**2``).  With
        ``cal = (cal_real, cal_imag)`` (``make_cal_broadcast_tensors`` /
        ``compact_cal`` layouts; RFI-masked gains zero) the gains are applied
        as each voltage ``fp16(code * scale)`` is loaded, rounded once to fp16,
        and correlated by fp16 MFMA with float32 accumulation
        (``x2/corr/gemm2x.py:158-168``); ``time_keep`` ``(n_acc, nchan, 2)``
        multiplies the gains of each accumulation's packets."""
        if raw.dtype != torch.uint8 or raw.dim() != 5 or tuple(raw.shape[1:]) != (self.nants, self.nchan, 2, 2):
            raise ValueError(f"raw must be uint8 (n_pkt, {self.nants}, {self.nchan}, 2, 2)")
        n_pk = int(raw.shape[0])
        if n_pk % (self.t_int_fast_native // NTIMES_PER_PACKET):
            raise ValueError("n_pkt must be a multiple of t_int_fast_native / 2")
        if cal is not None or time_keep is not None:
            return self._run_cal(raw.contiguous(), None, n_pk, 0, fuse_stokes_i, chan_sum_factor, cal,
                                 time_keep, scale)
        vis = self._run(raw.contiguous(), None, n_pk, 0, fuse_stokes_i, chan_sum_factor)
        return self._finish(vis, scale * scale)


def _compact_gains(cal, nchan: int, nants: int, device) -> Tuple[torch.Tensor, torch.Tensor]:
    """``(cal_real, cal_imag)`` in the broadcast layout ``(NCHAN or 1, 1, NPOL, 1,
    NANTS)`` (``make_cal_broadcast_tensors``) or the compact ``(NCHAN, NPOL,
    NANTS)`` (``compact_cal``) -> float32 ``(nchan, 2, nants)`` on ``device``."""
    if not (isinstance(cal, (tuple, list)) and len(cal) == 2):
        raise TypeError("cal must be a (cal_real, cal_imag) pair")
    out = []
    for c in cal:
        c = torch.as_tensor(c, device=device)
        if c.is_complex():
            raise TypeError("cal_real / cal_imag must be real tensors")
        if c.dim() == 5:
            if c.shape[1] != 1 or c.shape[3] != 1:
                raise ValueError(f"broadcast gains must be (NCHAN or 1, 1, NPOL, 1, NANTS), got {tuple(c.shape)}")
            c = c[:, 0, :, 0, :]
        if c.dim() != 3 or c.shape[1:] != (NPOL, nants) or c.shape[0] not in (1, nchan):
            raise ValueError(f"gains must be ({nchan} or 1, {NPOL}, {nants}), got {tuple(c.shape)}")
        out.append(c.to(torch.float32).expand(nchan, NPOL, nants).contiguous())
    return out[0], out[1]


def make_cal_broadcast_tensors(gains_fine, *, device="cpu", dtype: torch.dtype = torch.float16):
    """``(NANTS, NCHAN, NPOL)`` complex gains -> ``(cal_real, cal_imag)`` each
    ``(NCHAN, 1, NPOL, 1, NANTS)`` (``slow_corr_kernel.py:572-611``)."""
    g = np.asarray(gains_fine)
    if g.ndim != 3 or g.shape[2] != NPOL:
        raise ValueError(f"gains_fine shape {g.shape}, expected (NANTS, NCHAN, {NPOL})")
    real = torch.from_numpy(np.ascontiguousarray(g.real)).to(device=device, dtype=dtype)
    imag = torch.from_numpy(np.ascontiguousarray(g.imag)).to(device=device, dtype=dtype)
    real = real.permute(1, 2, 0).contiguous()[:, None, :, None, :]
    imag = imag.permute(1, 2, 0).contiguous()[:, None, :, None, :]
    return real, imag


def compact_cal(cal_real: torch.Tensor, cal_imag: torch.Tensor, cube_mask: Optional[torch.Tensor] = None):
    """Broadcast gains ``(NCHAN, 1, NPOL, 1, NANTS)`` -> float32 ``(NCHAN, NPOL,
    NANTS)`` with the RFI mask ``cube_mask`` ``(ANT, CH, POL)`` (True = flagged)
    folded in as zero gains (``x2/corr/gemm2x.py:158-168``)."""
    cr = cal_real[:, 0, :, 0, :].float()
    ci = cal_imag[:, 0, :, 0, :].float()
    if cube_mask is not None:
        keep = (~torch.as_tensor(cube_mask, device=cr.device, dtype=torch.bool)).permute(1, 2, 0).to(cr.dtype)
        cr = cr * keep
        ci = ci * keep
    return cr.contiguous(), ci.contiguous()


def apply_cal_split(real_v: torch.Tensor, imag_v: torch.Tensor, cal_real_fine: torch.Tensor,
                    cal_imag_fine: torch.Tensor):
    """Per-(ant, fine_ch, pol) complex gain on split fp16 voltages in GEMM layout
    ``(NCHAN, 2, NPOL, NPKT, NANTS)`` (``slow_corr_kernel.py:459-569``):
    ``real_out = fp16(v_re g_re - v_im g_im)``, ``imag_out = fp16(v_re g_im + v_im g_re)``,
    each product and sum rounded in float32, one fp16 rounding.  Gains are
    ``(NCHAN or 1, 1, NPOL, 1, NANTS)`` of the voltages' dtype.  Returns new
    tensors; raises ``ValueError`` on a shape / dtype / device mismatch."""
    if real_v.shape != imag_v.shape:
        raise ValueError(f"real_v shape {tuple(real_v.shape)} != imag_v shape {tuple(imag_v.shape)}")
    if real_v.dim() != 5 or real_v.shape[1] != NTIMES_PER_PACKET or real_v.shape[2] != NPOL:
        raise ValueError(f"voltage shape {tuple(real_v.shape)} is not (NCHAN, {NTIMES_PER_PACKET}, "
                         f"{NPOL}, NPKT, NANTS)")
    if real_v.dtype != imag_v.dtype:
        raise ValueError(f"real_v dtype {real_v.dtype} != imag_v dtype {imag_v.dtype}")
    if cal_real_fine.dtype != real_v.dtype or cal_imag_fine.dtype != real_v.dtype:
        raise ValueError(f"cal dtypes ({cal_real_fine.dtype}, {cal_imag_fine.dtype}) != "
                         f"voltage dtype {real_v.dtype}")
    if cal_real_fine.device != real_v.device or cal_imag_fine.device != real_v.device:
        raise ValueError(f"cal devices ({cal_real_fine.device}, {cal_imag_fine.device}) != "
                         f"voltage device {real_v.device}")
    if real_v.dtype != torch.float16:
        raise TypeError(f"apply_cal_split runs on fp16 voltages, got {real_v.dtype}")
    nch, _, _, npk, na = real_v.shape
    cr, ci = _compact_gains((cal_real_fine, cal_imag_fine), nch, na, real_v.device)
    re = real_v.contiguous()
    im = imag_v.contiguous()
    out_re = torch.empty_like(re)
    out_im = torch.empty_like(im)
    _lib.require_gpu(re, im, cr, ci, out_re, out_im)
    _lib.call("dsx_apply_cal", _lib.ptr(re), _lib.ptr(im), _lib.ptr(cr), _lib.ptr(ci), 1,
              _lib.ptr(out_re), _lib.ptr(out_im), nch, npk, na, _lib.stream())
    return out_re, out_im
