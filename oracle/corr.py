"""int64 correlator oracle (test infrastructure only).

Pinned by: the exhaustive 256-byte nibble table and the pack/unpack round
trip (``tests/test_slow_corr_synth.py:172-211`` in the reference), the sign
known answer ``E_1 = (1, 0)``, ``E_2 = (0, 1)`` -> ``V = +iK``
(``tests/test_slow_corr_synth.py:219-277``), zero-in/zero-out, real
non-negative autos, and an independent complex-loop brute force on tiny
inputs -- see ``tests/test_oracle_corr.py``.

The calibrated operands (:func:`cal_operands`, :func:`correlate_float`) are
pinned by a known answer with gains exactly representable in fp16
(``V_cal = conj(g_b) g_a V``, the gain application of
``slow_corr_kernel.py:459-502``) and by masked gains / zero keep factors
equalling zeroed voltages (``x2/corr/gemm2x.py:158-168,373-378``).
"""
from __future__ import annotations

from typing import Tuple

import numpy as np


def upper_tri(nants: int) -> Tuple[np.ndarray, np.ndarray]:
    """``(a, b)`` pairs with ``b <= a`` in ``bls = a(a+1)/2 + b`` order
    (slow_corr_kernel.py:161-178)."""
    a_l, b_l = [], []
    for a in range(nants):
        for b in range(a + 1):
            a_l.append(a)
            b_l.append(b)
    return np.asarray(a_l, dtype=np.int64), np.asarray(b_l, dtype=np.int64)


def unpack_bytes(raw: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """Low nibble = real, high nibble = imag, 4-bit two's complement
    (slow_corr_kernel.py:337-403: ``((int8)b << 4) >> 4`` and ``(int8)b >> 4``)."""
    b = np.asarray(raw, dtype=np.uint8).astype(np.int64)
    lo = b & 0x0F
    hi = (b >> 4) & 0x0F
    lo = np.where(lo >= 8, lo - 16, lo)
    hi = np.where(hi >= 8, hi - 16, hi)
    return lo, hi


def voltages(raw: np.ndarray, n_pkt: int, n_ant: int, n_chan: int) -> Tuple[np.ndarray, np.ndarray]:
    """On-wire ``(pkt, ant, chan, 2t, 2p)`` bytes -> int64 ``(sample, ant, chan, pol)``."""
    re, im = unpack_bytes(np.asarray(raw).reshape(n_pkt, n_ant, n_chan, 2, 2))
    re = re.transpose(0, 3, 1, 2, 4).reshape(n_pkt * 2, n_ant, n_chan, 2)
    im = im.transpose(0, 3, 1, 2, 4).reshape(n_pkt * 2, n_ant, n_chan, 2)
    return re, im


def correlate(raw: np.ndarray, *, n_pkt: int, n_ant: int, n_chan: int, t_int: int,
              csf: int = 1, fuse_pol: bool = True) -> np.ndarray:
    """Exact integer visibilities ``V[bls] = sum conj(E_b) E_a`` (b <= a).

    Sums over ``t_int`` consecutive samples, ``csf`` consecutive channels and,
    when ``fuse_pol``, both polarisations (Stokes I = XX + YY, no 1/2,
    corr_fast_kernel.py:764-792).  Returns int64
    ``(n_steps, n_chan // csf, nbase, 2)`` (re, im) when ``fuse_pol`` else
    ``(n_steps, n_chan // csf, 2, nbase, 2)``.
    """
    re, im = voltages(raw, n_pkt, n_ant, n_chan)
    S = n_pkt * 2
    n_steps = S // t_int
    nco = n_chan // csf
    # (step, t, cout, x, pol, ant)
    R = re.reshape(n_steps, t_int, n_ant, nco, csf, 2).transpose(0, 1, 3, 4, 5, 2)
    I = im.reshape(n_steps, t_int, n_ant, nco, csf, 2).transpose(0, 1, 3, 4, 5, 2)
    if fuse_pol:
        sub = "stcxpa,stcxpb->scab"
    else:
        sub = "stcxpa,stcxpb->scpab"
    # V[a, b] = sum conj(E_b) E_a : re = Rb Ra + Ib Ia ; im = Rb Ia - Ib Ra
    vre = np.einsum(sub, R, R) + np.einsum(sub, I, I)
    vim = np.einsum(sub, I, R) - np.einsum(sub, R, I)
    a_idx, b_idx = upper_tri(n_ant)
    out_re = vre[..., a_idx, b_idx]
    out_im = vim[..., a_idx, b_idx]
    return np.ascontiguousarray(np.stack([out_re, out_im], axis=-1), dtype=np.int64)


def correlate_blas(raw: np.ndarray, *, n_pkt: int, n_ant: int, n_chan: int, t_int: int,
                   csf: int = 1) -> np.ndarray:
    """Same visibilities as :func:`correlate` (Stokes I) through float32
    matrix products: samples are integers in [-8, 7], so every product and
    every partial sum (at most 2 * 64 * K for K = 2 * t_int * csf terms) is an
    integer below 2**24 while K <= 131072, and the float32 products are exact.
    Equality with the einsum form is checked in ``tests/test_oracle_corr.py``;
    this twin exists only to make 96-antenna oracle streams fast enough for
    the GPU parity tests."""
    K = t_int * csf * 2
    if 2 * 64 * K >= (1 << 24):
        raise ValueError("float32 products would not be exact for this K")
    b = np.arange(256)
    lut_re = (((b & 0xF) ^ 8) - 8).astype(np.float32)
    lut_im = ((((b >> 4) & 0xF) ^ 8) - 8).astype(np.float32)
    raw = np.asarray(raw, dtype=np.uint8)
    n_steps = n_pkt * 2 // t_int
    nco = n_chan // csf
    # on-wire (pkt, ant, chan, t2, pol) -> (step, cout, [pkt in step, t2, x, pol], ant)
    shp = (n_steps, t_int // 2, n_ant, nco, csf, 2, 2)
    order = (0, 3, 1, 5, 4, 6, 2)
    R = np.ascontiguousarray(lut_re[raw].reshape(shp).transpose(order)).reshape(n_steps, nco, K, n_ant)
    I = np.ascontiguousarray(lut_im[raw].reshape(shp).transpose(order)).reshape(n_steps, nco, K, n_ant)
    Rt = np.swapaxes(R, -1, -2)
    It = np.swapaxes(I, -1, -2)
    # [a, b] = sum_k X[k, a] Y[k, b]
    vre = Rt @ R + It @ I
    vim = It @ R - Rt @ I
    a_idx, b_idx = upper_tri(n_ant)
    out = np.stack([vre[..., a_idx, b_idx], vim[..., a_idx, b_idx]], axis=-1)
    return out.astype(np.int64)


# ---------------------------------------------------------------------------
# Calibrated correlator (cal + RFI mask on operand load)
# ---------------------------------------------------------------------------

def cal_operands(raw: np.ndarray, *, n_pkt: int, n_ant: int, n_chan: int, scale: float,
                 cal_re=None, cal_im=None, time_keep=None) -> Tuple[np.ndarray, np.ndarray]:
    """fp16 operands of the calibrated correlator, as float64
    ``(sample, ant, chan, pol)`` arrays.

    Restates the operand arithmetic of ``gemm2x_reference``
    (``src/dsart/x2/corr/gemm2x.py:346-390``): the unpacked voltage
    ``v = fp16(code * scale)`` (``unpack_int4_split``'s fp16 output,
    ``services/slow_corr_kernel.py:220-403``) times the per-(chan, pol, ant)
    complex gain ``g`` in float32 (products and sum rounded separately, no
    fused multiply-add), rounded once to fp16:
    ``x_re = fp16(v_re g_re - v_im g_im)``, ``x_im = fp16(v_re g_im + v_im g_re)``.
    ``cal_re``/``cal_im`` are ``(n_chan, 2, n_ant)`` (``compact_cal``'s layout,
    RFI-masked entries zero); ``time_keep`` ``(n_acc, n_chan, 2)`` multiplies the
    gain of every packet of accumulation ``pkt // (n_pkt // n_acc)``
    (``gemm2x.py:373-378``).  Without a gain the operands are the fp16
    voltages themselves.
    """
    re, im = voltages(raw, n_pkt, n_ant, n_chan)                  # int64 (S, ant, chan, pol)
    s32 = np.float32(scale)
    vr = (re.astype(np.float32) * s32).astype(np.float16).astype(np.float32)
    vi = (im.astype(np.float32) * s32).astype(np.float16).astype(np.float32)
    if cal_re is None:
        return vr.astype(np.float64), vi.astype(np.float64)
    gr = np.asarray(cal_re, np.float32).transpose(2, 0, 1)[None]     # (1, ant, chan, pol)
    gi = np.asarray(cal_im, np.float32).transpose(2, 0, 1)[None]
    if time_keep is not None:
        tk = np.asarray(time_keep, np.float32)
        n_acc = tk.shape[0]
        acc = (np.arange(2 * n_pkt) // 2) // (n_pkt // n_acc)       # accumulation of each sample
        tk = tk[acc][:, None, :, :]                                   # (S, 1, chan, pol)
        gr = gr * tk
        gi = gi * tk
    xr = ((vr * gr) - (vi * gi)).astype(np.float16)
    xi = ((vr * gi) + (vi * gr)).astype(np.float16)
    return xr.astype(np.float64), xi.astype(np.float64)


def correlate_float(xr: np.ndarray, xi: np.ndarray, *, t_int: int, csf: int = 1,
                    fuse_pol: bool = True) -> Tuple[np.ndarray, np.ndarray]:
    """float64 visibilities ``V[bls] = sum conj(x_b) x_a`` of float operands
    ``(sample, ant, chan, pol)`` (``gemm2x_reference``'s float64
    accumulation), and the bound ``sum |x_b| |x_a|`` of each output that a
    float32-accumulated result is compared against.  Layout as
    :func:`correlate`, complex128: ``(n_steps, n_chan // csf, nbase)`` or
    ``(n_steps, n_chan // csf, 2, nbase)``."""
    S, n_ant, n_chan, _ = xr.shape
    n_steps = S // t_int
    nco = n_chan // csf
    R = xr.reshape(n_steps, t_int, n_ant, nco, csf, 2).transpose(0, 1, 3, 4, 5, 2)
    I = xi.reshape(n_steps, t_int, n_ant, nco, csf, 2).transpose(0, 1, 3, 4, 5, 2)
    sub = "stcxpa,stcxpb->scab" if fuse_pol else "stcxpa,stcxpb->scpab"
    vre = np.einsum(sub, R, R) + np.einsum(sub, I, I)
    vim = np.einsum(sub, I, R) - np.einsum(sub, R, I)
    bound = np.einsum(sub, np.abs(R) + np.abs(I), np.abs(R) + np.abs(I))
    a_idx, b_idx = upper_tri(n_ant)
    v = vre[..., a_idx, b_idx] + 1j * vim[..., a_idx, b_idx]
    return v, bound[..., a_idx, b_idx]
