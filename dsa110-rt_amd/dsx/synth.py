"""Seeded synthetic node streams: 4-bit voltages with injected dispersed bursts.

The generator is counter-based and integer-exact, so the HIP kernel
(``dsx_synth_block``) and the numpy twin here produce bit-identical bytes for
any stream position, and a stream chopped into uneven blocks is the same
stream:

* noise: ``splitmix64(seed + i * golden)`` of the flat stream index ``i`` of
  each (sample, antenna, channel, pol, re/im) component; its top 16 bits
  index an inverse-normal table ``LUT[k] = rint(sigma * 1024 * Phi^-1((k+.5)/65536))``
  (fixed point, 1/1024 units);
* burst: a point source at pixel ``(row, col)`` adds the fixed-point phasor
  ``rint(1024 A e^{i phi[a, c]})`` to every sample of channel ``c`` inside its
  dispersed window (per-channel delay ``K DM (nu_c^-2 - nu_top^-2)``,
  cf. ``src/dsart/inject/online.py``);
* quantisation: round half to even of ``v / 1024``, clipped to [-7, 7]
  (``round(N(0, 2.5^2))`` clipped, as the configs specify), packed as one
  byte per complex sample, real in the low nibble.

On-wire layout: ``(n_pkt, n_ant, n_chan, 2 t, 2 pol)`` bytes
(``src/dsart/services/triton_unpack_onwire.py:110-114``).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

from .config import NodeConfig
from .constants import NATIVE_SAMPLE_US, NPOL, NTIMES_PER_PACKET, SPEED_OF_LIGHT_M_S
from .dispersion import delta_tau_us
from .pattern import CellMap, build_cell_map

FIXED_ONE = 1024
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def antenna_layout(cfg: NodeConfig) -> Tuple[np.ndarray, np.ndarray]:
    """Seeded ENU positions, uniform over ``array_ew_m x array_ns_m`` (U = 0)."""
    rng = np.random.default_rng(cfg.layout_seed)
    e = rng.uniform(-cfg.array_ew_m / 2, cfg.array_ew_m / 2, size=cfg.n_ant)
    n = rng.uniform(-cfg.array_ns_m / 2, cfg.array_ns_m / 2, size=cfg.n_ant)
    return e.astype(np.float64), n.astype(np.float64)


def noise_lut(sigma: float) -> np.ndarray:
    """``(65536,) int32`` inverse-normal table in 1/1024 units."""
    from scipy.special import ndtri
    k = np.arange(65536, dtype=np.float64)
    return np.rint(sigma * FIXED_ONE * ndtri((k + 0.5) / 65536.0)).astype(np.int32)


def quantised_variance(sigma: float) -> float:
    """Variance of one quantised voltage component (exact over the table)."""
    q = quantise_fixed(noise_lut(sigma).astype(np.int64))
    return float(np.mean(q.astype(np.float64) ** 2) - np.mean(q) ** 2)


def splitmix_top16(seed: int, idx: np.ndarray) -> np.ndarray:
    """Top 16 bits of splitmix64 at counter ``idx`` (uint64 arithmetic wraps)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + idx.astype(np.uint64) * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(48)).astype(np.int64)


def quantise_fixed(v: np.ndarray) -> np.ndarray:
    """Round half to even of ``v / 1024`` then clip to [-7, 7] (integer-exact)."""
    v = v.astype(np.int64)
    q = v >> 10
    rem = v - (q << 10)
    q = q + (rem > 512) + ((rem == 512) & ((q & 1) == 1))
    return np.clip(q, -7, 7)


@dataclass(frozen=True)
class Burst:
    """One dispersed point-source burst.

    ``t0_samp`` is the arrival (native samples since stream start) at the
    top channel; ``width`` is in steps of ``t_int`` samples.
    """

    dm: float
    t0_samp: int
    width: int
    amp: float
    row: int
    col: int
    snr_target: float = 0.0


@dataclass
class BurstTables:
    """Device-ready burst description (all integer)."""

    start: np.ndarray      # (n_burst, n_chan) int64 first active native sample
    length: np.ndarray     # (n_burst,) int32 active samples per channel
    phasor: np.ndarray     # (n_burst, n_ant, n_chan, 2) int32 fixed-point re/im


def burst_phases(cfg: NodeConfig, e: np.ndarray, n: np.ndarray, cell_lambda: float,
                 row: int, col: int) -> np.ndarray:
    """``(n_ant, n_chan)`` phase that images a point source at ``(row, col)``.

    A visibility ``conj(E_b) E_a`` gridded at ``U = -(N_a - N_b)/(lambda cell)``,
    ``W = -(E_a - E_b)/(lambda cell)`` images positive at ``(row, col)`` when
    ``phi_a - phi_b = -2 pi (U y0 + W x0) / N``, ``y0 = row - N/2``,
    ``x0 = col - N/2``; per antenna that is
    ``phi_a = 2 pi (N_a y0 + E_a x0) / (lambda cell N)``.
    """
    N = cfg.n_grid
    y0 = row - N // 2
    x0 = col - N // 2
    lam_out = SPEED_OF_LIGHT_M_S / (cfg.grid_freq_out_GHz * 1e9)       # (n_chan_out,)
    lam = np.repeat(lam_out, cfg.csf)                                    # (n_chan,)
    return 2.0 * np.pi * (n[:, None] * y0 + e[:, None] * x0) / (lam[None, :] * cell_lambda * N)


def burst_start_samples(cfg: NodeConfig, dm: float, t0_samp: int) -> np.ndarray:
    f = np.asarray(cfg.chan_freqs_GHz, dtype=np.float64)
    top = float(f[0])
    d = np.array([delta_tau_us(float(fc), top, dm) for fc in f]) / NATIVE_SAMPLE_US
    return (int(t0_samp) + np.rint(d)).astype(np.int64)


def build_burst_tables(cfg: NodeConfig, bursts: Sequence[Burst], e: np.ndarray,
                       n: np.ndarray, cell_lambda: float) -> BurstTables:
    nb = len(bursts)
    start = np.zeros((nb, cfg.n_chan), dtype=np.int64)
    length = np.zeros(nb, dtype=np.int32)
    ph = np.zeros((nb, cfg.n_ant, cfg.n_chan, 2), dtype=np.int32)
    for i, b in enumerate(bursts):
        start[i] = burst_start_samples(cfg, b.dm, b.t0_samp)
        length[i] = int(b.width) * cfg.t_int
        phi = burst_phases(cfg, e, n, cell_lambda, b.row, b.col)
        ph[i, :, :, 0] = np.rint(b.amp * FIXED_ONE * np.cos(phi)).astype(np.int32)
        ph[i, :, :, 1] = np.rint(b.amp * FIXED_ONE * np.sin(phi)).astype(np.int32)
    return BurstTables(start=start, length=length, phasor=ph)


def synth_block_np(cfg: NodeConfig, sample0: int, n_samp: int,
                   tables: Optional[BurstTables] = None, lut: Optional[np.ndarray] = None,
                   ) -> np.ndarray:
    """Numpy twin of ``dsx_synth_block``: raw bytes ``(n_pkt, n_ant, n_chan, 2, 2)``."""
    if sample0 % NTIMES_PER_PACKET or n_samp % NTIMES_PER_PACKET:
        raise ValueError("sample0 and n_samp must be packet aligned")
    if lut is None:
        lut = noise_lut(cfg.sigma_volt)
    na, nc = cfg.n_ant, cfg.n_chan
    samp = sample0 + np.arange(n_samp, dtype=np.int64)
    # flat component index ((((s * na + a) * nc + c) * 2 + p) * 2 + reim)
    base = (((samp[:, None, None, None] * na + np.arange(na)[None, :, None, None]) * nc
             + np.arange(nc)[None, None, :, None]) * NPOL + np.arange(NPOL)[None, None, None, :]) * 2
    v_re = lut[splitmix_top16(cfg.volt_seed, base)].astype(np.int64)
    v_im = lut[splitmix_top16(cfg.volt_seed, base + 1)].astype(np.int64)
    if tables is not None:
        for i in range(tables.length.shape[0]):
            st = tables.start[i]                                     # (nc,)
            act = (samp[:, None] >= st[None, :]) & (samp[:, None] < st[None, :] + tables.length[i])
            if not act.any():
                continue
            pr = tables.phasor[i, :, :, 0].astype(np.int64)          # (na, nc)
            pi = tables.phasor[i, :, :, 1].astype(np.int64)
            v_re += act[:, None, :, None] * pr[None, :, :, None]
            v_im += act[:, None, :, None] * pi[None, :, :, None]
    q_re = quantise_fixed(v_re)
    q_im = quantise_fixed(v_im)
    byte = ((q_re & 0xF) | ((q_im & 0xF) << 4)).astype(np.uint8)    # (S, na, nc, 2p)
    # (S, na, nc, p) -> (pkt, 2t, na, nc, p) -> (pkt, na, nc, 2t, p)
    byte = byte.reshape(n_samp // 2, 2, na, nc, NPOL).transpose(0, 2, 3, 1, 4)
    return np.ascontiguousarray(byte)


def unpack_nibbles(raw: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """``(re, im)`` int8 from packed bytes: low nibble real, high nibble imag."""
    i8 = raw.view(np.int8)
    re = ((i8 << 4).astype(np.int8) >> 4).astype(np.int8)
    im = (i8 >> 4).astype(np.int8)
    return re, im


# ---------------------------------------------------------------------------
# burst amplitude for a target matched-filter S/N
# ---------------------------------------------------------------------------


def expected_unit_snr(cfg: NodeConfig, cmap: CellMap, e: np.ndarray, n: np.ndarray,
                      burst: Burst, shifts: np.ndarray) -> float:
    """Expected best boxcar S/N of ``burst`` for ``amp = 1`` (S/N scales as amp^2).

    Signal: per output channel and step, ``2 A^2 cnt`` per baseline times the
    cosine of its cell-rounding residual phase, imaged with ``1/N^2``;
    noise: ``2 sigma_q^4 t_int npol csf`` per visibility component.  Both
    subband sums and boxcars are evaluated on the noise-free profile at the
    burst pixel for every DM trial; the largest S/N is returned.
    """
    N = cfg.n_grid
    y0 = burst.row - N // 2
    x0 = burst.col - N // 2
    phi = burst_phases(cfg, e, n, cmap.cell_lambda, burst.row, burst.col)   # (na, nch)
    from .pattern import upper_tri_indices
    a_idx, b_idx = upper_tri_indices(cfg.n_ant)
    st = burst_start_samples(cfg, burst.dm, burst.t0_samp)
    L = int(burst.width) * cfg.t_int
    t_lo = int(st.min()) // cfg.t_int
    t_hi = int(st.max() + L) // cfg.t_int + 1
    n_t = t_hi - t_lo + 1
    sig = np.zeros((cfg.n_sub, n_t))
    var_q = quantised_variance(cfg.sigma_volt)
    var_v = 2.0 * var_q * var_q * cfg.t_int * NPOL * cfg.csf
    var_img = np.zeros(cfg.n_sub)
    for co in range(cfg.n_chan_out):
        s = co // cfg.per_sub
        ok = cmap.row[:, co] >= 0
        if not ok.any():
            continue
        u = cmap.row[ok, co].astype(np.float64) - N // 2
        w = cmap.col[ok, co].astype(np.float64) - N // 2
        ch0 = co * cfg.csf
        dphi = phi[a_idx[ok], ch0] - phi[b_idx[ok], ch0]
        eps = float(np.sum(np.cos(dphi + 2.0 * np.pi * (u * y0 + w * x0) / N)))
        var_img[s] += ok.sum() * var_v / float(N) ** 4
        for ch in range(ch0, ch0 + cfg.csf):
            samp = np.arange(st[ch], st[ch] + L)
            cnt = np.bincount(samp // cfg.t_int - t_lo, minlength=n_t)[:n_t]
            sig[s] += 2.0 * cnt * eps / float(N) ** 2
    sigma_d = float(np.sqrt(var_img.sum()))
    best = 0.0
    pad = cfg.w_max + int(shifts.max()) + 2
    for d in range(shifts.shape[0]):
        D = np.zeros(n_t + 2 * pad)
        for s in range(cfg.n_sub):
            o = pad + int(shifts[d, s])
            D[o:o + n_t] += sig[s]
        P = np.concatenate([[0.0], np.cumsum(D)])
        for w in cfg.widths:
            box = P[w:] - P[:-w]
            best = max(best, float(box.max()) / (sigma_d * np.sqrt(w)))
    return best


def calibrate_amplitudes(cfg: NodeConfig, bursts: Sequence[Burst], cmap: CellMap,
                         e: np.ndarray, n: np.ndarray) -> List[Burst]:
    shifts = cfg.shift_table()
    out = []
    for b in bursts:
        unit = expected_unit_snr(cfg, cmap, e, n, Burst(b.dm, b.t0_samp, b.width, 1.0,
                                                        b.row, b.col), shifts)
        amp = float(np.sqrt(b.snr_target / unit)) if unit > 0 else 0.0
        out.append(Burst(b.dm, b.t0_samp, b.width, amp, b.row, b.col, b.snr_target))
    return out


def dedispersed_step(cfg: NodeConfig, b: Burst) -> float:
    """Step index (at the band bottom) where the burst's leading edge lands."""
    top = float(cfg.chan_freqs_GHz[0])
    d_us = delta_tau_us(cfg.nu_bot_GHz, top, b.dm)
    return (b.t0_samp + d_us / NATIVE_SAMPLE_US) / cfg.t_int


@dataclass
class Scenario:
    """A seeded stream: config, antenna layout, cell map, calibrated bursts."""

    cfg: NodeConfig
    antpos_e: np.ndarray
    antpos_n: np.ndarray
    cmap: CellMap
    bursts: List[Burst]
    tables: BurstTables

    def block(self, sample0: int, n_samp: int) -> np.ndarray:
        return synth_block_np(self.cfg, sample0, n_samp, self.tables)


def make_scenario(cfg: NodeConfig, bursts: Sequence[Burst] = ()) -> Scenario:
    cfg.validate(detect=False)
    e, n = antenna_layout(cfg)
    cmap = build_cell_map(e, n, cfg.grid_freq_out_GHz, cfg.n_grid, cell_lambda=cfg.cell_lambda)
    cal = calibrate_amplitudes(cfg, bursts, cmap, e, n) if bursts else []
    tables = build_burst_tables(cfg, cal, e, n, cmap.cell_lambda)
    return Scenario(cfg=cfg, antpos_e=e, antpos_n=n, cmap=cmap, bursts=list(cal), tables=tables)


def small_node_bursts(cfg: NodeConfig) -> List[Burst]:
    """Config 1: one burst, DM 250, width 2 steps, S/N 20, seeded pixel, in block 5.

    Block 5 is the first block whose DM-250 statistics hold 512 complete
    samples (the DM-250 sweep is ~530 steps of a 256-step block)."""
    rng = np.random.default_rng(0xB0B1)
    N = cfg.n_grid
    row = int(N // 2 + rng.integers(-N // 8, N // 8))
    col = int(N // 2 + rng.integers(-N // 8, N // 8))
    top = float(cfg.chan_freqs_GHz[0])
    sweep = delta_tau_us(cfg.nu_bot_GHz, top, 250.0) / NATIVE_SAMPLE_US
    t_dd = 5 * cfg.n_steps + cfg.n_steps // 3
    t0 = int(round(t_dd * cfg.t_int - sweep))
    return [Burst(dm=250.0, t0_samp=t0, width=2, amp=0.0, row=row, col=col, snr_target=20.0)]


def random_bursts(cfg: NodeConfig, n_burst: int, seed: int, *, dm_lo: float = 50.0,
                  dm_hi: float = 1400.0, w_lo: int = 1, w_hi: int = 16, snr_lo: float = 10.0,
                  snr_hi: float = 50.0, t_lo_step: int = 0, t_hi_step: int = 0,
                  radius_frac: float = 0.125) -> List[Burst]:
    """Seeded bursts with dedispersed steps uniform in ``[t_lo_step, t_hi_step)``."""
    rng = np.random.default_rng(seed)
    N = cfg.n_grid
    top = float(cfg.chan_freqs_GHz[0])
    out = []
    r = max(1, int(N * radius_frac))
    for _ in range(n_burst):
        dm = float(rng.uniform(dm_lo, dm_hi))
        w = int(rng.integers(w_lo, w_hi + 1))
        snr = float(rng.uniform(snr_lo, snr_hi))
        row = int(N // 2 + rng.integers(-r, r))
        col = int(N // 2 + rng.integers(-r, r))
        t_dd = int(rng.integers(t_lo_step, t_hi_step))
        sweep = delta_tau_us(cfg.nu_bot_GHz, top, dm) / NATIVE_SAMPLE_US
        t0 = int(round(t_dd * cfg.t_int - sweep))
        out.append(Burst(dm=dm, t0_samp=t0, width=w, amp=0.0, row=row, col=col, snr_target=snr))
    return out


class GpuSynth:
    """Device generator bit-identical to :func:`synth_block_np` (``dsx_synth_block``)."""

    def __init__(self, scenario: "Scenario", device: str = "cuda"):
        import torch
        from . import _lib
        self._torch = torch
        self._lib = _lib
        cfg = scenario.cfg
        self.cfg = cfg
        dev = torch.device(device)
        self.lut = torch.from_numpy(noise_lut(cfg.sigma_volt)).to(dev)
        tb = scenario.tables
        self.n_burst = int(tb.length.shape[0])
        if self.n_burst:
            self.bstart = torch.from_numpy(np.ascontiguousarray(tb.start)).to(dev)
            self.blen = torch.from_numpy(np.ascontiguousarray(tb.length)).to(dev)
            self.bph = torch.from_numpy(np.ascontiguousarray(tb.phasor)).to(dev)
        else:
            self.bstart = self.blen = self.bph = None
        self.dev = dev

    def block(self, sample0: int, n_samp: int, out=None):
        torch = self._torch
        cfg = self.cfg
        n_pkt = n_samp // 2
        shape = (n_pkt, cfg.n_ant, cfg.n_chan, 2, 2)
        if out is None:
            out = torch.empty(shape, dtype=torch.uint8, device=self.dev)
        elif tuple(out.shape) != shape or out.dtype != torch.uint8:
            raise ValueError(f"out must be uint8 {shape}")
        L = self._lib
        L.require_gpu(out)
        L.call("dsx_synth_block", L.ptr(out), int(cfg.volt_seed) & 0xFFFFFFFFFFFFFFFF, int(sample0),
               n_pkt, cfg.n_ant, cfg.n_chan, L.ptr(self.lut), self.n_burst, L.ptr(self.bstart),
               L.ptr(self.blen), L.ptr(self.bph), L.stream())
        return out
