"""Node configuration: array, band, block, grid, DM plan and detector knobs.

The reference hard-wires one op-point (96 antennas, 384 channels,
``src/dsart/common/constants.py:53-208``); the fast path here takes every
shape from a :class:`NodeConfig` so the same kernels serve the 16-antenna
CPU-sized case and the 96-antenna node.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace
from typing import Optional, Tuple

import numpy as np

from .constants import (
    DELTA_NU_CH_GHZ,
    NATIVE_SAMPLE_US,
    NTIMES_PER_PACKET,
    freq_GHz,
    nbase,
)
from .dispersion import subband_shift_table

#: Dedispersed centres trail the newest imaged step by this many steps, and
#: candidate maxima are taken over aligned chunks of the same length.
DETECT_DELAY_STEPS = 64
DETECT_CHUNK_STEPS = 64


@dataclass(frozen=True)
class NodeConfig:
    """Everything the fast path needs to know about one node stream."""

    name: str
    n_ant: int
    chan_freqs_GHz: Tuple[float, ...]      # (n_chan,) descending
    n_samp: int                            # native samples per full block
    t_int: int = 8                         # native samples per step
    csf: int = 1                           # channels summed inside the correlator
    n_sub: int = 1                         # uv planes (subbands) per step
    n_grid: int = 256
    grid_freq_GHz: Optional[float] = None  # None: each output channel's own frequency
    cell_lambda: Optional[float] = None    # None: auto-fit to the longest baseline
    dm_max: float = 500.0
    n_dm: int = 64
    dm_min: float = 0.0
    dm_endpoint: bool = True               # trials include dm_max
    # uneven plan: ((dm_lo, dm_hi, n_dm, time downsampling, endpoint), ...); empty = one range
    dm_ranges: Tuple[Tuple[float, float, int, int, bool], ...] = ()
    #: explicit DM trials (overrides dm_min / dm_max / n_dm spacing when non-empty)
    dm_list: Tuple[float, ...] = ()
    # ---- stage 1 (per-channel coarse-DM shifts before gridding)
    #: coarse DM whose delay inside each subband is removed before gridding
    #: (None: no stage 1); each group of ``stage1_csf`` channels is read
    #: ``stage1_bins`` steps later, aligning it to its subband's first group
    stage1_dm: Optional[float] = None
    stage1_csf: int = 1
    #: steps the stage-1 stream trails the raw stream (0: the largest stage-1 bin)
    stage1_delay_steps: int = 0
    #: coarse DM trials of a multi-coarse-DM node: every fine trial is searched on
    #: the stage-1 stream of its nearest coarse DM
    coarse_dms: Tuple[float, ...] = ()
    # ---- static-sky subtraction (per-pixel mean of the previous blocks)
    static_sky_blocks: int = 0             # window in blocks (0: off)
    static_sky_warmup: int = 8             # blocks passed through before subtracting
    widths: Tuple[int, ...] = (1, 2, 4, 8)
    threshold: float = 8.0
    min_stat_count: int = 512              # complete samples per DM before it is scored
    #: e-folding time of the running statistics (None: cumulative for ever);
    #: after each block the state's weight and M2 are scaled by
    #: exp(-block duration / stats_tau_s)
    stats_tau_s: Optional[float] = None
    n_top: int = 256
    radius_t: int = 128
    radius_dm: int = 5
    radius_lm: int = 3
    kernel_support: int = 1
    array_ew_m: float = 1500.0
    array_ns_m: float = 400.0
    layout_seed: int = 0xD5A1107E
    volt_seed: int = 0xD5A1107F
    sigma_volt: float = 2.5

    # ---- derived -------------------------------------------------------
    @property
    def n_chan(self) -> int:
        return len(self.chan_freqs_GHz)

    @property
    def n_pkt(self) -> int:
        return self.n_samp // NTIMES_PER_PACKET

    @property
    def n_steps(self) -> int:
        return self.n_samp // self.t_int

    @property
    def nbase(self) -> int:
        return nbase(self.n_ant)

    @property
    def n_chan_out(self) -> int:
        return self.n_chan // self.csf

    @property
    def per_sub(self) -> int:
        return self.n_chan_out // self.n_sub

    @property
    def t_step_us(self) -> float:
        return self.t_int * NATIVE_SAMPLE_US

    @property
    def block_seconds(self) -> float:
        return self.n_samp * NATIVE_SAMPLE_US * 1e-6

    @property
    def freq_out_GHz(self) -> np.ndarray:
        """Band-centre frequency of each correlator output channel.

        Mean of its ``csf`` fine channels (sparsity_pattern.py:1065-1073).
        """
        f = np.asarray(self.chan_freqs_GHz, dtype=np.float64)
        if self.csf == 1:
            return f
        return f.reshape(self.n_chan_out, self.csf).mean(axis=1)

    @property
    def grid_freq_out_GHz(self) -> np.ndarray:
        """Frequency used to scale baselines into wavelengths per output channel."""
        if self.grid_freq_GHz is None:
            return self.freq_out_GHz
        return np.full(self.n_chan_out, float(self.grid_freq_GHz), dtype=np.float64)

    @property
    def subband_ref_GHz(self) -> np.ndarray:
        """Reference frequency of each subband: its first (highest) output channel,
        or with stage 1 its first stage-1 channel group (the frequency stage 1
        aligns the subband to).

        Same rule as ``coarse_dm/dm_plan.py:179-191``.
        """
        if self.stage1_dm is not None:
            f = self.stage1_freqs_GHz
            return np.ascontiguousarray(f[:: f.size // self.n_sub][: self.n_sub])
        return np.ascontiguousarray(self.freq_out_GHz[:: self.per_sub][: self.n_sub])

    @property
    def stage1_freqs_GHz(self) -> np.ndarray:
        """Mean frequency of each group of ``stage1_csf`` channels (the summed
        channel table of ``coarse_dm/dm_plan.py:160-176``)."""
        f = np.asarray(self.chan_freqs_GHz, dtype=np.float64)
        if self.stage1_csf == 1:
            return f
        return f.reshape(-1, self.stage1_csf).mean(axis=1)

    def stage1_bins_for(self, dm: float) -> np.ndarray:
        """``(n_chan,)`` int32 stage-1 shift of every channel for coarse DM ``dm``
        (``DMPlan.delay_bins_per_subband``, ``coarse_dm/dm_plan.py:479-515``):
        ``rint(rint(K dm (nu_j^-2 - nu_ref^-2) 1e3 / 32.768 us) / t_int)`` for the
        channel's group ``j``, ``nu_ref`` the first group of its subband; two
        roundings (native sample, then step), every subband's first group 0."""
        from .constants import K_DM_MS_GHZ2_PC
        f = self.stage1_freqs_GHz
        per = f.size // self.n_sub
        ref = np.repeat(f[::per][: self.n_sub], per)
        inv_diff = (1.0 / f ** 2) - (1.0 / ref ** 2)
        delay_us = K_DM_MS_GHZ2_PC * float(dm) * inv_diff * 1e3
        nat = np.rint(delay_us / NATIVE_SAMPLE_US)
        bins = np.rint(nat / self.t_int).astype(np.int32)
        return np.ascontiguousarray(np.repeat(bins, self.stage1_csf))

    @property
    def stage1_bins(self) -> np.ndarray:
        if self.stage1_dm is None:
            return np.zeros(self.n_chan, np.int32)
        return self.stage1_bins_for(self.stage1_dm)

    @property
    def stage1_delay(self) -> int:
        """Steps the stage-1 stream trails the raw stream: output step ``t`` of
        a block reads channel ``c`` at ``t + bins[c]``, so a block emits the
        steps ``[T0 - M, T0 + L - M)`` and keeps its last ``M`` raw steps."""
        if self.stage1_dm is None:
            return 0
        return max(int(self.stage1_delay_steps), int(self.stage1_bins.max()))

    def coarse_groups(self) -> Tuple["NodeConfig", ...]:
        """One stage-1 configuration per coarse DM of a multi-coarse-DM node:
        the fine trials whose nearest coarse DM it is (ties to the lower one),
        all groups delayed by the largest stage-1 bin of any coarse DM."""
        coarse = np.asarray(self.coarse_dms, dtype=np.float64)
        fine = self.dm_trials
        f2c = np.argmin(np.abs(fine[:, None] - coarse[None, :]), axis=1)
        M = max(int(self.stage1_bins_for(float(c)).max()) for c in coarse)
        out = []
        for k, c in enumerate(coarse):
            sel = fine[f2c == k]
            if sel.size == 0:
                continue
            out.append(replace(self, coarse_dms=(), dm_ranges=(), dm_list=tuple(float(x) for x in sel),
                               n_dm=int(sel.size), stage1_dm=float(c), stage1_delay_steps=M))
        return tuple(out)

    @property
    def fine_to_coarse(self) -> np.ndarray:
        coarse = np.asarray(self.coarse_dms, dtype=np.float64)
        return np.argmin(np.abs(self.dm_trials[:, None] - coarse[None, :]), axis=1)

    @property
    def nu_bot_GHz(self) -> float:
        return float(self.chan_freqs_GHz[-1])

    @property
    def dm_trials(self) -> np.ndarray:
        """DM trials of the node: the concatenated ranges of an uneven plan, else
        ``n_dm`` evenly spaced values from ``dm_min`` (to ``dm_max`` when
        ``dm_endpoint``)."""
        if self.dm_ranges:
            return np.concatenate([r.dm_trials for r in self.range_configs()])
        if self.dm_list:
            return np.asarray(self.dm_list, dtype=np.float64)
        return np.linspace(float(self.dm_min), float(self.dm_max), int(self.n_dm), endpoint=self.dm_endpoint)

    def range_configs(self) -> Tuple["NodeConfig", ...]:
        """One single-range configuration per DM range of an uneven plan: the
        range's trials, its time resolution (``t_int`` times the downsampling)
        and the node's widths in samples of that resolution."""
        out = []
        for lo, hi, n, f, endpoint in self.dm_ranges:
            out.append(replace(self, t_int=self.t_int * int(f), dm_min=float(lo), dm_max=float(hi), n_dm=int(n),
                               dm_endpoint=bool(endpoint), dm_ranges=()))
        return tuple(out)

    @property
    def range_factors(self) -> Tuple[int, ...]:
        return tuple(int(r[3]) for r in self.dm_ranges) if self.dm_ranges else (1,)

    def shift_table(self) -> np.ndarray:
        return subband_shift_table(self.dm_trials, self.subband_ref_GHz,
                                   self.nu_bot_GHz, self.t_step_us)

    @property
    def w_max(self) -> int:
        return int(max(self.widths))

    def stats_decay(self, n_steps: int) -> float:
        """Weight kept by the running statistics after a block of ``n_steps``
        steps: ``exp(-n_steps * t_step / stats_tau_s)``, 1 without forgetting."""
        if self.stats_tau_s is None:
            return 1.0
        return math.exp(-n_steps * self.t_step_us * 1e-6 / float(self.stats_tau_s))

    def validate(self, detect: bool = True) -> None:
        """Check shapes; ``detect`` adds the block/width rules of the detector."""
        if self.n_ant < 2:
            raise ValueError("n_ant must be >= 2")
        if self.n_samp % (NTIMES_PER_PACKET * self.t_int) != 0:
            raise ValueError("n_samp must be a multiple of 2 * t_int")
        if self.t_int % NTIMES_PER_PACKET != 0:
            raise ValueError("t_int must be a multiple of 2")
        if self.n_chan % self.csf != 0:
            raise ValueError("csf must divide n_chan")
        if self.n_chan_out % self.n_sub != 0:
            raise ValueError("n_sub must divide n_chan / csf")
        if self.n_grid & (self.n_grid - 1) or not 8 <= self.n_grid <= 1024:
            raise ValueError("n_grid must be a power of two in [8, 1024]")
        if self.kernel_support != 1:
            raise ValueError("only the nearest-cell gridder (kernel_support=1) is supported")
        if not detect:
            return
        if self.n_steps % DETECT_CHUNK_STEPS != 0:
            raise ValueError(f"steps per block must be a multiple of {DETECT_CHUNK_STEPS}")
        ws = list(self.widths)
        if ws != sorted(set(ws)) or ws[0] != 1 or any(w & (w - 1) for w in ws):
            raise ValueError("widths must be ascending powers of two starting at 1")
        if ws != [1 << k for k in range(len(ws))]:
            raise ValueError("widths must be 1, 2, 4, ... without gaps")
        if self.w_max > 2 * DETECT_DELAY_STEPS:
            raise ValueError("widest boxcar exceeds twice the detection delay")
        if self.stats_tau_s is not None:
            if not self.stats_tau_s > 0:
                raise ValueError(f"stats_tau_s={self.stats_tau_s} must be > 0")
            lam = self.stats_decay(self.n_steps)
            n_sat = self.n_steps * lam / (1.0 - lam)        # weight the state saturates at
            if n_sat < self.min_stat_count:
                raise ValueError(f"stats_tau_s={self.stats_tau_s}: the decayed sample weight saturates at "
                                 f"{n_sat:.0f} < min_stat_count={self.min_stat_count}; no DM would be scored")
        if self.dm_list and len(self.dm_list) != self.n_dm:
            raise ValueError(f"dm_list holds {len(self.dm_list)} trials, n_dm={self.n_dm}")
        if self.stage1_dm is not None or self.coarse_dms:
            if self.stage1_csf < 1 or self.n_chan % self.stage1_csf or (self.n_chan // self.stage1_csf) % self.n_sub:
                raise ValueError(f"stage1_csf={self.stage1_csf} must divide the channels into whole groups "
                                 f"per subband")
            if self.stage1_csf > self.csf and self.stage1_csf % self.csf:
                raise ValueError("a stage-1 channel group must hold whole correlator output channels")
            if self.dm_ranges:
                raise ValueError("stage 1 and an uneven DM plan do not combine")
        if self.stage1_dm is not None and self.stage1_delay > self.n_steps:
            raise ValueError(f"stage-1 delay {self.stage1_delay} steps exceeds the block ({self.n_steps} steps)")
        if self.static_sky_blocks < 0 or self.static_sky_warmup < 0:
            raise ValueError("static_sky_blocks and static_sky_warmup must be >= 0")
        for lo, hi, n, f, _ in self.dm_ranges:
            if int(f) < 1 or self.n_steps % (DETECT_CHUNK_STEPS * int(f)) != 0:
                raise ValueError(f"DM range {lo}-{hi}: downsampling {f} must leave a multiple of "
                                 f"{DETECT_CHUNK_STEPS} steps per block")
            if int(n) < 1 or not hi > lo:
                raise ValueError(f"DM range {lo}-{hi} with {n} trials")

    def with_(self, **kw) -> "NodeConfig":
        return replace(self, **kw)


def small_node_config() -> NodeConfig:
    """16 ant x 64 ch (2.93 MHz, 1498.75 MHz down) x 2048 samples, 64 DMs 0-500."""
    dnu = 96 * DELTA_NU_CH_GHZ
    top = freq_GHz(0, 0)
    freqs = tuple(top - c * dnu for c in range(64))
    return NodeConfig(
        name="16ant_64ch", n_ant=16, chan_freqs_GHz=freqs, n_samp=2048, t_int=8,
        csf=1, n_sub=64, n_grid=64, grid_freq_GHz=None,
        dm_max=500.0, n_dm=64, widths=(1, 2, 4, 8), threshold=8.0,
    )


def search_node_config() -> NodeConfig:
    """96 ant x 384 ch (chgroup 0) x 16384 samples; 16 subbands, 256^2 grid, 256 DMs."""
    freqs = tuple(freq_GHz(0, c) for c in range(384))
    return NodeConfig(
        name="96ant_384ch", n_ant=96, chan_freqs_GHz=freqs, n_samp=16384, t_int=8,
        csf=24, n_sub=16, n_grid=256, grid_freq_GHz=1.4,
        dm_max=1500.0, n_dm=256, widths=(1, 2, 4, 8, 16, 32), threshold=8.0,
        layout_seed=0xD5A11096, volt_seed=0xD5A11097,
    )


def uneven_dm_config() -> NodeConfig:
    """96-antenna 384-channel inputs with the uneven DM plan: 0-300 (128 trials, full time
    resolution), 300-800 (96 trials, 2x downsampled), 800-1500 (64 trials, 4x
    downsampled); boxcars 1-64 samples of each range's resolution."""
    return search_node_config().with_(
        name="96ant_384ch_uneven", widths=(1, 2, 4, 8, 16, 32, 64), n_dm=288, dm_max=1500.0,
        dm_ranges=((0.0, 300.0, 128, 1, False), (300.0, 800.0, 96, 2, False), (800.0, 1500.0, 64, 4, True)))


def coarse_dm_config() -> NodeConfig:
    """The 96-antenna 384-channel node with stage 1: four coarse DMs (187.5, 562.5,
    937.5, 1312.5) remove the delay inside each subband per 8-channel group
    (the reference's 48 summed channels of a chgroup, ``chan_sum_factor`` 8)
    before gridding; each of the 256 fine trials 0-1500 is searched on its
    nearest coarse DM's stream; static sky subtracted over 8 blocks."""
    return search_node_config().with_(
        name="96ant_384ch_stage1", coarse_dms=(187.5, 562.5, 937.5, 1312.5), stage1_csf=8,
        static_sky_blocks=8, static_sky_warmup=8)


def search_node_short_config(n_samp: int = 1024) -> NodeConfig:
    """The 96-antenna node at a short block, for parity tests."""
    return search_node_config().with_(name=f"96ant_384ch_{n_samp}", n_samp=n_samp)
