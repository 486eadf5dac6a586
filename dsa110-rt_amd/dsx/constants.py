"""Numerical conventions of the DSA-110 search-node fast path.

Values follow ``src/dsart/common/constants.py`` of dsa110-rt (native sample
period ``:24``, band table ``:53-156``, K_DM ``:163``, fluff scale ``:373``,
grid defaults ``:409-417``).  Only the constants the hot path needs live here.
"""
from __future__ import annotations

from typing import Final

#: SNAP voltage sample period (us).  constants.py:24
NATIVE_SAMPLE_US: Final[float] = 32.768

#: Native samples per SNAP packet (the ``2t`` axis of the on-wire layout).
NTIMES_PER_PACKET: Final[int] = 2

#: Voltage polarisations.
NPOL: Final[int] = 2

#: Dispersion constant, ms GHz^2 cm^3 / pc.  constants.py:163
K_DM_MS_GHZ2_PC: Final[float] = 4.148808

#: Vacuum speed of light, m/s.
SPEED_OF_LIGHT_M_S: Final[float] = 299_792_458.0

#: Legacy 4-bit fluff scale applied by ``unpack_int4_split``.  constants.py:373
LEGACY_FLUFF_SCALE: Final[float] = 0.05

#: Configured band (corr_setup_96.yaml): top 1.530 GHz, 250 MHz, 8192 channels.
F0_CONF_GHZ: Final[float] = 1.530
BW_CONF_GHZ: Final[float] = 0.250
N_CHAN_CONF: Final[int] = 8192
DELTA_NU_CH_GHZ: Final[float] = BW_CONF_GHZ / N_CHAN_CONF

#: Processed band: 16 chgroups of 384 channels starting at system channel 1024.
N_CHGROUP: Final[int] = 16
NCHAN_PER_CHGROUP: Final[int] = 384
CH0_CHGROUP_0: Final[int] = 1024

#: uv grid side and gridding-kernel support (nearest cell).  constants.py:409-417
N_GRID_DEFAULT: Final[int] = 256
KERNEL_SUPPORT_DEFAULT: Final[int] = 1


def freq_GHz(chgroup: int, local_ch: int) -> float:
    """Frequency of ``(chgroup, local_ch)``; channels descend in frequency.

    Same closed form as ``constants.py:136-145``:
    ``f0 - (1024 + 384 g + ch) * dnu``.
    """
    if not 0 <= chgroup < N_CHGROUP:
        raise ValueError(f"chgroup={chgroup}, expected 0..{N_CHGROUP - 1}")
    if not 0 <= local_ch < NCHAN_PER_CHGROUP:
        raise ValueError(f"local_ch={local_ch}, expected 0..{NCHAN_PER_CHGROUP - 1}")
    ch_sys = CH0_CHGROUP_0 + chgroup * NCHAN_PER_CHGROUP + local_ch
    return F0_CONF_GHZ - ch_sys * DELTA_NU_CH_GHZ


def nbase(n_ant: int) -> int:
    """Baselines including autocorrelations, ``n (n + 1) / 2``."""
    return n_ant * (n_ant + 1) // 2
