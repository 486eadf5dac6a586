"""Shift-table arithmetic pinned by the reference's own DM plan file.

``tests/golden/dm_plan_tables.npz`` holds the three integer tables of
dsa110-rt's ``configs/dm_plan.npz`` (built by ``tools/build_dm_plan.py:322-387``
for 690 fine / 16 coarse DM trials, t_fast 262.144 us, t_search 524.288 us);
``tests/golden/make_dm_plan_fixture.py`` extracted them.  Every entry has to
land on the same integer, so a change in the delay's operation order or the
rounding rule shows up here."""
from pathlib import Path

import numpy as np
import pytest

from dsx import config as dcfg
from dsx.constants import NCHAN_PER_CHGROUP, N_CHGROUP, freq_GHz
from dsx.dedisp import (NU_BOT_PROC_GHZ, NU_CHGROUP_BOT_GHZ, NU_CHGROUP_TOP_GHZ, build_time_shifts,
                        compute_delay_native_samples_table, compute_time_shift_search, node_shift_table)
from oracle import search as osearch

G = np.load(Path(__file__).parent / "golden" / "dm_plan_tables.npz", allow_pickle=False)
T_FAST = float(G["t_int_fast_us"])
T_SEARCH = float(G["t_int_search_us"])


def test_fixture_shapes():
    assert G["time_shift_corr_stage1"].shape == (16, 384, 16)
    assert G["time_shift_corr_stage2"].shape == (16, 16)
    assert G["time_shift_search"].shape == (690, 16)
    assert (T_FAST, T_SEARCH) == (262.144, 524.288)


def test_build_time_shifts_reproduces_reference_plan():
    s1, s2, ss = build_time_shifts(G["fine_dm"], G["coarse_dm"], G["fine_to_coarse"], T_FAST, T_SEARCH)
    assert np.array_equal(s1, G["time_shift_corr_stage1"])
    assert np.array_equal(s2, G["time_shift_corr_stage2"])
    assert np.array_equal(ss, G["time_shift_search"])


def test_oracle_shift_table_reproduces_reference_plan():
    """The oracle's delay arithmetic lands on every reference integer."""
    bot = np.asarray(NU_CHGROUP_BOT_GHZ)
    ddm = G["fine_dm"] - G["coarse_dm"][G["fine_to_coarse"]]
    assert np.array_equal(osearch.shift_table(ddm, bot, NU_BOT_PROC_GHZ, T_SEARCH), G["time_shift_search"])
    assert np.array_equal(osearch.shift_table(G["coarse_dm"], bot, NU_BOT_PROC_GHZ, T_FAST).T,
                          G["time_shift_corr_stage2"])
    for g in (0, 7, 15):
        nu = np.array([freq_GHz(g, c) for c in range(NCHAN_PER_CHGROUP)])
        assert np.array_equal(osearch.shift_table(G["coarse_dm"], nu, bot[g], T_FAST).T,
                              G["time_shift_corr_stage1"][g])


def test_compute_time_shift_search_modes():
    fine, coarse, f2c = G["fine_dm"], G["coarse_dm"], G["fine_to_coarse"]
    # residual mode with the chgroup-bottom reference is the plan file's table
    tab = compute_time_shift_search(coarse_dm_pc_cm3=coarse, fine_dm_pc_cm3=fine, fine_to_coarse=f2c,
                                    t_int_search_us=T_SEARCH, nu_chgroup_top_GHz=NU_CHGROUP_BOT_GHZ)
    assert np.array_equal(tab.shifts, G["time_shift_search"])
    assert not tab.coarse_offset_baked and tab.shifts.dtype == np.int32
    # include_coarse_offset: one rounding of the full fine delay from each TOP
    full = compute_time_shift_search(coarse_dm_pc_cm3=coarse, fine_dm_pc_cm3=fine, fine_to_coarse=f2c,
                                     t_int_search_us=T_SEARCH, include_coarse_offset=True)
    assert full.coarse_offset_baked
    assert np.array_equal(full.shifts, osearch.shift_table(fine, np.asarray(NU_CHGROUP_TOP_GHZ),
                                                           NU_BOT_PROC_GHZ, T_SEARCH))
    # merged rounding: fine rint minus the corr-side coarse rint; |merged - residual| <= 1
    mer = compute_time_shift_search(coarse_dm_pc_cm3=coarse, fine_dm_pc_cm3=fine, fine_to_coarse=f2c,
                                    t_int_search_us=T_SEARCH, merge_coarse_rounding=True, t_int_corr_us=T_SEARCH)
    top = np.asarray(NU_CHGROUP_TOP_GHZ)
    want = (osearch.shift_table(fine, top, NU_BOT_PROC_GHZ, T_SEARCH)
            - osearch.shift_table(coarse, top, NU_BOT_PROC_GHZ, T_SEARCH)[f2c])
    assert np.array_equal(mer.shifts, want)
    res = compute_time_shift_search(coarse_dm_pc_cm3=coarse, fine_dm_pc_cm3=fine, fine_to_coarse=f2c,
                                    t_int_search_us=T_SEARCH)
    assert np.abs(mer.shifts.astype(int) - res.shifts).max() <= 1
    # n_sub > 1: subband streams g * n_sub + s
    refs = np.array([[freq_GHz(g, 96 * s) for s in range(4)] for g in range(N_CHGROUP)])
    sub = compute_time_shift_search(coarse_dm_pc_cm3=coarse, fine_dm_pc_cm3=fine[:40], fine_to_coarse=f2c[:40],
                                    t_int_search_us=T_SEARCH, merge_coarse_rounding=True, t_int_corr_us=T_SEARCH,
                                    n_sub=4, nu_subband_ref_GHz=refs)
    assert sub.shifts.shape == (40, 64)
    want = (osearch.shift_table(fine[:40], refs.reshape(-1), NU_BOT_PROC_GHZ, T_SEARCH)
            - np.repeat(osearch.shift_table(coarse, top, NU_BOT_PROC_GHZ, T_SEARCH)[f2c[:40]], 4, axis=1))
    assert np.array_equal(sub.shifts, want)


def test_compute_time_shift_search_errors():
    fine, coarse, f2c = G["fine_dm"], G["coarse_dm"], G["fine_to_coarse"]
    with pytest.raises(ValueError):
        compute_time_shift_search(coarse_dm_pc_cm3=coarse, fine_dm_pc_cm3=fine, fine_to_coarse=f2c[:-1])
    with pytest.raises(ValueError):
        compute_time_shift_search(coarse_dm_pc_cm3=coarse, fine_dm_pc_cm3=fine, fine_to_coarse=f2c,
                                  merge_coarse_rounding=True)
    with pytest.raises(ValueError):
        compute_time_shift_search(coarse_dm_pc_cm3=coarse, fine_dm_pc_cm3=fine, fine_to_coarse=f2c,
                                  merge_coarse_rounding=True, include_coarse_offset=True, t_int_corr_us=1.0)
    with pytest.raises(ValueError):
        compute_time_shift_search(coarse_dm_pc_cm3=coarse, fine_dm_pc_cm3=fine, fine_to_coarse=f2c, n_sub=2)


def test_delay_native_samples_table():
    nu = np.array([[freq_GHz(g, c) for c in range(NCHAN_PER_CHGROUP)] for g in range(N_CHGROUP)])
    tab = compute_delay_native_samples_table(G["coarse_dm"], nu)
    assert tab.shape == (16, 384, 16) and tab.dtype == np.int64
    assert (tab[:, 0, :] == 0).all() and (tab >= 0).all()
    assert (np.diff(tab, axis=1) >= 0).all() and (np.diff(tab, axis=2) >= 0).all()
    for g in (0, 15):      # dtau(nu_c, nu_top) = -dtau(nu_top, nu_c); rint is odd
        assert np.array_equal(tab[g], -osearch.shift_table(G["coarse_dm"], nu[g], nu[g, 0], 32.768).T)


def test_node_shift_table_matches_oracle():
    cfg = dcfg.search_node_config()
    assert np.array_equal(node_shift_table(cfg.dm_trials, cfg.subband_ref_GHz, cfg.nu_bot_GHz, cfg.t_step_us),
                          osearch.shift_table(cfg.dm_trials, cfg.subband_ref_GHz, cfg.nu_bot_GHz, cfg.t_step_us))
    assert np.array_equal(node_shift_table(cfg.dm_trials, cfg.subband_ref_GHz, cfg.nu_bot_GHz, cfg.t_step_us),
                          cfg.shift_table())
