"""The dedispersion LDS plan (``dsx_dedisp_plan``, host code) checked on the
CPU: a numpy model of the kernel's staging and summation, driven only by the
plan's tables, must reproduce direct shift-and-add dedispersion for every DM
and every step a tile reads, with every LDS index inside its row."""
import numpy as np
import pytest
import torch

from dsx import config as dcfg
from dsx.dedisp import plan_dedisp
from oracle import search as osearch


def _bf16(x):
    """float -> BF16 (round to nearest even) -> float."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _model(plan, sh_sd, n_widths, tt, rng):
    """Stage random subband windows, build pair rows, sum the terms."""
    n_sub, n_dm = sh_sd.shape
    h2 = (1 << (n_widths - 1)) // 2 if n_widths > 1 else 0
    ww = 2 * h2
    R = np.zeros(plan.r_total)
    shmax = sh_sd.max(axis=1)
    # absolute step tau of subband s sits at R[off_s + tau - (c0 - h2 - shmax_s)], c0 = 0
    T0 = -h2 - int(shmax.max()) - 4
    series = rng.standard_normal((n_sub, tt + ww + int(shmax.max()) - int(min(sh_sd.min(), 0)) + 8))  # tau at tau - T0
    for s in range(n_sub):
        off, smax, ln = plan.sub_info[s]
        assert smax == shmax[s] and ln == tt + ww + shmax[s] - sh_sd[s].min()
        tau = np.arange(ln) - h2 - smax
        R[off:off + ln] = series[s, tau - T0]
    if plan.paired:
        n_el = plan.q_total * (2 if plan.rows_bf16 else 1)      # BF16 rows: two per 4-byte word
        X = np.zeros(n_el)
        for qoff, ln, a0, b0, blo, bhi in plan.rows:
            assert 0 <= a0 and a0 + ln <= plan.r_total and qoff + ln <= n_el
            e = np.arange(ln)
            v = R[a0 + e].copy()
            ok = (b0 + e >= blo) & (b0 + e < bhi)
            v[ok] += R[(b0 + e)[ok]]
            X[qoff:qoff + ln] = _bf16(v) if plan.rows_bf16 else v
    else:
        X = R
    t_rel = np.arange(-h2, tt + h2)
    D = np.zeros((n_dm, t_rel.size))
    for g in range(plan.n_terms):
        idx = plan.tbase[g][:, None] + t_rel[None, :]
        assert idx.min() >= 0 and idx.max() < X.size
        D += X[idx]
    want = np.zeros_like(D)
    for s in range(n_sub):
        want += series[s, (t_rel[None, :] - sh_sd[s][:, None]) - T0]
    return D, want


@pytest.mark.parametrize("name,mode,dtype", [("96ant", 0, torch.float16), ("96ant", 2, torch.float16),
                                              ("96ant", 1, torch.float32), ("96ant", 3, torch.float32),
                                              ("96ant", 4, torch.float32), ("96ant", 5, torch.float16),
                                              ("16ant", 0, torch.float16)])
def test_plan_reproduces_direct_dedispersion(name, mode, dtype):
    cfg = dcfg.search_node_config() if name == "96ant" else dcfg.small_node_config()
    sh_sd = np.ascontiguousarray(cfg.shift_table().T)
    nw = len(cfg.widths)
    plan = plan_dedisp(sh_sd, n_widths=nw, n_centres=cfg.n_steps, ring_dtype=dtype,
                       lds_budget=80 * 1024 if name == "96ant" else 160 * 1024, mode=mode)
    assert plan.lds_bytes <= (80 if name == "96ant" else 160) * 1024
    assert cfg.n_steps % plan.tile == 0 and plan.tile % 64 == 0
    if mode == 1:
        assert plan.paired
    if mode == 2:
        assert not plan.paired and plan.n_terms == cfg.n_sub and plan.q_total == 0
    if mode == 3:     # two window buffers of the same layout; the second sits where pair rows would
        assert not plan.paired and plan.n_terms == cfg.n_sub and plan.q_total == plan.r_total
        assert plan.lds_bytes >= 2 * 4 * plan.r_total
    assert plan.dma == (mode in (3, 4)) and plan.rows_bf16 == (mode == 5)
    if mode == 4:
        assert plan.paired and plan.n_terms == (cfg.n_sub + 1) // 2
    D, want = _model(plan, sh_sd, nw, plan.tile, np.random.default_rng(3))
    if plan.rows_bf16:      # each of the n_sub / 2 rows carries one BF16 rounding (2^-9 relative)
        assert np.abs(D - want).max() <= plan.n_terms * 2.0 ** -8 * 4.0
        assert np.abs(D - want).max() > 0
    else:
        np.testing.assert_allclose(D, want, rtol=1e-12, atol=1e-12)


def _bank_conflicts(plan, n_dm, t_span):
    """Extra LDS cycles of the lanes' term reads: per term, wave of 64 DMs
    and step, (largest number of distinct addresses that share one of 64
    4-byte banks) - 1."""
    extra = 0
    for g in range(plan.n_terms):
        for w0 in range(0, n_dm, 64):
            base = plan.tbase[g][w0:w0 + 64]
            for t in range(t_span):
                addr = np.unique(base + t)
                extra += np.bincount(addr % 64).max() - 1
    return extra


def test_mode4_rows_spread_a_wave_over_the_banks():
    """The node's plan: with mode 4's row offsets the 64 DM lanes of a wave
    reading one pair's rows at one step rarely share a bank at different
    addresses (rounding makes a DM's delta jitter between neighbouring rows,
    so one wave can touch rows whose lanes span more than 64 addresses);
    the same rows packed back to back (mode 1) conflict on half the reads."""
    cfg = dcfg.search_node_config()
    sh_sd = cfg.shift_table().T
    kw = dict(n_widths=6, n_centres=2048, ring_dtype=torch.float32, lds_budget=160 * 1024)
    p4 = plan_dedisp(sh_sd, mode=4, **kw)
    p1 = plan_dedisp(sh_sd, mode=1, **kw)
    assert p4.paired and p1.paired and p4.n_rows == p1.n_rows == 96
    base_reads = p4.n_terms * (cfg.n_dm // 64) * 64
    c4, c1 = _bank_conflicts(p4, cfg.n_dm, 64), _bank_conflicts(p1, cfg.n_dm, 64)
    print(f"extra bank cycles per read: mode 4 {c4 / base_reads:.3f}, mode 1 {c1 / base_reads:.3f}")
    assert c4 <= 0.15 * base_reads and c4 * 3 <= c1
    # the padding costs at most 63 floats per row
    assert p4.q_total - p1.q_total <= 63 * p4.n_rows


def test_plan_96ant_is_paired_with_12_deltas():
    cfg = dcfg.search_node_config()
    plan = plan_dedisp(cfg.shift_table().T, n_widths=6, n_centres=2048, ring_dtype=torch.float16,
                       lds_budget=80 * 1024)
    assert plan.paired and plan.n_terms == 8 and plan.n_rows == 96 and plan.lds_bytes <= 80 * 1024


def test_plan_odd_subbands_and_signed_shifts():
    rng = np.random.default_rng(11)
    sh = np.sort(rng.integers(-20, 40, size=(5, 37)), axis=1).astype(np.int32)
    for mode in (1, 2, 4):
        plan = plan_dedisp(sh, n_widths=4, n_centres=256, ring_dtype=torch.float32, lds_budget=160 * 1024,
                           mode=mode)
        D, want = _model(plan, sh, 4, plan.tile, rng)
        np.testing.assert_allclose(D, want, rtol=1e-12, atol=1e-12)


def test_plan_errors():
    sh = np.zeros((4, 8), np.int32)
    with pytest.raises(RuntimeError, match="multiple of 64"):
        plan_dedisp(sh, n_widths=3, n_centres=100, ring_dtype=torch.float16, lds_budget=65536)
    with pytest.raises(RuntimeError, match="LDS"):
        plan_dedisp(sh, n_widths=3, n_centres=64, ring_dtype=torch.float16, lds_budget=100)
    with pytest.raises(RuntimeError, match="float32 history"):
        plan_dedisp(sh, n_widths=3, n_centres=64, ring_dtype=torch.float16, lds_budget=65536, mode=3)
