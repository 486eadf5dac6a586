"""GPU dedispersion + detection parity against the float64 oracle, and the
whole 16-antenna stream (16 ant x 64 ch, 7 blocks, DM-250 burst) end to end.

Bars: chunk-maximum records identical in (DM, chunk, pixel, time, width)
with S/N within 1e-4 relative for float32 images; decoded candidate keys
(DM index, time bin, width, row, col) identical with S/N within 2e-3
relative; float64 statistics within 1e-5 relative (the mean, which sits near
zero for noise, within 1e-5 relative or 2e-6 of the dedispersed series'
sigma, sqrt(n_sub) for unit-variance images: the kernel sums float32 samples
centred on the running mean, and a DM whose first block held only a few
complete samples starts from a rough mean).
"""
import numpy as np
import pytest
import torch

from dsx import config as dcfg
from dsx import synth
from dsx.pipeline import SearchNode
from oracle.stream import OracleStream

pytestmark = pytest.mark.gpu


def _oracle_records(o, thr):
    L = o.last
    d, c, p = np.nonzero(L["best_snr"] > thr)
    return {(int(dd), int(L["best_t"][dd, cc, pp]), int(pp)): (int(L["best_k"][dd, cc, pp]),
                                                                  float(L["best_snr"][dd, cc, pp]))
            for dd, cc, pp in zip(d, c, p)}


def _gpu_records(rec):
    return {(int(r["d"]), int(r["t"]), int(r["p"])): (int(r["k"]), float(r["snr"])) for r in rec}


def _compare_records(g, o, thr, rtol):
    near = lambda v: abs(v - thr) < 1e-3 * max(1.0, abs(thr))
    gk = {k for k, v in g.items() if not near(v[1])}
    ok = {k for k, v in o.items() if not near(v[1])}
    assert gk == ok, f"record sets differ: gpu-only {sorted(gk - ok)[:5]} oracle-only {sorted(ok - gk)[:5]}"
    for k in gk:
        assert g[k][0] == o[k][0], f"{k}: width {g[k][0]} vs {o[k][0]}"
        assert abs(g[k][1] - o[k][1]) <= rtol * abs(o[k][1]), f"{k}: snr {g[k][1]} vs {o[k][1]}"


def _cand_keys(cands):
    return [(c.dm_idx, c.time_bin, c.width_samples, int(c.l), int(c.m)) for c in cands]


@pytest.mark.parametrize("layout", ["unpaired", "paired", "async", "autotuned", "fp16", "fp16-autotuned"])
def test_dedisp_detect_random_images_vs_oracle(layout):
    """The LDS layouts the planner offers (unpaired subband windows, subband
    pair rows, and whichever the autotuner keeps) give the oracle's records,
    for float32 and fp16 histories (fp16 cases feed images already rounded
    to fp16, so both sides sum the same values)."""
    cfg = dcfg.small_node_config().with_(n_sub=8, n_grid=16, dm_max=100.0, n_dm=64, threshold=4.5,
                                         min_stat_count=256, widths=(1, 2, 4, 8, 16))
    rng = np.random.default_rng(123)
    ant_e = rng.uniform(-500, 500, cfg.n_ant)
    ant_n = rng.uniform(-100, 100, cfg.n_ant)
    print(f"layout {layout}", flush=True)
    fp16 = layout.startswith("fp16")
    dt = torch.float16 if fp16 else torch.float32
    node = SearchNode(cfg, ant_e, ant_n, ring_dtype=dt,
                      dedisp_mode={"paired": 1, "async": 3}.get(layout, 2))
    o = OracleStream(cfg, ant_e, ant_n)
    sh = cfg.shift_table()
    L = cfg.n_steps
    N = cfg.n_grid
    total = 0
    for blk in range(5):
        img = rng.standard_normal((cfg.n_sub, L, N * N)).astype(np.float32)
        if blk >= 3:       # dispersed pulses at DM index 40, pixel 77, width 4
            t0 = blk * L + 100
            for s in range(cfg.n_sub):
                tt = t0 - blk * L - int(sh[40, s])
                if 0 <= tt < L - 4:
                    img[s, tt:tt + 4, 77] += 2.5
        if fp16:
            img = img.astype(np.float16).astype(np.float32)
        img_t = torch.from_numpy(img.reshape(cfg.n_sub, L, N, N)).cuda().to(dt)
        node.push_images(img_t, L)
        if layout.endswith("autotuned") and blk == 1:
            rows = node.autotune(L)
            assert len(rows) >= 2 and {r["paired"] for r in rows} == {False, True}
            if not fp16:
                assert any(r["async"] for r in rows)
            best = min(rows, key=lambda r: r["ms"])
            kept = node.plans[L][0]
            assert (kept.tile, kept.paired, bool(kept.q_total > 0 and not kept.paired)) == \
                (best["tile"], best["paired"], best["async"])
        if layout == "paired":
            assert node._plan_for(L)[0].paired
        if layout == "async":
            pl = node._plan_for(L)[0]
            assert not pl.paired and pl.q_total == pl.r_total > 0
        res = node.detect(L)
        o.push_images(img.astype(np.float64))
        g = _gpu_records(res.records)
        ro = _oracle_records(o, cfg.threshold)
        _compare_records(g, ro, cfg.threshold, 1e-4)
        total += len(g)
        mean = node.mean.cpu().numpy().T
        m2 = node.m2.cpu().numpy().T
        np.testing.assert_allclose(mean, o.st_mean, rtol=1e-5, atol=2e-6 * np.sqrt(cfg.n_sub))
        np.testing.assert_allclose(m2, o.st_m2, rtol=1e-5, atol=1e-5)
    assert total > 0


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_dedisp_detect_ragged_shapes_vs_oracle(mode):
    """Ragged shapes: 300 DM trials (two work-groups along DMs, the second
    with 44 live lanes), 7 subbands (the last subband pair has no second
    member), shifts up to 208 steps against 256-step blocks; pair rows
    (mode 1) and unpaired windows (mode 2) against the float64 oracle."""
    base = dcfg.small_node_config()
    cfg = base.with_(chan_freqs_GHz=base.chan_freqs_GHz[:63], n_sub=7, n_grid=16, dm_max=100.0, n_dm=300,
                     threshold=4.5, min_stat_count=256, widths=(1, 2, 4, 8, 16))
    rng = np.random.default_rng(77)
    ant_e = rng.uniform(-500, 500, cfg.n_ant)
    ant_n = rng.uniform(-100, 100, cfg.n_ant)
    print(f"mode {mode}", flush=True)
    node = SearchNode(cfg, ant_e, ant_n, ring_dtype=torch.float32, dedisp_mode=mode)
    o = OracleStream(cfg, ant_e, ant_n)
    sh = cfg.shift_table()
    L, N = cfg.n_steps, cfg.n_grid
    assert node._plan_for(L)[0].paired == (mode == 1)
    assert (node._plan_for(L)[0].q_total > 0) == (mode != 2)
    total = 0
    for blk in range(5):
        img = rng.standard_normal((cfg.n_sub, L, N * N)).astype(np.float32)
        if blk >= 3:       # dispersed pulses at DM index 290 (second DM work-group), pixel 201
            t0 = blk * L + 220      # arrival at the band bottom; every subband's copy is in this block
            for s in range(cfg.n_sub):
                tt = t0 - blk * L - int(sh[290, s])
                if 0 <= tt < L - 8:
                    img[s, tt:tt + 8, 201] += 2.0
        node.push_images(torch.from_numpy(img.reshape(cfg.n_sub, L, N, N)).cuda(), L)
        res = node.detect(L)
        o.push_images(img.astype(np.float64))
        g = _gpu_records(res.records)
        _compare_records(g, _oracle_records(o, cfg.threshold), cfg.threshold, 1e-4)
        total += len(g)
        np.testing.assert_allclose(node.mean.cpu().numpy().T, o.st_mean, rtol=1e-5, atol=2e-6 * np.sqrt(cfg.n_sub))
        np.testing.assert_allclose(node.m2.cpu().numpy().T, o.st_m2, rtol=1e-5, atol=1e-5)
    assert any(k[0] >= 256 for k in g), "no record from the second DM work-group"
    assert total > 0


@pytest.mark.parametrize("ring_dtype", [torch.float16, torch.float32])
def test_16ant_stream_candidates_match_oracle(ring_dtype):
    cfg = dcfg.small_node_config()
    sc = synth.make_scenario(cfg, synth.small_node_bursts(cfg))
    gs = synth.GpuSynth(sc)
    node = SearchNode(cfg, sc.antpos_e, sc.antpos_n, ring_dtype=ring_dtype)
    o = OracleStream(cfg, sc.antpos_e, sc.antpos_n)
    found = []
    for blk in range(7):
        print(f"block {blk}", flush=True)
        s0 = blk * cfg.n_samp
        res = node.process_block(gs.block(s0, cfg.n_samp), cfg.n_samp)
        want = o.push_raw(sc.block(s0, cfg.n_samp), cfg.n_samp)
        gk, ok = _cand_keys(res.candidates), [(c.dm_idx, c.t, c.width, c.row, c.col) for c in want]
        assert gk == ok, f"block {blk}: gpu {gk} oracle {ok}"
        for gc, oc in zip(res.candidates, want):
            assert abs(gc.snr - oc.snr) <= 2e-3 * abs(oc.snr), (gc, oc)
        found += res.candidates
    b = sc.bursts[0]
    assert any(abs(c.l - b.row) <= 1 and abs(c.m - b.col) <= 1 and c.snr > 15 for c in found), found


def _small_search_cfg():
    return dcfg.small_node_config().with_(n_sub=8, n_grid=16, dm_max=100.0, n_dm=64, threshold=4.5,
                                          min_stat_count=256, widths=(1, 2, 4, 8, 16))


@pytest.mark.parametrize("mode", [2, 3])
def test_steady_sources_keep_float64_statistics(mode):
    """Pixels holding steady sources: per-pixel offsets up to 25 sigma in every
    subband (dedispersed means up to ~70 sigma).  The kernel centres its
    float32 sums on the running mean, so records and the float64 state match
    the oracle as tightly as for zero-mean noise."""
    cfg = _small_search_cfg()
    rng = np.random.default_rng(404)
    ant_e, ant_n = rng.uniform(-500, 500, cfg.n_ant), rng.uniform(-100, 100, cfg.n_ant)
    node = SearchNode(cfg, ant_e, ant_n, dedisp_mode=mode)
    o = OracleStream(cfg, ant_e, ant_n)
    L, N = cfg.n_steps, cfg.n_grid
    offsets = rng.uniform(-25.0, 25.0, N * N).astype(np.float32)
    sh = cfg.shift_table()
    total = 0
    for blk in range(5):
        img = rng.standard_normal((cfg.n_sub, L, N * N)).astype(np.float32) + offsets[None, None, :]
        if blk >= 3:
            for s in range(cfg.n_sub):
                tt = 100 - int(sh[40, s])
                img[s, tt:tt + 4, 77] += 2.5
        node.push_images(torch.from_numpy(img.reshape(cfg.n_sub, L, N, N)).cuda(), L)
        res = node.detect(L)
        o.push_images(img.astype(np.float64))
        g = _gpu_records(res.records)
        _compare_records(g, _oracle_records(o, cfg.threshold), cfg.threshold, 1e-4)
        total += len(g)
        np.testing.assert_allclose(node.mean.cpu().numpy().T, o.st_mean, rtol=1e-5, atol=2e-6 * np.sqrt(cfg.n_sub))
        np.testing.assert_allclose(node.m2.cpu().numpy().T, o.st_m2, rtol=1e-5, atol=1e-5)
    assert total > 0


def test_record_times_past_int32():
    """A stream whose steps are past 2^31 (6.5 days at 262 us per step): the
    kernel's records hold block-relative times and the host adds the int64
    block start, so record times equal the oracle's absolute steps."""
    cfg = _small_search_cfg()
    rng = np.random.default_rng(505)
    ant_e, ant_n = rng.uniform(-500, 500, cfg.n_ant), rng.uniform(-100, 100, cfg.n_ant)
    node = SearchNode(cfg, ant_e, ant_n)
    o = OracleStream(cfg, ant_e, ant_n)
    T_start = (1 << 31) + 1000
    node.T0 = T_start
    o.start_at(T_start)           # zeros before the stream start, like the node's ring
    L, N = cfg.n_steps, cfg.n_grid
    sh = cfg.shift_table()
    total = 0
    for blk in range(4):
        img = rng.standard_normal((cfg.n_sub, L, N * N)).astype(np.float32)
        if blk >= 2:
            for s in range(cfg.n_sub):
                tt = 120 - int(sh[30, s])
                img[s, tt:tt + 4, 33] += 2.5
        node.push_images(torch.from_numpy(img.reshape(cfg.n_sub, L, N, N)).cuda(), L)
        res = node.detect(L)
        o.push_images(img.astype(np.float64))
        g = _gpu_records(res.records)
        _compare_records(g, _oracle_records(o, cfg.threshold), cfg.threshold, 1e-4)
        total += len(g)
        assert all(k[1] >= T_start - 64 for k in g)
        assert all(c.time_bin >= T_start - 64 for c in res.candidates)
    assert total > 0


def test_stats_decay_vs_oracle():
    """Running statistics with forgetting (e-folding 0.1 s, about half a
    block): six blocks whose noise level triples after the third; the GPU's
    weights, means, M2 and records equal the oracle's, and the state's sigma
    moves most of the way to the new level within three blocks (a
    cumulative state would stay near the old one)."""
    cfg = dcfg.small_node_config().with_(n_sub=8, n_grid=16, dm_max=100.0, n_dm=64, threshold=4.5,
                                         min_stat_count=128, widths=(1, 2, 4, 8, 16), stats_tau_s=0.1)
    rng = np.random.default_rng(31)
    ant_e = rng.uniform(-500, 500, cfg.n_ant)
    ant_n = rng.uniform(-100, 100, cfg.n_ant)
    node = SearchNode(cfg, ant_e, ant_n)
    o = OracleStream(cfg, ant_e, ant_n)
    L, N = cfg.n_steps, cfg.n_grid
    lam = cfg.stats_decay(L)
    assert 0.4 < lam < 0.6
    sig_hist = []
    for blk in range(6):
        img = (rng.standard_normal((cfg.n_sub, L, N * N)) * (1.0 if blk < 3 else 3.0)).astype(np.float32)
        node.push_images(torch.from_numpy(img.reshape(cfg.n_sub, L, N, N)).cuda(), L)
        res = node.detect(L)
        o.push_images(img.astype(np.float64))
        _compare_records(_gpu_records(res.records), _oracle_records(o, cfg.threshold), cfg.threshold, 1e-4)
        n_g = node.n_prev.cpu().numpy()
        np.testing.assert_allclose(n_g, o.st_n, rtol=1e-12)
        mean = node.mean.cpu().numpy().T
        m2 = node.m2.cpu().numpy().T
        np.testing.assert_allclose(mean, o.st_mean, rtol=1e-5, atol=6e-6 * np.sqrt(cfg.n_sub))
        np.testing.assert_allclose(m2, o.st_m2, rtol=1e-5, atol=1e-5)
        ok = n_g > 0
        sig_hist.append(float(np.median(np.sqrt(m2[ok] / n_g[ok, None]))))
    s_lo = np.sqrt(cfg.n_sub)                  # unit-variance images: sigma of the dedispersed series
    assert abs(sig_hist[2] / s_lo - 1.0) < 0.1
    assert sig_hist[5] / s_lo > 2.5           # cumulative weighting would give about sqrt(5) = 2.24
