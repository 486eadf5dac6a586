"""The dedispersion kernel instances of the uneven DM plan, against the
float64 oracle: each range of the 96-antenna node's plan (0-300 pc cm^-3,
128 trials at 1x; 300-800, 96 trials at 2x; 800-1500, 64 trials at 4x time
downsampling), 16 subbands (16 terms per dedispersed sample unpaired, 8 with
pair rows), boxcars 1-64 samples of the range's resolution (7 widths), one
full block per range (2048 / 1024 / 512 steps) for four blocks of random
images on a 16 x 16 grid, under every LDS layout the planner offers.

Bars: chunk-maximum records identical in (DM, time, pixel, width) with S/N
within 1e-4 relative; float64 mean / M2 within 1e-5 relative (the mean also within 2e-6 of the
dedispersed series' sigma)."""
import numpy as np
import pytest
import torch

from dsx import config as dcfg
from dsx import synth
from dsx.pipeline import SearchNode
from oracle.stream import OracleStream
from parity_util import compare_records, gpu_records, oracle_records

pytestmark = pytest.mark.gpu

N_BLOCKS = 4


@pytest.mark.parametrize("rng_idx", [0, 1, 2])
def test_uneven_range_every_layout_vs_oracle(rng_idx):
    full = dcfg.uneven_dm_config().with_(n_grid=16, threshold=4.5, min_stat_count=256)
    rc = full.range_configs()[rng_idx]
    f = full.dm_ranges[rng_idx][3]
    assert rc.n_sub == 16 and rc.widths == (1, 2, 4, 8, 16, 32, 64) and rc.n_steps == 2048 // f
    e, n = synth.antenna_layout(rc)
    L, N = rc.n_steps, rc.n_grid
    probe = SearchNode(rc, e, n)
    plans = probe.layouts(L)
    del probe
    assert {(p.paired, p.q_total > 0 and not p.paired) for p in plans} == \
        {(False, False), (True, False), (False, True)}
    nodes = []
    for p in plans:
        nd = SearchNode(rc, e, n)
        nd.use_plan(L, p)
        nodes.append((f"tile {p.tile} {'pairs' if p.paired else 'two-buffer' if p.q_total else 'one-buffer'}", nd))
    print(f"range {rng_idx} ({rc.n_dm} DMs, {L} steps): {[k for k, _ in nodes]}", flush=True)
    o = OracleStream(rc, e, n)
    sh = rc.shift_table()
    rng = np.random.default_rng(31 + rng_idx)
    d_pulse = rc.n_dm * 3 // 4
    total = 0
    for blk in range(N_BLOCKS):
        img = rng.standard_normal((rc.n_sub, L, N * N)).astype(np.float32)
        if blk >= 2:
            t0 = blk * L + L // 2
            for s in range(rc.n_sub):
                tt = t0 - blk * L - int(sh[d_pulse, s])
                img[s, tt:tt + 8, 141] += 1.5
        img_t = torch.from_numpy(img.reshape(rc.n_sub, L, N, N)).cuda()
        o.push_images(img.astype(np.float64))
        want = oracle_records(o, rc.threshold)
        for name, nd in nodes:
            nd.push_images(img_t, L)
            res = nd.detect(L)
            total += compare_records(gpu_records(res.records), want, rc.threshold, 1e-4, f"block {blk} {name}")
            np.testing.assert_allclose(nd.mean.cpu().numpy().T, o.st_mean, rtol=1e-5, atol=2e-6 * np.sqrt(rc.n_sub))
            np.testing.assert_allclose(nd.m2.cpu().numpy().T, o.st_m2, rtol=1e-5, atol=1e-5)
        print(f"block {blk}: {len(want)} oracle records", flush=True)
    assert total > 0
    assert any(abs(k[0] - d_pulse) <= 2 and k[2] == 141 for k in want), "pulse not among the records"
