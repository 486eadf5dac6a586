"""The dedispersion kernel instances the 96-antenna node runs, against the
float64 oracle: the node's own DM plan (16 subbands of 24 x 30.5 kHz
channels, 256 DM trials 0-1500 pc cm^-3, boxcars 1-32 = 6 widths, 16 terms
per dedispersed sample unpaired / 8 with subband-pair rows), four blocks of
2048 steps of random images on a 16 x 16 grid, under every LDS layout the
planner offers (one window buffer, two window buffers with LDS-direct loads,
subband-pair rows; every time tile) for float32 and fp16 histories.

The images carry per-pixel offsets of up to 8 sigma in every subband
(dedispersed means up to 32 sigma), so the float32 centring of the
statistics and boxcars is exercised, and a dispersed pulse at DM index 200
in blocks 2 and 3.  Bars: chunk-maximum records identical in (DM, time,
pixel, width) with S/N within 1e-4 relative; float64 mean / M2 within 1e-5
relative (the mean also within 2e-6 of the dedispersed series' sigma)."""
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
DM_PULSE, PIX_PULSE = 200, 77


def _images(cfg, rng, blk, offsets, sh):
    L, N = cfg.n_steps, cfg.n_grid
    img = rng.standard_normal((cfg.n_sub, L, N * N)).astype(np.float32)
    img += offsets[None, None, :]
    if blk >= 2:
        t0 = blk * L + 700                      # arrival at the band bottom
        for s in range(cfg.n_sub):
            tt = t0 - blk * L - int(sh[DM_PULSE, s])
            img[s, tt:tt + 4, PIX_PULSE] += 2.0
    return img


@pytest.mark.parametrize("ring_dtype", [torch.float32, torch.float16])
def test_node_plan_every_layout_vs_oracle(ring_dtype):
    cfg = dcfg.search_node_config().with_(n_grid=16, threshold=4.5, min_stat_count=1024)
    assert cfg.n_sub == 16 and cfg.n_dm == 256 and cfg.widths == (1, 2, 4, 8, 16, 32) and cfg.n_steps == 2048
    e, n = synth.antenna_layout(cfg)
    L, N = cfg.n_steps, cfg.n_grid
    probe = SearchNode(cfg, e, n, ring_dtype=ring_dtype)
    plans = probe.layouts(L)
    del probe
    kinds = {(p.paired, p.dma) for p in plans}
    if ring_dtype == torch.float32:
        assert kinds == {(False, False), (True, False), (False, True), (True, True)}, kinds
        assert {64, 128, 256} <= {p.tile for p in plans if p.dma and not p.paired}
    else:
        assert kinds == {(False, False), (True, False)}, kinds
    nodes = []
    for p in plans:
        nd = SearchNode(cfg, e, n, ring_dtype=ring_dtype)
        nd.use_plan(L, p)
        nodes.append((f"tile {p.tile} {'pairs' if p.paired else 'two-buffer' if p.q_total else 'one-buffer'}"
                      f"{' lds-direct' if p.dma else ''}"
                      f" terms {p.n_terms}", nd))
    print(f"{ring_dtype}: {len(nodes)} layouts: {[k for k, _ in nodes]}", flush=True)
    o = OracleStream(cfg, e, n, pix=np.arange(N * N))
    sh = cfg.shift_table()
    rng = np.random.default_rng(2024)
    offsets = rng.uniform(-8.0, 8.0, N * N).astype(np.float32)
    pulse_seen = False
    for blk in range(N_BLOCKS):
        img = _images(cfg, rng, blk, offsets, sh)
        if ring_dtype == torch.float16:
            img = img.astype(np.float16).astype(np.float32)
        img_t = torch.from_numpy(img.reshape(cfg.n_sub, L, N, N)).cuda().to(ring_dtype)
        o.push_images(img.astype(np.float64))
        want = oracle_records(o, cfg.threshold)
        for name, nd in nodes:
            nd.push_images(img_t, L)
            res = nd.detect(L)
            k = compare_records(gpu_records(res.records), want, cfg.threshold, 1e-4, f"block {blk} {name}")
            np.testing.assert_allclose(nd.mean.cpu().numpy().T, o.st_mean, rtol=1e-5, atol=2e-6 * np.sqrt(cfg.n_sub),
                                       err_msg=f"block {blk} {name} mean")
            np.testing.assert_allclose(nd.m2.cpu().numpy().T, o.st_m2, rtol=1e-5, atol=1e-5,
                                       err_msg=f"block {blk} {name} m2")
            if blk >= 2:
                pulse_seen = pulse_seen or any(abs(c.dm_idx - DM_PULSE) <= 2 and
                                               int(c.l) * N + int(c.m) == PIX_PULSE for c in res.candidates)
        print(f"block {blk}: {len(want)} oracle records, every layout identical", flush=True)
    assert pulse_seen, "the injected pulse was not decoded"
