"""The 96-antenna, 384-channel node end to end against the oracle: raw 4-bit
voltages (seeded noise + three dispersed bursts of DM 200 / 600 / 1200,
widths 2 / 4 / 8 steps) in blocks of 2048 samples through the correlator,
the fused gridder + imager, the history and the fused dedispersion +
detection kernel (the node's own 16-subband, 256-DM, 6-width plan and its
default two-window-buffer layout), four blocks (statistics burn-in: 256
complete samples per DM).

The oracle chain (exact integer visibilities, float64 gridding and
``fftshift(Re(ifft2))`` imaging, float64 dedispersion, boxcars and running
statistics) runs on a subset of pixels: the bursts' 5 x 5 neighbourhoods
plus 256 seeded pixels.  Bars: visibilities bit-exact; the chunk-maximum
records of those pixels identical in (DM, time, pixel, width) with S/N within
2e-3 relative (the images carry float32 rounding); each burst found."""
import numpy as np
import pytest
import torch

from dsx import config as dcfg
from dsx import synth
from dsx.pipeline import SearchNode
from oracle import corr as ocorr
from oracle import image as oimage
from oracle.stream import OracleStream
from parity_util import compare_records, gpu_records, oracle_records

pytestmark = pytest.mark.gpu

N_BLOCKS = 4


def test_node96_short_blocks_vs_oracle():
    cfg = dcfg.search_node_short_config(2048).with_(min_stat_count=256)
    L, N = cfg.n_steps, cfg.n_grid
    assert cfg.n_ant == 96 and cfg.n_chan == 384 and L == 256
    bursts = []
    for i, (dm, w, row, col) in enumerate(((200.0, 2, 120, 140), (600.0, 4, 100, 97), (1200.0, 8, 150, 160))):
        t_dd = 3 * L + 40 + 70 * i                       # dedispersed step at the band bottom, block 3
        sweep = synth.delta_tau_us(cfg.nu_bot_GHz, float(cfg.chan_freqs_GHz[0]), dm) / synth.NATIVE_SAMPLE_US
        bursts.append(synth.Burst(dm=dm, t0_samp=int(round(t_dd * cfg.t_int - sweep)), width=w, amp=0.0,
                                  row=row, col=col, snr_target=25.0))
    sc = synth.make_scenario(cfg, bursts)
    gs = synth.GpuSynth(sc)
    rng = np.random.default_rng(96)
    pix = set(int(x) for x in rng.choice(N * N, 256, replace=False))
    for b in sc.bursts:
        for dr in range(-2, 3):
            for dc in range(-2, 3):
                pix.add((b.row + dr) * N + b.col + dc)
    pix = np.array(sorted(pix))
    node = SearchNode(cfg, sc.antpos_e, sc.antpos_n, max_block_steps=L)
    o = OracleStream(cfg, sc.antpos_e, sc.antpos_n, pix=pix)
    n_rec = 0
    found = set()
    for blk in range(N_BLOCKS):
        print(f"block {blk}", flush=True)
        raw = gs.block(blk * cfg.n_samp, cfg.n_samp)
        raw_np = raw.cpu().numpy()
        vis = node.correlate(raw, cfg.n_samp)
        img = node.image(vis, L)
        want_vis = ocorr.correlate_blas(raw_np, n_pkt=cfg.n_pkt, n_ant=cfg.n_ant, n_chan=cfg.n_chan,
                                        t_int=cfg.t_int, csf=cfg.csf)
        assert np.array_equal(vis.cpu().numpy().astype(np.int64), want_vis), f"block {blk}: visibilities"
        node.push_images(img, L)
        res = node.detect(L)
        o_img = oimage.subband_images_sparse(want_vis, o.rows, o.cols, N, cfg.n_sub, o.mask, pix=pix)
        o.push_images(o_img)
        n_rec += compare_records(gpu_records(res.records, pix), oracle_records(o, cfg.threshold),
                                 cfg.threshold, 2e-3, f"block {blk}")
        for i, b in enumerate(sc.bursts):
            d_b = int(np.argmin(np.abs(cfg.dm_trials - b.dm)))
            if any(abs(r["d"] - d_b) <= 3 and abs(int(r["p"]) // N - b.row) <= 1 and
                   abs(int(r["p"]) % N - b.col) <= 1 and r["snr"] > 15 for r in res.records):
                found.add(i)
    assert n_rec > 0
    assert found == {0, 1, 2}, f"bursts found: {sorted(found)}"
