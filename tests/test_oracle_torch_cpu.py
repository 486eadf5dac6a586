"""The PyTorch-CPU restatement used as the benchmark's CPU baseline agrees
with the numpy / C oracle: exact visibilities, images within 1e-5 of the
float64 imager relative to the peak, and the same chunk-maximum records and
decoded candidates on a 16-antenna stream segment (CPU only)."""
import numpy as np
import torch

from dsx import config as dcfg
from dsx import synth
from oracle import corr as ocorr
from oracle import image as oimage
from oracle import search as osearch
from oracle import torch_cpu as tc


def test_correlate_and_images_match_oracle():
    cfg = dcfg.small_node_config().with_(n_samp=256)
    sc = synth.make_scenario(cfg)
    raw = sc.block(0, cfg.n_samp)
    want = ocorr.correlate(raw, n_pkt=cfg.n_pkt, n_ant=cfg.n_ant, n_chan=cfg.n_chan, t_int=cfg.t_int, csf=cfg.csf)
    got = tc.correlate(raw, n_pkt=cfg.n_pkt, n_ant=cfg.n_ant, n_chan=cfg.n_chan, t_int=cfg.t_int, csf=cfg.csf)
    assert np.array_equal(got.numpy().astype(np.int64), want)
    rows, cols, _ = oimage.cell_map(sc.antpos_e, sc.antpos_n, cfg.freq_out_GHz, cfg.n_grid)
    mask = oimage.edge_mask(cfg.n_grid, 1)
    wi = oimage.subband_images(want, rows, cols, cfg.n_grid, cfg.n_sub, mask).reshape(cfg.n_sub, cfg.n_steps, -1)
    gi = tc.grid_images(got, rows, cols, cfg.n_grid, cfg.n_sub, mask).numpy()
    assert np.abs(gi - wi).max() <= 1e-5 * np.abs(wi).max()


def test_search_block_matches_oracle():
    rng = np.random.default_rng(4)
    n_sub, n_pix, n_dm = 6, 40, 24
    sh = np.sort(rng.integers(0, 30, size=(n_dm, n_sub)), axis=0).astype(np.int32)
    widths = [1, 2, 4, 8]
    L, hist_len = 128, 128 + 64 + 4 + 31
    img = rng.standard_normal((n_sub, hist_len, n_pix))
    img[:, 150:154, 7] += 3.0
    st_n = np.full(n_dm, 1000.0)
    st_mean = rng.normal(0, 0.1, (n_dm, n_pix))
    st_m2 = np.full((n_dm, n_pix), 1000.0 * n_sub)
    tau0, c_lo = 0, hist_len - L - 4 - 64 + 64
    got = tc.search_block(torch.from_numpy(img.astype(np.float32)), tau0, sh, c_lo, L - 64, widths, st_n,
                          torch.from_numpy(st_mean), torch.from_numpy(st_m2), 0)
    imgf = img.astype(np.float32).astype(np.float64)
    want = osearch.search_block_c(imgf, tau0, sh, c_lo, L - 64, 64, widths, st_n, st_mean, st_m2, 0)
    ok = np.isfinite(want[0])
    assert np.array_equal(ok, np.isfinite(got[0]))
    np.testing.assert_allclose(got[0][ok], want[0][ok], rtol=1e-5)
    assert np.array_equal(got[1][ok], want[1][ok]) and np.array_equal(got[2][ok], want[2][ok])
    np.testing.assert_allclose(got[3], want[3], rtol=1e-5, atol=1e-4)      # float32 sums
    np.testing.assert_allclose(got[4], want[4], rtol=1e-5)
    assert np.array_equal(got[5], want[5])
    cw = osearch.decode(want[0], want[1], want[2], widths, 8, threshold=3.0, n_top=64, radius_t=8,
                        radius_dm=2, radius_lm=1)
    cg = tc.decode(got[0], got[1], got[2], 8, threshold=3.0, n_top=64, radius_t=8, radius_dm=2, radius_lm=1)
    assert [(c.dm_idx, c.t, widths.index(c.width), c.row, c.col) for c in cw] == [x[:5] for x in cg]
    assert len(cg) > 0
