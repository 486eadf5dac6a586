"""Precision of the stored formats against the float64 oracle (config 1
inputs, 3 blocks): FP16 subband images (computed in FP32 from the exact int32
uv grid) and the dedispersed series summed in FP32 from those FP16 images,
each measured on the central 32 x 32 pixel crop for every (subband, step)
image and every (DM trial, pixel) series.  Visibilities are exact.

Bars: image error <= 1e-3 of the crop's peak; dedispersed-series error
<= 2e-3 of the series' peak.  The figures are printed as one JSON line."""
import json

import numpy as np
import pytest
import torch

from dsx import config as dcfg
from dsx import synth
from dsx.dedisp import combine_chgroups
from dsx.imager import SubbandImager
from dsx.pipeline import SearchNode
from oracle import corr as ocorr
from oracle import image as oimage
from oracle import search as osearch

pytestmark = pytest.mark.gpu


def test_precision_report_16ant():
    cfg = dcfg.small_node_config()
    sc = synth.make_scenario(cfg, synth.small_node_bursts(cfg))
    n_blk = 3
    raw = np.concatenate([sc.block(b * cfg.n_samp, cfg.n_samp) for b in range(n_blk)])
    n_samp = n_blk * cfg.n_samp
    vis = ocorr.correlate(raw, n_pkt=n_samp // 2, n_ant=cfg.n_ant, n_chan=cfg.n_chan, t_int=cfg.t_int,
                          csf=cfg.csf)
    node = SearchNode(cfg, sc.antpos_e, sc.antpos_n, max_block_steps=n_samp // cfg.t_int)
    vis_gpu = node.correlate(torch.from_numpy(raw).cuda(), n_samp).cpu().numpy()
    vis_err = int(np.abs(vis_gpu.astype(np.int64) - vis).max())
    assert vis_err == 0
    rows, cols, _ = oimage.cell_map(sc.antpos_e, sc.antpos_n, cfg.grid_freq_out_GHz, cfg.n_grid)
    ref = oimage.subband_images(vis, rows, cols, cfg.n_grid, cfg.n_sub, oimage.edge_mask(cfg.n_grid))
    img = SubbandImager(sc.cmap, cfg.n_sub, out_dtype=torch.float16)(
        torch.from_numpy(vis.astype(np.int32)).cuda())
    N = cfg.n_grid
    lo = N // 2 - 16
    crop = (slice(None), slice(None), slice(lo, lo + 32), slice(lo, lo + 32))
    g = img.float().cpu().numpy()[crop]
    r = ref[crop]
    img_err = np.abs(g - r).max(axis=(2, 3)) / np.maximum(np.abs(r).max(axis=(2, 3)), 1e-30)
    # dedispersed series of every DM trial over the crop, FP32 sums of the FP16 images
    sh = cfg.shift_table()
    n_steps = g.shape[1]
    gpu_streams = {s: torch.from_numpy(np.ascontiguousarray(g[s].reshape(n_steps, -1))).cuda()
                   for s in range(cfg.n_sub)}
    t_lo, t_hi = 0, n_steps                      # zero before the first step in both paths
    ref_D = osearch.dedisperse_np(r.reshape(cfg.n_sub, n_steps, -1), 0, sh, t_lo, t_hi)   # (dm, t, pix)
    ser_err = np.zeros((cfg.n_dm, 1024))
    for d in range(cfg.n_dm):
        D = combine_chgroups(per_chgroup_streams={s: v.reshape(n_steps, 32, 32) for s, v in gpu_streams.items()},
                             time_shift_per_chgroup=sh[d].astype(np.int32), t_window=(t_lo, t_hi),
                             n_chgroup=cfg.n_sub).cpu().numpy().reshape(t_hi - t_lo, -1)
        ser_err[d] = np.abs(D - ref_D[d]).max(axis=0) / np.maximum(np.abs(ref_D[d]).max(axis=0), 1e-30)
    report = {"config": "16-antenna 64-channel inputs, 3 blocks (768 steps), 64 subbands, 64 DMs, crop 32x32",
              "visibilities_max_abs_err": vis_err,
              "image_fp16_rel_err": {"max": float(img_err.max()), "median": float(np.median(img_err)),
                                     "n_images": int(img_err.size)},
              "dedispersed_fp32_from_fp16_rel_err": {"max": float(ser_err.max()),
                                                      "median": float(np.median(ser_err)),
                                                      "n_series": int(ser_err.size)}}
    print(json.dumps(report), flush=True)
    assert img_err.max() <= 1e-3
    assert ser_err.max() <= 2e-3
