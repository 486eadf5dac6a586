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
from oracle.stream import OracleStream

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


def _chunk_snr(rec, c_lo):
    """GPU records -> {(DM, chunk, pixel): snr}."""
    return {(int(r["d"]), (int(r["t"]) - c_lo) // 64, int(r["p"])): float(r["snr"]) for r in rec}


def _oracle_chunk_snr(o, pix_map):
    L = o.last
    d, c, p = np.nonzero(np.isfinite(L["best_snr"]) & (L["best_snr"] > -1e29))
    return {(int(dd), int(cc), int(pix_map[pp])): float(L["best_snr"][dd, cc, pp]) for dd, cc, pp in zip(d, c, p)}


def _snr_errors(got, want):
    keys = sorted(set(got) & set(want))
    g = np.array([got[k] for k in keys])
    w = np.array([want[k] for k in keys])
    big = np.abs(w) >= 1.0
    return len(keys), float(np.abs(g - w).max()), float((np.abs(g - w)[big] / np.abs(w[big])).max())


@pytest.mark.parametrize("ring_dtype,mode", [(torch.float32, 3), (torch.float16, 2)])
def test_precision_through_fused_kernel_16ant(ring_dtype, mode):
    """Config-1 inputs through the whole fused path (correlator, imager into
    the history dtype, dedispersion + boxcars + float64 statistics): every
    chunk maximum of every DM over the central 32 x 32 crop (threshold below
    any S/N) against the float64 oracle.  Bars on S/N in units of sigma:
    1e-4 with float32 images, 5e-3 with FP16 images."""
    cfg = dcfg.small_node_config().with_(threshold=-1e30, min_stat_count=256)
    sc = synth.make_scenario(cfg, synth.small_node_bursts(cfg))
    N = cfg.n_grid
    lo = N // 2 - 16
    crop = np.array([(lo + r) * N + lo + c for r in range(32) for c in range(32)])
    node = SearchNode(cfg, sc.antpos_e, sc.antpos_n, ring_dtype=ring_dtype, dedisp_mode=mode, cand_cap=1 << 22)
    o = OracleStream(cfg, sc.antpos_e, sc.antpos_n, pix=crop)
    keep = set(int(x) for x in crop)
    n_cmp, max_abs, max_rel = 0, 0.0, 0.0
    for blk in range(4):
        raw = sc.block(blk * cfg.n_samp, cfg.n_samp)
        res = node.process_block(torch.from_numpy(raw).cuda())
        assert not res.overflow
        o.push_raw(raw, cfg.n_samp)
        c_lo = res.T0 - 64
        got = _chunk_snr(res.records[np.isin(res.records["p"], crop)], c_lo)
        want = _oracle_chunk_snr(o, crop)
        if blk >= 2:
            assert set(got) == set(want), f"block {blk}: scored chunks differ"
        if want:
            n, a, r = _snr_errors(got, want)
            n_cmp, max_abs, max_rel = n_cmp + n, max(max_abs, a), max(max_rel, r)
    print(json.dumps({"history": str(ring_dtype), "plan_mode": mode, "chunk_maxima_compared": n_cmp,
                      "snr_max_abs_err_sigma": max_abs, "snr_max_rel_err_above_1sigma": max_rel}), flush=True)
    assert n_cmp > 0
    assert max_abs <= (1e-4 if ring_dtype == torch.float32 else 5e-3)


def test_precision_bf16_pair_rows_node_plan():
    """BF16 dedispersion partial sums: the 16-subband, 256-DM node plan with
    its subband-pair rows stored as BF16 (plan mode 5) against the same plan
    with float32 rows (mode 1) and the float64 oracle, FP16 images (the
    stored-precision variant), 3 blocks of 2048 steps of unit-variance
    images on a 16 x 16 grid.  Every chunk maximum is compared (threshold
    below any S/N).  Bars on S/N in units of sigma: float32 rows 1e-4
    (FP16 images), BF16 rows 5e-2."""
    cfg = dcfg.search_node_config().with_(n_grid=16, threshold=-1e30, min_stat_count=1024)
    e, n = synth.antenna_layout(cfg)
    L, N = cfg.n_steps, cfg.n_grid
    nodes = {m: SearchNode(cfg, e, n, ring_dtype=torch.float16, dedisp_mode=m, cand_cap=1 << 22) for m in (1, 5)}
    assert nodes[5]._plan_for(L)[0].rows_bf16 and not nodes[1]._plan_for(L)[0].rows_bf16
    o = OracleStream(cfg, e, n, pix=np.arange(N * N))
    rng = np.random.default_rng(31)
    errs = {m: (0, 0.0, 0.0) for m in nodes}
    for blk in range(3):
        img = rng.standard_normal((cfg.n_sub, L, N * N)).astype(np.float16).astype(np.float32)
        o.push_images(img.astype(np.float64))
        want = _oracle_chunk_snr(o, np.arange(N * N))
        img_t = torch.from_numpy(img.reshape(cfg.n_sub, L, N, N)).cuda().to(torch.float16)
        for m, nd in nodes.items():
            nd.push_images(img_t, L)
            res = nd.detect(L)
            got = _chunk_snr(res.records, res.T0 - 64)
            if want:
                c, a, r = _snr_errors(got, want)
                errs[m] = (errs[m][0] + c, max(errs[m][1], a), max(errs[m][2], r))
    report = {f"mode{m}_{'bf16' if m == 5 else 'f32'}_rows": {"compared": c, "snr_max_abs_err_sigma": a,
                                                                "snr_max_rel_err_above_1sigma": r}
              for m, (c, a, r) in errs.items()}
    print(json.dumps(report), flush=True)
    assert errs[1][0] > 0 and errs[5][0] > 0
    assert errs[1][1] <= 1e-4 and errs[5][1] <= 5e-2
