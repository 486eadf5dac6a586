"""GPU parity of each stage against the CPU oracle (int64 / float64).

Bars: bit-exact for the synthetic bytes and the integer visibilities;
images within 1e-5 (float32) / 1e-3 (float16) of the float64 oracle,
relative to the image's peak magnitude.
"""
import numpy as np
import pytest
import torch

from dsx import config as dcfg
from dsx import synth
from dsx.imager import SubbandImager
from dsx.pattern import build_cell_map
from oracle import corr as ocorr
from oracle import image as oimage

pytestmark = pytest.mark.gpu


def _scenario_small():
    cfg = dcfg.small_node_config()
    return synth.make_scenario(cfg, synth.small_node_bursts(cfg))


def test_synth_gpu_matches_numpy_twin():
    sc = _scenario_small()
    cfg = sc.cfg
    gs = synth.GpuSynth(sc)
    for blk in (0, 5):          # block 5 carries the burst
        s0 = blk * cfg.n_samp
        got = gs.block(s0, cfg.n_samp).cpu().numpy()
        want = sc.block(s0, cfg.n_samp)
        assert got.shape == want.shape
        assert np.array_equal(got, want), f"block {blk}: {np.count_nonzero(got != want)} bytes differ"


def test_synth_gpu_96ant_slice_matches_numpy():
    cfg = dcfg.search_node_short_config(64)
    sc = synth.make_scenario(cfg, synth.random_bursts(cfg, 2, 7, t_lo_step=0, t_hi_step=8))
    gs = synth.GpuSynth(sc)
    got = gs.block(128, 64).cpu().numpy()
    assert np.array_equal(got, sc.block(128, 64))


def _corr_gpu(raw_np, cfg, *, fuse_pol=True, out_layout=1, n_samp=None, split=False):
    from dsx import _lib
    n_samp = n_samp or raw_np.shape[0] * 2
    n_pkt = n_samp // 2
    raw = torch.from_numpy(raw_np).cuda()
    npol = 1 if fuse_pol else 2
    n_steps = n_samp // cfg.t_int
    if out_layout == 1:
        vis = torch.zeros((n_steps, cfg.n_chan_out, npol, cfg.nbase, 2), dtype=torch.int32, device="cuda")
    else:
        vis = torch.zeros((n_steps, cfg.nbase, cfg.n_chan_out, npol, 2), dtype=torch.int32, device="cuda")
    if split:
        re = torch.empty((cfg.n_chan, 2, 2, n_pkt, cfg.n_ant), dtype=torch.int8, device="cuda")
        im = torch.empty_like(re)
        _lib.call("dsx_unpack_i4", _lib.ptr(raw), _lib.ptr(re), _lib.ptr(im), n_pkt, cfg.n_ant,
                  cfg.n_chan, _lib.DSX_I8, 1.0, _lib.stream())
        _lib.call("dsx_corr_i8", _lib.ptr(re), _lib.ptr(im), _lib.ptr(vis), n_pkt, cfg.n_ant, cfg.n_chan,
                  cfg.t_int, cfg.csf, int(fuse_pol), 1, out_layout, _lib.stream())
    else:
        _lib.call("dsx_corr_i8", _lib.ptr(raw), None, _lib.ptr(vis), n_pkt, cfg.n_ant, cfg.n_chan,
                  cfg.t_int, cfg.csf, int(fuse_pol), 0, out_layout, _lib.stream())
    torch.cuda.synchronize()
    return vis.cpu().numpy()


def test_corr_16ant_bit_exact():
    sc = _scenario_small()
    cfg = sc.cfg
    raw = sc.block(5 * cfg.n_samp, cfg.n_samp)
    got = _corr_gpu(raw, cfg)[:, :, 0]                       # (steps, cout, nbase, 2)
    want = ocorr.correlate(raw, n_pkt=cfg.n_pkt, n_ant=cfg.n_ant, n_chan=cfg.n_chan, t_int=cfg.t_int,
                           csf=cfg.csf)
    assert np.array_equal(got.astype(np.int64), want)


@pytest.mark.parametrize("fuse_pol,out_layout,split", [(False, 1, False), (True, 0, False),
                                                        (False, 0, True), (True, 1, True)])
def test_corr_layouts_bit_exact(fuse_pol, out_layout, split):
    cfg = dcfg.small_node_config().with_(csf=4, n_sub=16, n_samp=512)
    sc = synth.make_scenario(cfg)
    raw = sc.block(0, 512)
    got = _corr_gpu(raw, cfg, fuse_pol=fuse_pol, out_layout=out_layout, split=split)
    want = ocorr.correlate(raw, n_pkt=256, n_ant=cfg.n_ant, n_chan=cfg.n_chan, t_int=cfg.t_int,
                           csf=cfg.csf, fuse_pol=fuse_pol)
    if fuse_pol:
        want = want[:, :, None]                                # (steps, cout, 1, nbase, 2)
    if out_layout == 0:
        got = np.transpose(got, (0, 2, 3, 1, 4))               # -> (steps, cout, pol, nbase, 2)
    assert np.array_equal(got.astype(np.int64), want)


def test_corr_96ant_shape_bit_exact():
    """96 antennas x 384 channels, csf 24 (16 subbands), 8 steps."""
    cfg = dcfg.search_node_short_config(64)
    sc = synth.make_scenario(cfg)
    raw = sc.block(0, 64)
    got = _corr_gpu(raw, cfg)[:, :, 0]
    want = ocorr.correlate(raw, n_pkt=32, n_ant=96, n_chan=384, t_int=8, csf=24)
    assert np.array_equal(got.astype(np.int64), want)


def test_corr_sign_known_answer():
    """E_1 = (1, 0), E_2 = (0, 1) -> V = conj(E_1) E_2 = +i per sample and pol."""
    cfg = dcfg.small_node_config().with_(n_samp=64)
    re = np.zeros((64, cfg.n_ant, cfg.n_chan, 2), dtype=np.int64)
    im = np.zeros_like(re)
    re[:, 3] = 1
    im[:, 5] = 1
    byte = ((re & 0xF) | ((im & 0xF) << 4)).astype(np.uint8)
    raw = np.ascontiguousarray(byte.reshape(32, 2, cfg.n_ant, cfg.n_chan, 2).transpose(0, 2, 3, 1, 4))
    got = _corr_gpu(raw, cfg)[:, :, 0]
    bls = 5 * 6 // 2 + 3
    assert np.all(got[:, :, bls, 0] == 0)
    assert np.all(got[:, :, bls, 1] == cfg.t_int * 2)


@pytest.mark.parametrize("dtype,tol", [(torch.float32, 1e-5), (torch.float16, 1e-3)])
def test_imager_16ant_vs_float64(dtype, tol):
    sc = _scenario_small()
    cfg = sc.cfg
    raw = sc.block(5 * cfg.n_samp, cfg.n_samp)[: 64]          # 16 steps
    vis = ocorr.correlate(raw, n_pkt=64, n_ant=cfg.n_ant, n_chan=cfg.n_chan, t_int=cfg.t_int, csf=cfg.csf)
    rows, cols, cell = oimage.cell_map(sc.antpos_e, sc.antpos_n, cfg.grid_freq_out_GHz, cfg.n_grid)
    want = oimage.subband_images(vis, rows, cols, cfg.n_grid, cfg.n_sub, oimage.edge_mask(cfg.n_grid))
    imager = SubbandImager(sc.cmap, cfg.n_sub, out_dtype=dtype)
    got = imager(torch.from_numpy(vis.astype(np.int32)).cuda()).float().cpu().numpy()
    err = np.abs(got - want).max() / np.abs(want).max()
    assert err <= tol, f"max err {err:.3e} > {tol}"


def test_imager_96ant_vs_float64():
    cfg = dcfg.search_node_short_config(64)
    sc = synth.make_scenario(cfg)
    raw = sc.block(0, 64)
    vis = ocorr.correlate(raw, n_pkt=32, n_ant=96, n_chan=384, t_int=8, csf=24)
    rows, cols, cell = oimage.cell_map(sc.antpos_e, sc.antpos_n, cfg.grid_freq_out_GHz, cfg.n_grid)
    assert np.array_equal(rows, sc.cmap.row) and np.array_equal(cols, sc.cmap.col)
    want = oimage.subband_images(vis, rows, cols, cfg.n_grid, cfg.n_sub, oimage.edge_mask(cfg.n_grid))
    for dtype, tol in ((torch.float32, 1e-5), (torch.float16, 1e-3)):
        imager = SubbandImager(sc.cmap, cfg.n_sub, out_dtype=dtype)
        got = imager(torch.from_numpy(vis.astype(np.int32)).cuda()).float().cpu().numpy()
        err = np.abs(got - want).max() / np.abs(want).max()
        assert err <= tol, f"{dtype}: max err {err:.3e} > {tol}"
