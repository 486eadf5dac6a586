"""Calibrated correlator (cal + RFI mask on operand load, fp16 MFMA, float32
accumulation) against the oracle's restatement of the reference's
``gemm2x_reference`` operand arithmetic (src/dsart/x2/corr/gemm2x.py:346-390).

Bars: fp16 operands (``apply_cal_split`` output) bit-exact; visibilities
within 1e-5 of the float64 sum relative to the per-output bound
sum |x_b| |x_a| (float32 accumulation of at most a few hundred exact
products); unit gains on integer codes bit-exact with the int8 correlator;
the fused raw-byte path bit-identical to apply_cal_split + compute_split."""
import numpy as np
import pytest
import torch

from dsx import config as dcfg
from dsx import synth
from dsx.corr import (FastCorrKernel, apply_cal_split, compact_cal, make_cal_broadcast_tensors,
                      unpack_int4_split)
from oracle import corr as ocorr

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _gains(rng, n_ant, n_chan, masked=0.05):
    g = (rng.normal(1.0, 0.3, (n_ant, n_chan, 2)) + 1j * rng.normal(0.0, 0.3, (n_ant, n_chan, 2)))
    mask = rng.random((n_ant, n_chan, 2)) < masked                  # (ANT, CH, POL) flagged
    return g, mask


def _oracle(raw, n_pkt, n_ant, n_chan, scale, cal_re=None, cal_im=None, keep=None, t_int=8, csf=1,
            fuse=True):
    xr, xi = ocorr.cal_operands(raw, n_pkt=n_pkt, n_ant=n_ant, n_chan=n_chan, scale=scale,
                                cal_re=cal_re, cal_im=cal_im, time_keep=keep)
    v, bound = ocorr.correlate_float(xr, xi, t_int=t_int, csf=csf, fuse_pol=fuse)
    order = (0, 2, 1) if fuse else (0, 3, 1, 2)                    # -> (fv, nbase, nch[, pol])
    return np.transpose(v, order), np.transpose(bound, order), (xr, xi)


def _check(got, want, bound):
    err = np.abs(got - want)
    lim = TOL * bound + 1e-30
    worst = float((err / np.maximum(bound, 1e-30)).max())
    assert np.all(err <= lim), f"max error / bound = {worst:.3e}"
    return worst


@pytest.mark.parametrize("fuse,csf", [(True, 1), (False, 1), (True, 4)])
def test_compute_raw_cal_vs_oracle_16ant(fuse, csf):
    cfg = dcfg.small_node_config().with_(n_samp=256)
    sc = synth.make_scenario(cfg)
    raw = sc.block(0, 256)                                         # (128, 16, 64, 2, 2)
    rng = np.random.default_rng(21)
    g, mask = _gains(rng, 16, 64)
    cr16, ci16 = make_cal_broadcast_tensors(g, device="cuda", dtype=torch.float16)
    cr, ci = compact_cal(cr16, ci16, torch.from_numpy(mask).cuda())
    keep = np.ones((4, 64, 2), np.float32)
    keep[1, 3:9, 0] = 0.0                                          # an array-burst time mask
    k = FastCorrKernel(device="cuda", t_int_fast_native=8, nants=16, nchan=64)
    got = k.compute_raw(torch.from_numpy(raw).cuda(), fuse_stokes_i=fuse, chan_sum_factor=csf,
                        scale=0.05, cal=(cr, ci), time_keep=torch.from_numpy(keep).cuda())
    want, bound, _ = _oracle(raw, 128, 16, 64, 0.05, cr.cpu().numpy(), ci.cpu().numpy(), keep,
                             csf=csf, fuse=fuse)
    worst = _check(got.cpu().numpy(), want, bound)
    print(f"fuse={fuse} csf={csf}: max |error| / bound = {worst:.2e}", flush=True)


def test_compute_raw_cal_vs_oracle_96ant_csf24():
    """The node's shape: 96 antennas x 384 channels summed 24 to a subband
    (two 12-channel LDS chunks per work item)."""
    cfg = dcfg.search_node_short_config(64)
    sc = synth.make_scenario(cfg)
    raw = sc.block(0, 64)                                          # (32, 96, 384, 2, 2)
    rng = np.random.default_rng(22)
    g, mask = _gains(rng, 96, 384)
    cr16, ci16 = make_cal_broadcast_tensors(g, device="cuda", dtype=torch.float16)
    cr, ci = compact_cal(cr16, ci16, torch.from_numpy(mask).cuda())
    k = FastCorrKernel(device="cuda", t_int_fast_native=8, nants=96, nchan=384)
    got = k.compute_raw(torch.from_numpy(raw).cuda(), fuse_stokes_i=True, chan_sum_factor=24,
                        scale=0.05, cal=(cr, ci))
    want, bound, _ = _oracle(raw, 32, 96, 384, 0.05, cr.cpu().numpy(), ci.cpu().numpy(), csf=24)
    worst = _check(got.cpu().numpy(), want, bound)
    print(f"96 ant csf 24: max |error| / bound = {worst:.2e}", flush=True)


def test_unit_gains_on_codes_equal_int8_correlator():
    """Unit gains, scale 1: fp16 operands are the integer codes and every
    float32 partial sum is an exact integer -> bit-exact with the int8 path."""
    cfg = dcfg.search_node_short_config(64)
    sc = synth.make_scenario(cfg)
    raw_t = torch.from_numpy(sc.block(0, 64)).cuda()
    ones = torch.ones((384, 2, 96), device="cuda")
    zeros = torch.zeros_like(ones)
    k = FastCorrKernel(device="cuda", t_int_fast_native=8, nants=96, nchan=384, accum_dtype=torch.int32)
    exact = k.compute_raw(raw_t, fuse_stokes_i=True, chan_sum_factor=24).cpu().numpy()
    cal = k.compute_raw(raw_t, fuse_stokes_i=True, chan_sum_factor=24, cal=(ones, zeros)).cpu().numpy()
    assert np.array_equal(cal.real, exact[..., 0].astype(np.float32))
    assert np.array_equal(cal.imag, exact[..., 1].astype(np.float32))


def test_apply_cal_split_then_compute_split_equals_fused():
    cfg = dcfg.small_node_config().with_(n_samp=256)
    sc = synth.make_scenario(cfg)
    raw = sc.block(0, 256)
    rng = np.random.default_rng(23)
    g, mask = _gains(rng, 16, 64)
    cr16, ci16 = make_cal_broadcast_tensors(np.where(mask, 0, g), device="cuda", dtype=torch.float16)
    re, im = unpack_int4_split(raw, scale=0.05, out_dtype=torch.float16, shape=(128, 16, 64))
    cre, cim = apply_cal_split(re, im, cr16, ci16)
    # fp16 operands bit-exact with the oracle's
    _, _, (xr, xi) = _oracle(raw, 128, 16, 64, 0.05, cr16[:, 0, :, 0, :].float().cpu().numpy(),
                             ci16[:, 0, :, 0, :].float().cpu().numpy())
    # oracle (S, ant, chan, pol) -> GEMM layout (chan, t, pol, pkt, ant)
    S = xr.shape[0]
    to_gemm = lambda x: np.transpose(x.reshape(S // 2, 2, 16, 64, 2), (3, 1, 4, 0, 2))
    assert np.array_equal(cre.float().cpu().numpy(), to_gemm(xr))
    assert np.array_equal(cim.float().cpu().numpy(), to_gemm(xi))
    k = FastCorrKernel(device="cuda", t_int_fast_native=8, nants=16, nchan=64)
    split = k.compute_split(cre, cim, fuse_stokes_i=True)
    fused = k.compute_raw(torch.from_numpy(raw).cuda(), fuse_stokes_i=True, scale=0.05, cal=(cr16, ci16))
    assert torch.equal(split, fused)


def test_float_planes_use_fp16_operands():
    """unpack_int4_split's float output through compute_split: the reference's
    fp16 operands (fp16(0.05 k)), float32 accumulation."""
    cfg = dcfg.small_node_config().with_(n_samp=256)
    sc = synth.make_scenario(cfg)
    raw = sc.block(0, 256)
    k = FastCorrKernel(device="cuda", t_int_fast_native=8, nants=16, nchan=64)
    for dt in (torch.float16, torch.float32):
        re, im = unpack_int4_split(raw, scale=0.05, out_dtype=dt, shape=(128, 16, 64))
        got = k.compute_split(re, im).cpu().numpy()                   # (fv, nbase, nch, 2)
        want, bound, _ = _oracle(raw, 128, 16, 64, 0.05, fuse=False)
        _check(got, want, bound)


def test_cal_argument_errors():
    k = FastCorrKernel(device="cuda", t_int_fast_native=8, nants=16, nchan=64)
    raw = torch.zeros((8, 16, 64, 2, 2), dtype=torch.uint8, device="cuda")
    ones = torch.ones((64, 2, 16), device="cuda")
    with pytest.raises(ValueError):
        k.compute_raw(raw, cal=(ones[:, :, :8], ones[:, :, :8]))
    with pytest.raises(ValueError):
        k.compute_raw(raw, cal=(ones, ones), time_keep=torch.ones((3, 64, 2), device="cuda"))
    with pytest.raises(ValueError):
        k.compute_raw(raw, time_keep=torch.ones((4, 64, 2), device="cuda"))
    with pytest.raises(TypeError):
        k.compute_raw(raw, cal=ones)


def test_search_node_calibrated_images_vs_oracle():
    """The node stream with gains on the correlator's operand load: float32
    visibilities, float32 LDS-atomic gridding, C2R imaging -- images within
    1e-5 of the float64 oracle relative to the image peak; the stream then
    runs detection on them."""
    from dsx.pipeline import SearchNode
    from oracle import image as oimage
    from oracle.stream import OracleStream
    cfg = dcfg.small_node_config().with_(n_samp=512)
    sc = synth.make_scenario(cfg, synth.small_node_bursts(cfg))
    raw = sc.block(0, cfg.n_samp)
    rng = np.random.default_rng(24)
    g, mask = _gains(rng, cfg.n_ant, cfg.n_chan)
    keep = np.ones((2, cfg.n_chan, 2), np.float32)
    keep[1, 10:20, 1] = 0.0
    cr16, ci16 = make_cal_broadcast_tensors(g, device="cuda", dtype=torch.float16)
    node = SearchNode(cfg, sc.antpos_e, sc.antpos_n)
    node.set_calibration(cr16, ci16, cube_mask=torch.from_numpy(mask).cuda(), time_keep=keep, scale=0.05)
    vis = node.correlate(torch.from_numpy(raw).cuda(), cfg.n_samp)
    img = node.image(vis, cfg.n_steps)
    torch.cuda.synchronize()
    cr, ci = compact_cal(cr16, ci16, torch.from_numpy(mask).cuda())
    xr, xi = ocorr.cal_operands(raw, n_pkt=cfg.n_pkt, n_ant=cfg.n_ant, n_chan=cfg.n_chan, scale=0.05,
                                cal_re=cr.cpu().numpy(), cal_im=ci.cpu().numpy(), time_keep=keep)
    v, bound = ocorr.correlate_float(xr, xi, t_int=cfg.t_int, csf=cfg.csf)     # (steps, nco, nbase)
    got_v = torch.view_as_complex(vis.view(cfg.n_steps, cfg.n_chan_out, cfg.nbase, 2)).cpu().numpy()
    assert np.all(np.abs(got_v - v) <= TOL * bound + 1e-30)
    o = OracleStream(cfg, sc.antpos_e, sc.antpos_n)
    want = oimage.subband_images(np.stack([v.real, v.imag], axis=-1), o.rows, o.cols, cfg.n_grid,
                                 cfg.n_sub, o.mask)
    got = img.float().cpu().numpy()
    err = float(np.abs(got - want).max() / np.abs(want).max())
    print(f"calibrated images: max error / peak = {err:.2e}", flush=True)
    assert err <= 1e-5
    node.push_images(img, cfg.n_steps)
    res = node.detect(cfg.n_steps)
    assert res.T0 == 0 and not res.overflow
    node.set_calibration()                       # back to the exact integer path
    assert node.correlate(torch.from_numpy(raw).cuda(), cfg.n_samp).dtype == torch.int32
