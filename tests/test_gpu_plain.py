"""The plain FP32 form (library GEMM correlator, index_add_ gridder, rocFFT
imager, one dedispersion pass per DM trial) against the float64 oracle and
the fused form on the 16-antenna stream (16 ant x 64 ch, 7 blocks, DM-250
burst).

Bars: visibilities bit-exact (integer-valued complex64); decoded candidate
keys (DM index, time bin, width, row, col) identical to the oracle's with S/N
within 2e-3 relative."""
import numpy as np
import pytest
import torch

from dsx import config as dcfg
from dsx import synth
from dsx.plain import PlainNode
from oracle import corr as ocorr
from oracle.stream import OracleStream

pytestmark = pytest.mark.gpu


def test_plain_correlator_exact():
    cfg = dcfg.small_node_config().with_(n_samp=512)
    sc = synth.make_scenario(cfg)
    raw = sc.block(0, 256)
    node = PlainNode(cfg, sc.antpos_e, sc.antpos_n)
    v = node.correlate(torch.from_numpy(raw).cuda(), 256).cpu().numpy()     # (steps, ch, bls)
    want = ocorr.correlate(raw, n_pkt=128, n_ant=cfg.n_ant, n_chan=cfg.n_chan, t_int=cfg.t_int,
                           csf=cfg.csf, fuse_pol=True)                       # (steps, ch, bls, 2)
    assert np.array_equal(v.real.astype(np.int64), want[..., 0])
    assert np.array_equal(v.imag.astype(np.int64), want[..., 1])


def test_plain_16ant_stream_candidates_match_oracle():
    cfg = dcfg.small_node_config()
    sc = synth.make_scenario(cfg, synth.small_node_bursts(cfg))
    gs = synth.GpuSynth(sc)
    node = PlainNode(cfg, sc.antpos_e, sc.antpos_n)
    o = OracleStream(cfg, sc.antpos_e, sc.antpos_n)
    found = []
    for blk in range(7):
        print(f"block {blk}", flush=True)
        s0 = blk * cfg.n_samp
        res = node.process_block(gs.block(s0, cfg.n_samp), cfg.n_samp)
        want = o.push_raw(sc.block(s0, cfg.n_samp), cfg.n_samp)
        gk = [(c.dm_idx, c.time_bin, c.width_samples, int(c.l), int(c.m)) for c in res.candidates]
        ok = [(c.dm_idx, c.t, c.width, c.row, c.col) for c in want]
        assert gk == ok, f"block {blk}: plain {gk} oracle {ok}"
        for gc, oc in zip(res.candidates, want):
            assert abs(gc.snr - oc.snr) <= 2e-3 * abs(oc.snr), (gc, oc)
        found += res.candidates
    b = sc.bursts[0]
    assert any(abs(c.l - b.row) <= 1 and abs(c.m - b.col) <= 1 and c.snr > 15 for c in found), found
