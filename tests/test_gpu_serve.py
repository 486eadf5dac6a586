"""Serving: uneven blocks from two node streams through :class:`NodeServer`
(overlapping front/back stages on two HIP streams, HBM-budget admission)
give every block the same result as the stream processed alone.

Bar: per block, the chunk-maximum records (sorted) and the decoded candidate
lists are identical (same kernels on the same data and history), and the
float64 noise state after the last block is identical."""
import numpy as np
import pytest
import torch

from dsx import config as dcfg
from dsx import synth
from dsx.pipeline import SearchNode
from dsx.serve import BlockRequest, NodeServer, poisson_arrivals

pytestmark = pytest.mark.gpu

LENGTHS = (1024, 2048, 4096)


def _stream(seed_volt, seed_len, n_blocks, bursts):
    cfg = dcfg.small_node_config().with_(volt_seed=seed_volt)
    sc = synth.make_scenario(cfg, bursts(cfg) if bursts else ())
    rng = np.random.default_rng(seed_len)
    lens = [int(x) for x in rng.choice(LENGTHS, size=n_blocks)]
    starts = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
    return cfg, sc, lens, starts


def _sorted(rec):
    return np.sort(rec, order=["t", "p", "d", "k"]) if rec.size else rec


def _keys(cands):
    return [(c.dm_idx, c.time_bin, c.width_samples, int(c.l), int(c.m), c.snr) for c in cands]


def test_alone_equals_in_the_mix():
    cap = max(LENGTHS) // 8
    streams = [_stream(0xD5A1107F, 11, 10, synth.small_node_bursts), _stream(0x5EED0002, 12, 8, None)]
    # alone: each stream on its own node, one block at a time
    alone = []
    for cfg, sc, lens, starts in streams:
        node = SearchNode(cfg, sc.antpos_e, sc.antpos_n, max_block_steps=cap)
        gs = synth.GpuSynth(sc)
        res = [node.process_block(gs.block(int(s0), n), n) for s0, n in zip(starts, lens)]
        alone.append((res, node.mean.cpu().numpy(), node.m2.cpu().numpy()))
        del node
    torch.cuda.empty_cache()
    # in the mix: both streams interleaved, bursty arrivals, overlapping stages
    nodes, reqs = [], []
    for sid, (cfg, sc, lens, starts) in enumerate(streams):
        nodes.append(SearchNode(cfg, sc.antpos_e, sc.antpos_n, max_block_steps=cap, n_slots=2))
        gs = synth.GpuSynth(sc)
        arr = poisson_arrivals([n * cfg.t_step_us / cfg.t_int * 1e-6 for n in lens], 50.0, 7 + sid)
        for k, (s0, n) in enumerate(zip(starts, lens)):
            reqs.append(BlockRequest(stream_id=sid, seq=k, n_samp=n, t_arrival=float(arr[k]),
                                     make=lambda st, g=gs, a=int(s0), b=n: g.block(a, b)))
    server = NodeServer(nodes, hbm_budget_bytes=64 << 30)
    done = server.serve(reqs, realtime=True)
    assert len(done) == len(reqs)
    by = {(d.stream_id, d.seq): d.result for d in done}
    n_cands = 0
    for sid, (res_alone, mean_a, m2_a) in enumerate(alone):
        for k, ra in enumerate(res_alone):
            rm = by[(sid, k)]
            assert rm.T0 == ra.T0 and rm.n_steps == ra.n_steps
            assert np.array_equal(_sorted(rm.records), _sorted(ra.records)), (sid, k)
            assert _keys(rm.candidates) == _keys(ra.candidates), (sid, k)
            n_cands += len(ra.candidates)
        assert np.array_equal(nodes[sid].mean.cpu().numpy(), mean_a)
        assert np.array_equal(nodes[sid].m2.cpu().numpy(), m2_a)
    assert n_cands > 0
