"""Serving: uneven blocks from two node streams through :class:`NodeServer`
(overlapping front/back stages on two HIP streams, HBM-budget admission)
give every block the same result as the stream processed alone, and that
result is the oracle's.

Bars: per block, the chunk-maximum records (sorted) and the decoded
candidate lists are identical alone and in the mix (same kernels on the same
data and history), and the float64 noise state after the last block is
identical; every block of the burst stream decodes the oracle's candidate
keys (DM index, time bin, width, row, col) with S/N within 2e-3; with an HBM
budget that holds one block in flight, admission waits for memory and the
results do not change."""
import numpy as np
import pytest
import torch

from dsx import config as dcfg
from dsx import synth
from dsx.pipeline import SearchNode
from dsx.serve import BlockRequest, NodeServer, poisson_arrivals
from oracle.stream import OracleStream

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


def _alone(streams, cap):
    out = []
    for cfg, sc, lens, starts in streams:
        node = SearchNode(cfg, sc.antpos_e, sc.antpos_n, max_block_steps=cap)
        gs = synth.GpuSynth(sc)
        res = [node.process_block(gs.block(int(s0), n), n) for s0, n in zip(starts, lens)]
        out.append((res, node.mean.cpu().numpy(), node.m2.cpu().numpy()))
        del node
    torch.cuda.empty_cache()
    return out


def _served(streams, cap, rate, budget):
    nodes, reqs = [], []
    for sid, (cfg, sc, lens, starts) in enumerate(streams):
        nodes.append(SearchNode(cfg, sc.antpos_e, sc.antpos_n, max_block_steps=cap, n_slots=2))
        gs = synth.GpuSynth(sc)
        arr = poisson_arrivals([n * cfg.t_step_us / cfg.t_int * 1e-6 for n in lens], rate, 7 + sid)
        for k, (s0, n) in enumerate(zip(starts, lens)):
            reqs.append(BlockRequest(stream_id=sid, seq=k, n_samp=n, t_arrival=float(arr[k]),
                                     make=lambda st, g=gs, a=int(s0), b=n: g.block(a, b)))
    if budget is None:
        budget = 64 << 30
    elif budget < 0:          # room for one block in flight next to the resident state
        budget = sum(nd.device_bytes() for nd in nodes) + NodeServer.block_bytes(nodes[0], max(LENGTHS)) + 1
    server = NodeServer(nodes, hbm_budget_bytes=budget)
    done = server.serve(reqs, realtime=True)
    assert len(done) == len(reqs)
    return server, nodes, {(d.stream_id, d.seq): d.result for d in done}


def _same(alone, nodes, by):
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
    return n_cands


def test_alone_equals_in_the_mix_and_the_oracle():
    cap = max(LENGTHS) // 8
    streams = [_stream(0xD5A1107F, 11, 10, synth.small_node_bursts), _stream(0x5EED0002, 12, 8, None)]
    alone = _alone(streams, cap)
    # the burst stream, block by block, against the oracle (uneven block lengths)
    cfg, sc, lens, starts = streams[0]
    o = OracleStream(cfg, sc.antpos_e, sc.antpos_n)
    for k, (s0, n) in enumerate(zip(starts, lens)):
        want = o.push_raw(sc.block(int(s0), n), n)
        got = alone[0][0][k].candidates
        ok = [(c.dm_idx, c.t, c.width, c.row, c.col) for c in want]
        assert [c.key for c in got] == ok, f"block {k} ({n} samples): gpu {[c.key for c in got]} oracle {ok}"
        for gc, oc in zip(got, want):
            assert abs(gc.snr - oc.snr) <= 2e-3 * abs(oc.snr), (k, gc, oc)
    print(f"oracle: {sum(len(r.candidates) for r in alone[0][0])} candidates over {len(lens)} blocks", flush=True)
    # in the mix: both streams interleaved, bursty arrivals, overlapping stages
    _, nodes, by = _served(streams, cap, 50.0, None)
    assert _same(alone, nodes, by) > 0


def test_admission_waits_for_memory():
    cap = max(LENGTHS) // 8
    streams = [_stream(0xD5A1107F, 21, 6, synth.small_node_bursts)]
    alone = _alone(streams, cap)
    server, nodes, by = _served(streams, cap, 200.0, -1)
    print(f"admissions that waited for memory: {server.budget_waits}", flush=True)
    assert server.budget_waits > 0
    _same(alone, nodes, by)
