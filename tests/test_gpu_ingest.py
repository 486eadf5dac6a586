"""Host ingest: raw pages in pinned host memory, copied on a copy stream
while the previous block computes (double-buffered device slots), give the
same results as blocks already resident in HBM -- records, candidates and
the float64 state bit-identical over a 16-antenna stream of uneven blocks --
and the H2D copy of each block overlaps the previous block's compute."""
import numpy as np
import pytest
import torch

from dsx import config as dcfg
from dsx import synth
from dsx.ingest import HostIngest, PinnedPagePool, page_bytes, process_host_stream
from dsx.pipeline import SearchNode

pytestmark = pytest.mark.gpu


def test_host_pages_equal_device_blocks():
    cfg = dcfg.small_node_config()
    sc = synth.make_scenario(cfg, synth.small_node_bursts(cfg))
    lens = [2048, 1024, 4096, 2048, 2048, 1024, 2048, 4096]
    starts = np.concatenate([[0], np.cumsum(lens)[:-1]])
    cap = max(lens) // cfg.t_int
    gs = synth.GpuSynth(sc)
    dev_node = SearchNode(cfg, sc.antpos_e, sc.antpos_n, max_block_steps=cap)
    want = [dev_node.process_block(gs.block(int(s0), n), n) for s0, n in zip(starts, lens)]
    pool = PinnedPagePool(len(lens), page_bytes(cfg, max(lens)))
    pages = []
    for s0, n in zip(starts, lens):
        p = pool.next_page()
        p[:page_bytes(cfg, n)].copy_(torch.from_numpy(sc.block(int(s0), n).reshape(-1)))
        pages.append(p)
    ingest = HostIngest(cfg, max_samp=max(lens))
    host_node = SearchNode(cfg, sc.antpos_e, sc.antpos_n, max_block_steps=cap)
    got = process_host_stream(host_node, ingest, pages, lens)
    n_c = 0
    for k, (a, b) in enumerate(zip(got, want)):
        assert a.T0 == b.T0 and a.n_steps == b.n_steps
        assert np.array_equal(np.sort(a.records, order=["t", "p", "d", "k"]),
                              np.sort(b.records, order=["t", "p", "d", "k"])), k
        assert [c.key + (c.snr,) for c in a.candidates] == [c.key + (c.snr,) for c in b.candidates], k
        n_c += len(a.candidates)
    assert np.array_equal(host_node.mean.cpu().numpy(), dev_node.mean.cpu().numpy())
    assert np.array_equal(host_node.m2.cpu().numpy(), dev_node.m2.cpu().numpy())
    assert n_c > 0


def test_copy_overlaps_compute():
    """Block i + 1's H2D copy starts before block i's dedispersion ends."""
    cfg = dcfg.search_node_short_config(4096)
    sc = synth.make_scenario(cfg)
    gs = synth.GpuSynth(sc)
    pool = PinnedPagePool(2, page_bytes(cfg, cfg.n_samp))
    for i, p in enumerate(pool.pages):
        p.copy_(gs.block(i * cfg.n_samp, cfg.n_samp).reshape(-1).cpu())
    ingest = HostIngest(cfg)
    node = SearchNode(cfg, sc.antpos_e, sc.antpos_n)
    node.enable_timing()
    b0 = ingest.submit(pool.pages[0], cfg.n_samp, timed=True)
    b0.wait()
    node.process_block(b0.raw, cfg.n_samp)
    ingest.release(b0)
    b1 = ingest.submit(pool.pages[1], cfg.n_samp, timed=True)      # copies while block 1 computes
    b2_page = pool.pages[0]
    b1.wait()
    start = torch.cuda.Event(enable_timing=True)
    start.record()
    b2 = ingest.submit(b2_page, cfg.n_samp, timed=True)
    node.process_block(b1.raw, cfg.n_samp)
    end = torch.cuda.Event(enable_timing=True)
    end.record()
    torch.cuda.synchronize()
    c_start, c_end, _ = ingest.copy_events[-1]
    # the copy of block 2 began before block 1's chain ended
    assert start.elapsed_time(c_start) < start.elapsed_time(end)
    rate = ingest.copy_rate()
    print(f"H2D {rate}", flush=True)
    assert rate["h2d_GBps"] > 5.0
