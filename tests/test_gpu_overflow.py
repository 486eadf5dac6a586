"""Candidate-record overflow on 16-antenna 64-channel inputs.  With a record
capacity far below the number of chunk maxima above threshold, a block sets
``overflow`` and counts every record in ``n_raw``; it is then detected again
(records only, against the same history and noise state) at the lowest
record-count level (threshold x sqrt(2)^j) whose records fit, so the kept
records are exactly the uncapped run's records above that level -- the
strongest ones, independent of the order of the kernel's atomic appends --
and the candidates decode from them.  The statistics do not depend on the
capacity, so later blocks match too.  The uneven plan applies the policy per
DM range."""
import numpy as np
import pytest

from dsx import config as dcfg
from dsx import synth
from dsx.detector import record_levels
from dsx.pipeline import MultiRangeNode, SearchNode

pytestmark = pytest.mark.gpu

N_BLOCKS = 4


def _recs(rec):
    return {(int(r["d"]), int(r["t"]), int(r["p"]), int(r["k"])): float(r["snr"]) for r in rec}


def _run(node, cfg, gs):
    return [node.process_block(gs.block(b * cfg.n_samp, cfg.n_samp), cfg.n_samp) for b in range(N_BLOCKS)]


def _check_range(f_rec, c_rec, thr_used, cap, blk):
    want = {k: v for k, v in _recs(f_rec).items() if v > thr_used}
    got = _recs(c_rec)
    assert got == want, f"block {blk}: kept records are not the uncapped run's above {thr_used}"
    assert len(got) <= cap


def test_search_node_overflow_keeps_strongest_records():
    cfg = dcfg.small_node_config().with_(threshold=3.5, min_stat_count=128)
    sc = synth.make_scenario(cfg, synth.small_node_bursts(cfg))
    full = _run(SearchNode(cfg, sc.antpos_e, sc.antpos_n), cfg, synth.GpuSynth(sc))
    capped = _run(SearchNode(cfg, sc.antpos_e, sc.antpos_n, cand_cap=16), cfg, synth.GpuSynth(sc))
    levels = record_levels(cfg.threshold)
    overflowed = 0
    for blk, (f, c) in enumerate(zip(full, capped)):
        print(f"block {blk}: {f.n_raw} records, kept {len(c.records)} above {c.threshold_used}", flush=True)
        assert not f.overflow and f.threshold_used == np.float32(cfg.threshold)
        assert c.n_raw == f.n_raw == len(f.records)
        if c.overflow:
            overflowed += 1
            assert c.threshold_used in levels[1:] and c.threshold_used > cfg.threshold
            # the lowest level that fits: the level below it holds more than 16 records
            j = int(np.nonzero(levels == np.float32(c.threshold_used))[0][0])
            assert sum(1 for r in f.records if r["snr"] > levels[j - 1]) > 16
        _check_range(f.records, c.records, c.threshold_used, 16, blk)
        want = [x.key for x in f.candidates if x.snr > c.threshold_used]
        assert [x.key for x in c.candidates] == want, blk
    assert overflowed > 0


def test_multi_range_overflow_per_range():
    cfg = dcfg.small_node_config().with_(
        widths=(1, 2, 4, 8), threshold=3.5, min_stat_count=128, n_dm=72,
        dm_ranges=((0.0, 150.0, 32, 1, False), (150.0, 300.0, 24, 2, False), (300.0, 500.0, 16, 4, True)))
    sc = synth.make_scenario(cfg, synth.small_node_bursts(cfg))
    full = _run(MultiRangeNode(cfg, sc.antpos_e, sc.antpos_n), cfg, synth.GpuSynth(sc))
    capped = _run(MultiRangeNode(cfg, sc.antpos_e, sc.antpos_n, cand_cap=8), cfg, synth.GpuSynth(sc))
    bounds = np.cumsum([0] + [r[2] for r in cfg.dm_ranges])
    overflowed = 0
    for blk, (f, c) in enumerate(zip(full, capped)):
        assert c.n_raw == f.n_raw and not f.overflow
        overflowed += int(c.overflow)
        for r, thr in enumerate(c.range_thresholds):
            sel = lambda rec: rec[(rec["d"] >= bounds[r]) & (rec["d"] < bounds[r + 1])]
            _check_range(sel(f.records), sel(c.records), thr, 8, blk)
    assert overflowed > 0
