"""Cross-kernel merger: ``dsx.merger`` against the pairwise oracle and the
reference's own known answers.

The known-answer table restates the inputs and expected survivor counts of
the reference's ``tests/test_c1_merger.py`` (C1 OR-cross rule) and
``tests/test_detector_merger.py`` (legacy box rule), each row citing the
test it follows.  The randomised cases compare the vectorised product with
``oracle.merger`` and check the invariants that determine a greedy
suppression scan.
"""
from __future__ import annotations

import random

import pytest

from dsx.detector import Candidate, CandidateFlags
from dsx.merger import (DEFAULT_MERGE_RADIUS_FDM, DEFAULT_MERGE_RADIUS_LM, DEFAULT_MERGE_RADIUS_T,
                        MergerConfig, merge_across_kernels, merge_across_kernels_c1)
from oracle import merger as om


def cand(*, snr, l=32.0, m=32.0, dm_idx=10, event_specnum=256, kernel_id="unit:d1:b1", width_samples=1):
    return Candidate(l=l, m=m, dm_fine=float(dm_idx), dm_idx=dm_idx, event_specnum=event_specnum,
                     width_samples=width_samples, kernel_id=kernel_id, snr=snr, detector_version="v1.M5",
                     flags=int(CandidateFlags.NONE), search_node_id=0, gpu_half=0)


def c1_both(cands, cfg):
    a = merge_across_kernels_c1(cands, cfg)
    b = om.merge_c1(cands, lm_max=cfg.lm_max_cells, dm_max=cfg.dm_max_trials, t_frac=cfg.t_frac,
                    sps=cfg.sample_period_specnum)
    assert a == b
    return a


def box_both(cands, **kw):
    a = merge_across_kernels(cands, **kw)
    b = om.merge_box(cands, r_lm=kw.get("merge_radius_lm", 3), r_fdm=kw.get("merge_radius_fdm", 5),
                     r_t=kw.get("merge_radius_t", 128))
    assert a == b
    return a


# (reference test, cfg kwargs, second candidate kwargs, expected survivors); first is cand(snr=12)
C1_PAIRS = [
    ("or_lm_small_delta_l_large_delta_m_merges", {}, dict(l=32.0, m=132.0), 1),
    ("or_lm_small_delta_m_large_delta_l_merges", {}, dict(l=132.0, m=32.0), 1),
    ("or_lm_both_axes_far_no_merge", {}, dict(l=132.0, m=132.0), 2),
    ("or_lm_inclusive_at_exact_boundary_l", {}, dict(l=35.0, m=132.0), 1),
    ("or_lm_inclusive_at_exact_boundary_m", {}, dict(l=132.0, m=35.0), 1),
    ("or_lm_just_outside_l_and_far_m_no_merge", {}, dict(l=36.0, m=132.0), 2),
    ("fdm_within_window_merges", {}, dict(dm_idx=12), 1),
    ("fdm_outside_window_no_merge", {}, dict(dm_idx=13), 2),
    ("fdm_dominates_when_lm_satisfied", {}, dict(l=33.0, m=33.0, dm_idx=15), 2),
    ("time_window_width_one_separation_4_no_merge", {}, dict(event_specnum=256 + 64), 2),
    ("time_window_width_one_separation_within_merges", {}, dict(event_specnum=256 + 16), 1),
    ("time_window_width_one_separation_just_over_no_merge", {}, dict(event_specnum=256 + 17), 2),
    ("sample_period_specnum_scales_window", dict(sample_period_specnum=1), dict(event_specnum=258), 2),
    ("in_lm_or_pass_but_t_fail_no_merge", {}, dict(l=33.0, m=33.0, event_specnum=273), 2),
    ("in_t_and_fdm_pass_but_lm_or_fail_no_merge", {}, dict(l=132.0, m=132.0), 2),
]


@pytest.mark.parametrize("name,cfg_kw,b_kw,n", C1_PAIRS, ids=[r[0] for r in C1_PAIRS])
def test_c1_known_answers(name, cfg_kw, b_kw, n):
    a = cand(snr=12.0)
    b = cand(snr=10.0, **b_kw)
    out = c1_both([b, a], MergerConfig(**cfg_kw))
    assert len(out) == n
    assert out[0].snr == 12.0


@pytest.mark.parametrize("w_a,w_b,t_b,t_frac,n", [
    (32, 32, 10_000 + 256, 1.0, 1),        # time_window_width_32_separation_16_samples_merges
    (32, 1, 10_017, 1.0, 1),               # time_window_asymmetric_widths_uses_mean_overlap
    (32, 32, 10_000 + 257, 0.5, 2),        # t_frac_half_tightens_window
    (1, 1, 10_000 + 64, 1.0, 2),           # multi_pulsar_pulses_width_one_separated_4_samples
])
def test_c1_width_window(w_a, w_b, t_b, t_frac, n):
    a = cand(snr=12.0, event_specnum=10_000, width_samples=w_a)
    b = cand(snr=10.0, event_specnum=t_b, width_samples=w_b)
    assert len(c1_both([a, b], MergerConfig(t_frac=t_frac))) == n


def test_c1_all_three_pass_merges():
    a = cand(snr=12.0, l=32.0, m=32.0, dm_idx=10, event_specnum=256, width_samples=4)
    b = cand(snr=10.0, l=34.0, m=70.0, dm_idx=11, event_specnum=300, width_samples=4)
    out = c1_both([a, b], MergerConfig())
    assert len(out) == 1 and out[0].snr == 12.0


def test_c1_config_validation_and_defaults():
    cfg = MergerConfig()
    assert (cfg.lm_max_cells, cfg.dm_max_trials, cfg.t_frac, cfg.sample_period_specnum) == (3, 2, 1.0, 16)
    for kw, msg in ((dict(lm_max_cells=-1), "lm_max_cells"), (dict(dm_max_trials=-1), "dm_max_trials"),
                    (dict(t_frac=-0.1), "t_frac"), (dict(sample_period_specnum=0), "sample_period_specnum")):
        with pytest.raises(ValueError, match=msg):
            MergerConfig(**kw)


def test_c1_trivial_and_ties():
    assert merge_across_kernels_c1([], MergerConfig()) == []
    c = cand(snr=9.0)
    assert merge_across_kernels_c1([c], MergerConfig()) == [c]
    a, b = cand(snr=10.0, kernel_id="unit:d1:b1"), cand(snr=10.0, kernel_id="psf:d1:b1")
    for order in ([a, b], [b, a]):
        out = c1_both(order, MergerConfig())
        assert [x.kernel_id for x in out] == ["psf:d1:b1"]
    chain = [cand(snr=20.0, m=32.0), cand(snr=15.0, m=33.0), cand(snr=10.0, m=34.0)]
    out = c1_both(chain, MergerConfig())
    assert [x.snr for x in out] == [20.0]


def test_c1_frb_burst_collapses_to_one():
    """test_c1_merger.py::test_frb_burst_collapses_to_one: 32 kernel
    triples of one pulse with width-scaled time jitter -> one survivor,
    the brightest."""
    cfg = MergerConfig()
    rng = random.Random(0)
    cands = []
    for img in ("unit", "psf", "psf_shift_lm", "psf_shift_l"):
        for dm in ("d1", "d3"):
            for tw in ("b4", "b8", "b16", "b32"):
                w = int(tw[1:])
                cands.append(cand(snr=14.0 + rng.uniform(-1.0, 1.0), l=32.0 + rng.choice([-1, 0, 0, 1]),
                                  m=32.0 + rng.choice([-1, 0, 0, 1]), dm_idx=20 + rng.choice([-1, 0, 0, 1]),
                                  event_specnum=30_000 + rng.choice([-w // 2, 0, 0, w // 2]) * 16,
                                  kernel_id=f"{img}:{dm}:{tw}", width_samples=w))
    out = c1_both(cands, cfg)
    assert len(out) == 1 and out[0].snr == max(c.snr for c in cands)
    far = [cand(snr=12.0, l=10.0, m=10.0, dm_idx=10, event_specnum=100),
           cand(snr=11.0, l=200.0, m=200.0, dm_idx=200, event_specnum=1_000_000)]
    assert len(c1_both(far, cfg)) == 2


def test_box_known_answers():
    """test_detector_merger.py cases on the legacy rule (radii lm 3, fdm 5, t 128)."""
    assert merge_across_kernels([]) == []
    hi, lo = cand(snr=12.0, kernel_id="psf:d3:b4", l=4.0, m=4.0), cand(snr=10.0, l=4.0, m=4.0)
    out = box_both([lo, hi])
    assert [(c.snr, c.kernel_id) for c in out] == [(12.0, "psf:d3:b4")]
    base = dict(l=4.0, m=4.0)
    cases = [(dict(m=4.0 + DEFAULT_MERGE_RADIUS_LM + 1), 2), (dict(dm_idx=10 + DEFAULT_MERGE_RADIUS_FDM + 1), 2),
             (dict(event_specnum=256 + DEFAULT_MERGE_RADIUS_T + 1), 2), (dict(m=4.0 + DEFAULT_MERGE_RADIUS_LM), 1)]
    for kw, n in cases:
        b = cand(snr=10.0, **{**base, **kw})
        assert len(box_both([cand(snr=12.0, **base), b])) == n
    with pytest.raises(ValueError, match="merge_radius"):
        merge_across_kernels([], merge_radius_lm=-1)
    rng = random.Random(0)
    burst = [cand(snr=14.0 + rng.uniform(-1, 1), l=32.0 + rng.choice([-1, 0, 0, 1]),
                  m=32.0 + rng.choice([-1, 0, 0, 1]), dm_idx=20 + rng.choice([-1, 0, 0, 1]),
                  event_specnum=300 + rng.choice([-2, -1, 0, 1, 2]), kernel_id=f"{i}:d1:b1")
             for i in ("unit", "psf", "psf_shift_lm", "psf_shift_l") for _ in range(4)]
    out = box_both(burst)
    assert len(out) == 1 and out[0].snr == max(c.snr for c in burst)


def _random_cands(rng, n, span_t=2000, span_dm=40, span_lm=40):
    ids = ["unit:d1:b1", "psf:d1:b1", "unit:d3:b4"]
    return [cand(snr=round(rng.uniform(8, 30), 1), l=float(rng.randrange(span_lm)), m=float(rng.randrange(span_lm)),
                 dm_idx=rng.randrange(span_dm), event_specnum=rng.randrange(span_t), kernel_id=rng.choice(ids),
                 width_samples=rng.choice([1, 2, 4, 8, 16, 32])) for _ in range(n)]


@pytest.mark.parametrize("seed", range(6))
def test_random_matches_oracle_and_invariants(seed):
    rng = random.Random(seed)
    cands = _random_cands(rng, 300)
    cfg = MergerConfig(lm_max_cells=rng.choice([0, 2, 3]), dm_max_trials=rng.choice([0, 2, 4]),
                       t_frac=rng.choice([0.5, 1.0]), sample_period_specnum=rng.choice([1, 16]))
    out = c1_both(cands, cfg)
    kw = dict(lm_max=cfg.lm_max_cells, dm_max=cfg.dm_max_trials, t_frac=cfg.t_frac, sps=cfg.sample_period_specnum)
    # no survivor lies in the window of an earlier survivor
    for i, c in enumerate(out):
        assert not any(om.c1_pair(c, s, **kw) for s in out[:i])
    # every dropped candidate lies in the window of a survivor that precedes it
    key = lambda c: (-c.snr, c.event_specnum, c.dm_idx, c.l, c.m, c.kernel_id)
    kept = set(map(id, out))
    for c in cands:
        if id(c) not in kept:
            assert any(om.c1_pair(c, s, **kw) and key(s) <= key(c) for s in out)
    shuf = cands[:]
    rng.shuffle(shuf)
    assert merge_across_kernels_c1(shuf, cfg) == out
    box = box_both(cands, merge_radius_lm=3, merge_radius_fdm=2, merge_radius_t=64)
    for i, c in enumerate(box):
        assert not any(om.box_pair(c, s, r_lm=3, r_fdm=2, r_t=64) for s in box[:i])
