"""Forgetting of the running statistics: the oracle's merge + decay against
the weighted two-pass closed form, and the configuration rule.

After blocks 1..B with per-block factors lam_b (scaling applied after each
merge), block b's samples carry weight W_b = lam_b * lam_{b+1} * ... * lam_B,
so the state must equal n = sum_b W_b n_b, mean = sum_b W_b sum(x_b) / n and
M2 = sum_b W_b sum((x_b - mean)^2) computed directly over all samples.
Decaying the mean, decaying before instead of after the merge, or dropping
the newest block's factor each break this equality.
"""
import math

import numpy as np
import pytest

from dsx import config as dcfg
from oracle import search as osearch


def test_merge_and_decay_equal_weighted_two_pass():
    rng = np.random.default_rng(9)
    n_dm, npx = 3, 5
    n = np.zeros(n_dm)
    mean = np.zeros((n_dm, npx))
    m2 = np.zeros((n_dm, npx))
    blocks, lams = [], []
    for b in range(6):
        L = int(rng.integers(20, 90))
        x = rng.normal(3.0 * b, 1.0 + b, size=(n_dm, L, npx))
        cnt = np.full(n_dm, float(L))
        cnt[1] = 0.0 if b == 0 else cnt[1]            # a DM without samples in its first block
        xs = x * (cnt[:, None, None] > 0)
        n, mean, m2 = osearch.merge_stats(n, mean, m2, cnt, xs.sum(1), (xs * xs).sum(1))
        lam = osearch.block_decay(L, 8, 0.05)
        n, m2 = osearch.decay_stats(n, m2, lam)
        blocks.append((x, cnt))
        lams.append(lam)
    for d in range(n_dm):
        w = [math.prod(lams[b:]) for b in range(len(blocks))]
        wn = sum(w[b] * blocks[b][1][d] for b in range(len(blocks)))
        ws = sum(w[b] * blocks[b][0][d].sum(0) for b in range(len(blocks)) if blocks[b][1][d] > 0)
        mu = ws / wn
        wm2 = sum(w[b] * ((blocks[b][0][d] - mu) ** 2).sum(0) for b in range(len(blocks)) if blocks[b][1][d] > 0)
        assert n[d] == pytest.approx(wn, rel=1e-13)
        np.testing.assert_allclose(mean[d], mu, rtol=1e-12)
        np.testing.assert_allclose(m2[d], wm2, rtol=1e-11)


def test_block_decay_closed_form_and_config():
    # 256 steps of 8 native samples (262.144 us) over tau 0.1 s
    assert osearch.block_decay(256, 8, 0.1) == pytest.approx(math.exp(-256 * 8 * 32.768e-6 / 0.1), rel=1e-15)
    cfg = dcfg.small_node_config()
    assert cfg.stats_decay(256) == 1.0
    c2 = cfg.with_(stats_tau_s=0.1, min_stat_count=128)
    c2.validate()
    assert c2.stats_decay(256) == pytest.approx(osearch.block_decay(256, 8, 0.1), rel=1e-15)
    with pytest.raises(ValueError, match="saturates"):
        cfg.with_(stats_tau_s=0.01).validate()
    with pytest.raises(ValueError, match="stats_tau_s"):
        cfg.with_(stats_tau_s=-1.0).validate()
