"""Pins of the dedispersion / detection oracle (CPU).

Known answers: 1697 < delta_tau(nu_bot_proc, nu_top_proc, DM 3000) < 1699 ms
(tests/test_numerical_conventions.py:85-102); the closed-form shift table
(tests/test_fine_dm_combiner.py:148-196); brute-force shift-and-add with
zero fill (tests/test_fine_dm_combiner.py:336-458); explicit sliding-window
boxcar sums (tests/test_detector_protocol.py:371-448); an injected boxcar of
amplitude snr/sqrt(w) on unit-variance statistics scoring exactly snr
(inject/cube_injection.py:285-329); Chan merge == two-pass float64
moments; decoder == an exhaustive greedy restatement.  The C inner loop is
checked against the numpy twin.
"""
import numpy as np
import pytest

from oracle import search as osearch

F0, DNU = 1.530, 0.250 / 8192


def _freq(g, ch):
    return F0 - (1024 + 384 * g + ch) * DNU


def test_delta_tau_known_answer():
    nu_bot, nu_top = _freq(15, 383), _freq(0, 0)
    d_ms = osearch.K_DM * 3000.0 * (1.0 / nu_bot ** 2 - 1.0 / nu_top ** 2)
    assert 1697.0 < d_ms < 1699.0


def test_shift_table_closed_form_and_monotone():
    dm = np.linspace(0, 1500, 17)
    refs = np.array([_freq(0, 24 * s) for s in range(16)])
    sh = osearch.shift_table(dm, refs, _freq(0, 383), 262.144)
    assert np.all(sh[0] == 0)
    for d in (3, 16):
        for s in (0, 7, 15):
            want = int(np.rint(4.148808 * dm[d] * (1 / _freq(0, 383) ** 2 - 1 / refs[s] ** 2) * 1000 / 262.144))
            assert sh[d, s] == want
    assert np.all(np.diff(sh, axis=1) <= 0)           # higher subbands read further back
    assert np.all(np.diff(sh, axis=0) >= 0)           # larger DMs read further back


def test_dedisperse_brute_force():
    rng = np.random.default_rng(0)
    img = rng.standard_normal((3, 40, 5))
    sh = np.array([[0, 0, 0], [5, 2, 0], [9, 4, 1]], dtype=np.int32)
    D = osearch.dedisperse_np(img, 0, sh, -3, 30)
    for d in range(3):
        for i, t in enumerate(range(-3, 30)):
            want = np.zeros(5)
            if t >= 0:
                for s in range(3):
                    if t - sh[d, s] >= 0:
                        want += img[s, t - sh[d, s]]
            np.testing.assert_allclose(D[d, i], want, atol=1e-12)


@pytest.mark.parametrize("w", [1, 2, 3, 4, 7, 8, 16])
def test_boxcar_matches_sliding_window(w):
    rng = np.random.default_rng(20260505)
    x = rng.standard_normal((64, 6))
    got = osearch.boxcar_np(x, w)
    pad_l, pad_r = w // 2, w - 1 - w // 2
    xp = np.concatenate([np.zeros((pad_l, 6)), x, np.zeros((pad_r, 6))])
    want = np.lib.stride_tricks.sliding_window_view(xp, w, axis=0).sum(axis=-1)
    np.testing.assert_allclose(got, want, atol=1e-12)


def test_merge_stats_equals_two_pass():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((3, 500, 4)) * 2 + 0.3
    n, mean, m2 = np.zeros(3), np.zeros((3, 4)), np.zeros((3, 4))
    for lo, hi in ((0, 7), (7, 200), (200, 500)):
        blk = x[:, lo:hi]
        n, mean, m2 = osearch.merge_stats(n, mean, m2, np.full(3, hi - lo), blk.sum(1), (blk ** 2).sum(1))
    np.testing.assert_allclose(mean, x.mean(1), rtol=1e-12)
    np.testing.assert_allclose(m2 / n[:, None], x.var(1), rtol=1e-10)


def _small_case(seed=5):
    rng = np.random.default_rng(seed)
    n_sub, n_pix, n_tau = 4, 6, 400
    img = rng.standard_normal((n_sub, n_tau, n_pix))
    sh = np.array([[0, 0, 0, 0], [6, 4, 2, 0], [12, 8, 4, 1], [20, 13, 7, 2]], dtype=np.int32)
    # a dispersed width-4 pulse for DM 2 at t = 300, pixel 3
    for s in range(n_sub):
        img[s, 300 - sh[2, s]:304 - sh[2, s], 3] += 3.0
    return img, sh


def test_search_c_matches_numpy_twin():
    img, sh = _small_case()
    widths = [1, 2, 4, 8]
    st_n = np.array([0.0, 600.0, 600.0, 600.0])
    st_mean = np.full((4, 6), 0.05)
    st_m2 = np.full((4, 6), 600.0 * 4.0)
    args = (img, 0, sh, 192, 192, 64, widths, st_n, st_mean, st_m2, 512)
    a = osearch.search_block_np(*args)
    b = osearch.search_block_c(*args)
    fin = np.isfinite(a[0])
    assert np.array_equal(fin, np.isfinite(b[0]))
    np.testing.assert_allclose(a[0][fin], b[0][fin], rtol=1e-10)
    np.testing.assert_array_equal(a[1][fin], b[1][fin])
    np.testing.assert_array_equal(a[2][fin], b[2][fin])
    np.testing.assert_allclose(a[3], b[3], rtol=1e-12)
    np.testing.assert_allclose(a[4], b[4], rtol=1e-12)
    np.testing.assert_array_equal(a[5], b[5])
    assert np.all(np.isinf(a[0][0]))                   # DM 0 has no statistics yet


def test_injected_boxcar_scores_exact_snr():
    n_sub, n_pix = 2, 1
    img = np.zeros((n_sub, 300, n_pix))
    snr, w = 12.0, 4
    img[0, 150:150 + w, 0] = snr / np.sqrt(w)      # unit sigma, zero mean, DM 0
    sh = np.zeros((1, n_sub), dtype=np.int32)
    bs, bt, bk, *_ = osearch.search_block_c(img, 0, sh, 128, 64, 64, [1, 2, 4, 8], np.array([1e6]),
                                            np.zeros((1, 1)), np.full((1, 1), 1e6), 1)
    assert abs(bs[0, 0, 0] - snr) < 1e-12
    assert bk[0, 0, 0] == 2 and bt[0, 0, 0] == 150 + w // 2


def _decode_exhaustive(recs, thr, n_top, rt, rd, rl, n_grid):
    """Greedy suppression written independently: repeatedly take the best
    remaining record (ties: smaller t, dm, pixel), drop its neighbours."""
    pool = [r for r in recs if r[4] > thr]
    pool.sort(key=lambda r: (-r[4], r[0], r[1], r[2]))
    pool = pool[:n_top]
    out = []
    for r in pool:
        rr, cc = divmod(r[2], n_grid)
        if all(not (abs(r[0] - o[0]) <= rt and abs(r[1] - o[1]) <= rd and abs(rr - o[5]) <= rl
                    and abs(cc - o[6]) <= rl) for o in out):
            out.append((r[0], r[1], r[2], r[3], r[4], rr, cc))
    return [(o[0], o[1], o[5], o[6]) for o in out]


def test_decode_matches_exhaustive_greedy():
    rng = np.random.default_rng(9)
    n_dm, n_ch, n_pix, n_grid = 12, 3, 64, 8
    bs = rng.uniform(0, 12, size=(n_dm, n_ch, n_pix))
    bs[3, 1, 10] = bs[4, 1, 10] = 11.9999                # exact tie -> lower DM first
    bt = np.broadcast_to((np.arange(n_ch) * 64 + 5)[None, :, None], bs.shape).astype(np.int32)
    bk = rng.integers(0, 4, size=bs.shape).astype(np.int32)
    got = osearch.decode(bs, bt, bk, [1, 2, 4, 8], n_grid, threshold=9.0, n_top=50, radius_t=10,
                         radius_dm=2, radius_lm=1)
    recs = [(int(bt[d, c, p]), d, p, int(bk[d, c, p]), float(bs[d, c, p]))
            for d in range(n_dm) for c in range(n_ch) for p in range(n_pix)]
    want = _decode_exhaustive(recs, 9.0, 50, 10, 2, 1, n_grid)
    assert [(c.t, c.dm_idx, c.row, c.col) for c in got] == want
