"""Pins of the stage-1 and static-sky oracle (CPU).

Known answers from the reference's own tests: every subband's first group
has shift 0, shifts are >= 0 and never exceed the whole-chgroup shifts
(tests/test_subband_n_sub.py:185-202); a per-channel impulse at
t0 + bins[c] lands on t0 in the output (tests/test_coarse_dm_stage1.py:
120-200).  Plus: a pulse dispersed at the coarse DM (continuous delays) is
aligned to within one step inside each subband after stage 1; the
static-sky window is the mean of the last `window` block means (brute
force), a constant sky vanishes after warm-up, and the current block is
never part of its own subtraction."""
import numpy as np

from oracle import corr as ocorr
from oracle import stage1 as os1

F64 = 1.49875 - 2.9296875e-3 * np.arange(64)          # 64 x 2.93 MHz from the band top


def test_stage1_bins_reference_invariants():
    for dm in (0.0, 100.0, 500.0, 1500.0):
        sub = os1.stage1_bins(F64, n_sub=16, group=1, t_int=8, dm=dm)
        whole = os1.stage1_bins(F64, n_sub=1, group=1, t_int=8, dm=dm)
        assert np.all(sub[::4] == 0)
        assert np.all(sub >= 0) and np.all(sub <= whole)
        assert np.all(np.diff(sub.reshape(16, 4), axis=1) >= 0)       # lower channel, later
    lo = os1.stage1_bins(F64, n_sub=16, group=1, t_int=8, dm=100.0)
    hi = os1.stage1_bins(F64, n_sub=16, group=1, t_int=8, dm=200.0)
    assert np.all(hi >= lo) and hi.max() > lo.max()
    # groups: every channel carries its group's shift, groups referenced to the subband's first
    g = os1.stage1_bins(F64, n_sub=4, group=4, t_int=8, dm=500.0)
    assert np.all(g.reshape(16, 4) == g.reshape(16, 4)[:, :1])
    assert np.all(g.reshape(4, 16)[:, :4] == 0)


def test_stage1_bins_two_roundings_at_dm_3000():
    """The delay across a 4-channel subband near the band top at DM 3000
    (closed form via the derivative), rounded to native samples, then steps."""
    b = os1.stage1_bins(F64, n_sub=16, group=1, t_int=8, dm=3000.0)
    d_us = 4.148808 * 3000 * (1 / F64[3] ** 2 - 1 / F64[0] ** 2) * 1e3
    assert b[3] == np.rint(np.rint(d_us / 32.768) / 8)
    approx = 4.148808 * 3000 * 2 * (F64[0] - F64[3]) / F64[0] ** 3 * 1e3 / 32.768 / 8
    assert abs(b[3] - approx) < 0.06 * approx + 1


def _pack(re, im):
    S, A, C, P = re.shape
    byte = ((re & 0xF) | ((im & 0xF) << 4)).astype(np.uint8)
    return np.ascontiguousarray(byte.reshape(S // 2, 2, A, C, P).transpose(0, 2, 3, 1, 4))


def test_stage1_impulse_lands_on_t0():
    """Channel c carries an impulse at step t0 + bins[c] (antenna 0 = (1, 0),
    antenna 1 = (0, 1)): after stage 1 every channel's baseline (1, 0) holds
    it at t0 only."""
    A, C, t_int, L = 2, 8, 8, 32
    f = F64[:C]
    bins = os1.stage1_bins(f, n_sub=2, group=1, t_int=t_int, dm=60.0)
    delay = int(bins.max())
    assert delay > 0
    S = (L + delay) * t_int
    re = np.zeros((S, A, C, 2), np.int64)
    im = np.zeros_like(re)
    t0 = 10
    for c in range(C):
        s0 = (delay + t0 + bins[c]) * t_int            # sample index in the concatenated stream
        re[s0:s0 + t_int, 0, c] = 1
        im[s0:s0 + t_int, 1, c] = 1
    raw = _pack(re, im)
    n_tail = delay * t_int // 2
    v = os1.stage1_visibilities(raw[:n_tail], raw[n_tail:], n_ant=A, n_chan=C, t_int=t_int, csf=1,
                                bins=bins, delay=delay)
    bls = 1 * 2 // 2 + 0
    hit = np.abs(v[:, :, bls, 1])
    assert hit.shape == (L, C)
    for c in range(C):
        assert hit[t0 + delay, c] == 2 * t_int and hit[:, c].sum() == 2 * t_int, c


def test_stage1_aligns_a_pulse_dispersed_at_the_coarse_dm():
    """Continuous per-channel delays of DM 800 (referenced to the band top):
    after stage 1 the channels of each subband peak within one step."""
    dm, t_int = 800.0, 8
    bins = os1.stage1_bins(F64, n_sub=16, group=1, t_int=t_int, dm=dm)
    arr_us = 4.148808 * dm * (1 / F64 ** 2 - 1 / F64[0] ** 2) * 1e3
    step = arr_us / 32.768 / t_int                      # arrival step per channel
    aligned = step - bins
    per_sub = aligned.reshape(16, 4)
    assert np.all(per_sub.max(axis=1) - per_sub.min(axis=1) < 1.0)
    raw_spread = step.reshape(16, 4)
    assert np.all(raw_spread.max(axis=1) - raw_spread.min(axis=1) > 3.0)


def test_static_sky_window_mean_warmup_and_causality():
    rng = np.random.default_rng(5)
    W, warm = 3, 2
    sky = os1.StaticSkyMean(window=W, warmup=warm)
    means = []
    for b in range(7):
        g = rng.normal(size=(5, 4, 4)) + 1j * rng.normal(size=(5, 4, 4))
        out = sky.apply(g)
        if b < warm:
            assert np.array_equal(out, g)
        else:
            want = g - np.mean(np.stack(means[-W:]), axis=0)[None]
            np.testing.assert_allclose(out, want, rtol=0, atol=1e-12)
        means.append(g.mean(axis=0))
    # a constant sky is removed exactly once the window holds it
    sky = os1.StaticSkyMean(window=2, warmup=1)
    c = np.full((3, 2, 2), 2.5 + 1j)
    outs = [sky.apply(c) for _ in range(4)]
    assert np.array_equal(outs[0], c) and all(np.abs(o).max() == 0 for o in outs[1:])
    # a burst in one block is not subtracted in that block
    sky = os1.StaticSkyMean(window=2, warmup=0)
    sky.apply(np.zeros((3, 2, 2), complex))
    burst = np.zeros((3, 2, 2), complex)
    burst[1, 0, 0] = 9.0
    assert sky.apply(burst)[1, 0, 0] == 9.0
