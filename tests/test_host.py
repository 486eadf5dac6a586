"""Host-side logic of the fast path (CPU): tables the kernels consume,
the imager's Hermitian-fold math (numpy twin of dsx_image_subbands), the
decoder, the generator's integer rules, and configuration checks."""
import numpy as np
import pytest

from dsx import config as dcfg
from dsx import synth
from dsx.detector import CAND_DTYPE, decode_chunk_maxima
from dsx.dispersion import delta_tau_ms
from dsx.pattern import (TAP_CELL_MASK, TAP_CONJ_BIT, TAP_NEG_BIT, build_cell_map,
                         build_imager_taps, upper_tri_indices)
from oracle import image as oimage
from oracle import search as osearch


@pytest.mark.parametrize("make", [dcfg.small_node_config, dcfg.search_node_config])
def test_cell_map_matches_oracle(make):
    cfg = make()
    e, n = synth.antenna_layout(cfg)
    cm = build_cell_map(e, n, cfg.grid_freq_out_GHz, cfg.n_grid)
    rows, cols, cell = oimage.cell_map(e, n, cfg.grid_freq_out_GHz, cfg.n_grid)
    assert cell == cm.cell_lambda
    assert np.array_equal(cm.row, rows) and np.array_equal(cm.col, cols)


@pytest.mark.parametrize("make", [dcfg.small_node_config, dcfg.search_node_config])
def test_shift_table_matches_oracle(make):
    cfg = make()
    want = osearch.shift_table(cfg.dm_trials, cfg.subband_ref_GHz, cfg.nu_bot_GHz, cfg.t_step_us)
    assert np.array_equal(cfg.shift_table(), want)


def _fold_image_numpy(vis_slab, taps, s, n_grid):
    """Float64 twin of one dsx_image_subbands work item: scatter the taps into
    the Hermitian half plane, row transforms, column C2R, / (2 N^2)."""
    N = n_grid
    H = np.zeros((N // 2 + 1, N), dtype=np.complex128)
    lo, hi = taps.offsets[s], taps.offsets[s + 1]
    for i in range(lo, hi):
        v = vis_slab[taps.src[i]]
        d = int(taps.dst[i])
        if d & TAP_CONJ_BIT:
            v = np.conj(v)
        if d & TAP_NEG_BIT:
            v = -v
        cell = d & TAP_CELL_MASK
        H[cell // N, cell % N] += v
    Y = np.fft.ifft(H, axis=1) * N                       # sum_w H e^{+2 pi i w j / N}
    r = np.arange(N // 2 + 1)
    c = np.where((r == 0) | (r == N // 2), 1.0, 2.0)
    ph = np.exp(2j * np.pi * np.outer(np.arange(N), r) / N)     # (i, r)
    out = np.real(ph @ (c[:, None] * Y)) / (2.0 * N * N)
    return out


@pytest.mark.parametrize("make", [dcfg.small_node_config, dcfg.search_node_config])
def test_imager_fold_matches_reference_image(make):
    """The tap table + fold + C2R equals mask * fftshift(Re(ifft2(G)))."""
    cfg = make()
    e, n = synth.antenna_layout(cfg)
    cm = build_cell_map(e, n, cfg.grid_freq_out_GHz, cfg.n_grid)
    taps = build_imager_taps(cm, cfg.n_sub)
    rng = np.random.default_rng(7)
    N = cfg.n_grid
    mask = oimage.edge_mask(N, 1)
    mask01 = np.abs(mask)
    for s in (0, cfg.n_sub - 1):
        vis = (rng.integers(-500, 500, size=(cfg.per_sub, cfg.nbase))
               + 1j * rng.integers(-500, 500, size=(cfg.per_sub, cfg.nbase)))
        sl = slice(s * cfg.per_sub, (s + 1) * cfg.per_sub)
        G = oimage.grid_plane(vis.T, cm.row[:, sl], cm.col[:, sl], N)
        want = oimage.image(G, mask)
        got = _fold_image_numpy(vis.reshape(-1), taps, s, N) * mask01
        np.testing.assert_allclose(got, want, atol=1e-9 * np.abs(want).max())
    assert taps.n_rows <= N // 2 + 1


def test_upper_tri_indices():
    a, b = upper_tri_indices(5)
    assert np.array_equal(a * (a + 1) // 2 + b, np.arange(15))


def test_quantise_half_even_and_clip():
    v = np.array([0, 511, 512, 513, 1536, -512, -1536, 100000, -100000, 2560])
    assert list(synth.quantise_fixed(v)) == [0, 0, 0, 1, 2, 0, -2, 7, -7, 2]


def test_splitmix_known_values():
    # splitmix64 reference values for seed 0, counters 1, 2 (top 16 bits)
    got = synth.splitmix_top16(0, np.array([1, 2], dtype=np.int64))
    assert list(got) == [0xE220, 0x6E78]


def test_noise_lut_and_quantised_variance():
    lut = synth.noise_lut(2.5)
    assert lut.shape == (65536,) and abs(lut.mean()) < 1
    assert np.all(np.diff(lut) >= 0)
    v = synth.quantised_variance(2.5)
    assert 5.9 < v < 6.5                                  # round(N(0, 2.5^2)) clipped to +-7


def test_synth_stream_independent_of_block_split():
    cfg = dcfg.small_node_config().with_(n_samp=256)
    sc = synth.make_scenario(cfg)
    whole = sc.block(0, 256)
    parts = np.concatenate([sc.block(0, 64), sc.block(64, 192)])
    assert np.array_equal(whole, parts)


def test_small_node_burst_calibration_targets_snr():
    cfg = dcfg.small_node_config()
    sc = synth.make_scenario(cfg, synth.small_node_bursts(cfg))
    b = sc.bursts[0]
    assert b.amp > 0
    unit = synth.expected_unit_snr(cfg, sc.cmap, sc.antpos_e, sc.antpos_n,
                                   synth.Burst(b.dm, b.t0_samp, b.width, 1.0, b.row, b.col),
                                   cfg.shift_table())
    assert abs(unit * b.amp ** 2 - 20.0) < 1e-9


def test_decode_chunk_maxima_matches_oracle_decode():
    rng = np.random.default_rng(3)
    n = 400
    rec = np.zeros(n, dtype=CAND_DTYPE)
    rec["t"] = rng.integers(0, 2000, n)
    rec["p"] = rng.integers(0, 64 * 64, n)
    rec["d"] = rng.integers(0, 64, n)
    rec["k"] = rng.integers(0, 4, n)
    rec["snr"] = rng.uniform(8, 30, n).astype(np.float32)
    cfg = dcfg.small_node_config()
    got = decode_chunk_maxima(rec, widths=cfg.widths, n_grid=64, dm_trials=cfg.dm_trials,
                              threshold=8.0, n_top=100, radius_t=128, radius_dm=5, radius_lm=3)
    # same records through the oracle decoder (one chunk per record)
    bs = np.full((64, n, 64 * 64), -np.inf)
    bt = np.zeros(bs.shape, dtype=np.int32)
    bk = np.zeros(bs.shape, dtype=np.int32)
    for i, r in enumerate(rec):
        bs[r["d"], i, r["p"]] = r["snr"]
        bt[r["d"], i, r["p"]] = r["t"]
        bk[r["d"], i, r["p"]] = r["k"]
    want = osearch.decode(bs, bt, bk, list(cfg.widths), 64, threshold=8.0, n_top=100, radius_t=128,
                          radius_dm=5, radius_lm=3)
    assert [(c.dm_idx, c.time_bin, c.width_samples, int(c.l), int(c.m)) for c in got] == \
           [(c.dm_idx, c.t, c.width, c.row, c.col) for c in want]


def test_config_validation():
    with pytest.raises(ValueError):
        dcfg.small_node_config().with_(widths=(1, 3)).validate()
    with pytest.raises(ValueError):
        dcfg.small_node_config().with_(n_sub=7).validate()
    with pytest.raises(ValueError):
        dcfg.small_node_config().with_(n_samp=1000).validate()
    dcfg.small_node_config().validate()
    dcfg.search_node_config().validate()


def test_dispersion_sign_and_zero():
    assert delta_tau_ms(1.0, 2.0, 0.0) == 0.0
    assert delta_tau_ms(1.0, 2.0, 100.0) > 0
