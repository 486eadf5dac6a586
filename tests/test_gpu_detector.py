"""The Detector protocol drop-in (DeterministicDetector.forward over a
[T_det, N_fdm, H, W] cube) against the float64 oracle, three cubes in a
row so the Layer-2 sigma_k state carries: boxcar widths 1-16, injected
boxes at several (t, fdm, pixel), float32 and float16 cubes.

Per cube: the GPU's cells above threshold (count) and its strongest
n_top cells, the decoded candidates (time, fine DM, pixel, width, S/N) and
the updated sigma_k equal the oracle's: S/N within 1e-4 relative, sigma_k
within 1e-5 relative (the oracle clips the same seeded sample of interior
cells).  A float16 cube of small integers makes exact S/N ties common and
pins their order (flat index ascending)."""
import numpy as np
import pytest
import torch

from dsx.detector import DeterministicDetector
from oracle import detect as od

pytestmark = pytest.mark.gpu

WIDTHS = (1, 2, 4, 8, 16)


def _cube(rng, T, F, N, blk, ints=False):
    if ints:
        x = rng.integers(-3, 4, size=(T, F, N, N)).astype(np.float32)
    else:
        x = rng.standard_normal((T, F, N, N)).astype(np.float32)
    for t, f, l, m, w, a in ((20 + blk, 2, 5, 6, 4, 4.0), (50, 4, 10, 3, 8, 3.0), (70 - blk, 0, 12, 12, 1, 9.0)):
        x[t:t + w, f, l, m] += a
    return x


@pytest.mark.parametrize("dtype,ints", [(torch.float32, False), (torch.float16, False), (torch.float16, True)])
def test_detector_forward_vs_oracle(dtype, ints):
    rng = np.random.default_rng(11 + int(ints))
    T, F, N = 96, 6, 16
    thr = 5.0 if not ints else 4.0
    det = DeterministicDetector(time_widths=WIDTHS, threshold_sigma=thr, merge_radius_t=6, merge_radius_fdm=1,
                                merge_radius_lm=1, layer2_n_burnin=2, cube_cadence_s=0.25, layer2_tau_s=1.0,
                                streaming_decoder_n_top=128, layer2_sigma_max_samples=4096)
    L2 = od.Layer2(len(WIDTHS), cadence_s=0.25, tau_s=1.0, n_burnin=2, seed=[np.sqrt(w) for w in WIDTHS])
    valid = torch.ones((T, F), dtype=torch.bool, device="cuda")
    sig1 = torch.ones((T, F), dtype=torch.float32, device="cuda")
    n_cands = 0
    for blk in range(3):
        x = _cube(rng, T, F, N, blk, ints)
        xt = torch.from_numpy(x).cuda().to(dtype)
        x64 = xt.float().cpu().numpy().astype(np.float64)
        s_prev = L2.s_k.copy()
        np.testing.assert_allclose(det.layer2_state.s_k, s_prev, rtol=1e-5)
        cands = det.forward(xt, valid, sig1)
        snr, win, t_base = od.cube_scores(x64, WIDTHS, s_prev)
        want = od.decode_topk(snr, win, threshold=thr, n_top=128, radius_t=6, radius_fdm=1, radius_lm=1)
        assert det.last_n_above == int((snr > thr).sum()), (det.last_n_above, int((snr > thr).sum()))
        got = [(c.time_bin - t_base, int(c.dm_idx), int(c.l), int(c.m), c.width_samples) for c in cands]
        ok = [(t, f, l, m, WIDTHS[k]) for t, f, l, m, k, _ in want]
        assert got == ok, f"cube {blk}: gpu {got[:6]} oracle {ok[:6]}"
        for c, w in zip(cands, want):
            assert abs(c.snr - w[5]) <= 1e-4 * abs(w[5]), (c, w)
            assert c.kernel_id == f"unit:d1:b{c.width_samples}"
            assert bool(c.flags & 8) == (blk < 2)                 # NOISE_WARMUP during burn-in
        n_cands += len(cands)
        # Layer-2 update from the same sample of interior cells
        cells = det._samples.cpu().numpy()
        B = od.box_sums(x64, WIDTHS)[:, t_base:T - t_base].reshape(len(WIDTHS), -1)
        L2.update(np.array([od.sigma_clipped_std(B[k, cells]) for k in range(len(WIDTHS))]))
        np.testing.assert_allclose(det.layer2_state.s_k, L2.s_k, rtol=1e-5)
    assert n_cands > 0


def test_detector_validation_and_invalid_cube():
    det = DeterministicDetector(time_widths=(1, 2, 4), layer2_n_burnin=1)
    x = torch.zeros((16, 2, 8, 8), device="cuda")
    with pytest.raises(ValueError, match="expected 4"):
        det.forward(x[0], torch.ones((16, 2), dtype=torch.bool, device="cuda"), None)
    with pytest.raises(TypeError, match="torch.bool"):
        det.forward(x, torch.ones((16, 2), device="cuda"), None)
    with pytest.raises(ValueError, match="N_fdm"):
        det.forward(x, torch.ones((16, 3), dtype=torch.bool, device="cuda"), None)
    s0 = det.layer2_state.s_k
    bad = torch.ones((16, 2), dtype=torch.bool, device="cuda")
    bad[3, 1] = False
    det.forward(torch.randn((16, 2, 8, 8), device="cuda"), bad, None)
    assert np.array_equal(det.layer2_state.s_k, s0) and det.layer2_state.cube_count == 0


def test_detector_c1_merger_and_wire():
    """With a MergerConfig the detector's survivors are the oracle decode's
    candidates after the pairwise C1 merge, and they encode to a batch
    that parses back to the same rows."""
    from dsx import wire
    from dsx.merger import MergerConfig
    from oracle import merger as om
    rng = np.random.default_rng(5)
    T, F, N = 96, 6, 16
    mc = MergerConfig(lm_max_cells=1, dm_max_trials=1, sample_period_specnum=1)
    kw = dict(time_widths=WIDTHS, threshold_sigma=4.5, layer2_n_burnin=2, streaming_decoder_n_top=256,
              layer2_sigma_max_samples=4096)
    det = DeterministicDetector(merger_config=mc, **kw)
    ref = DeterministicDetector(**kw)
    valid = torch.ones((T, F), dtype=torch.bool, device="cuda")
    for blk in range(2):
        xt = torch.from_numpy(_cube(rng, T, F, N, blk)).cuda()
        got = det.forward(xt, valid, None, event_specnum=1000 * blk)
        base = ref.forward(xt, valid, None, event_specnum=1000 * blk)
        want = om.merge_c1(base, lm_max=1, dm_max=1, t_frac=1.0, sps=1.0)
        assert [c.key for c in got] == [c.key for c in want] and len(got) <= len(base)
        geom = wire.CubeGeometry(cube_id=blk, specnum_start=1000 * blk, sample_period_specnum=1, t_det=T, n_grid=N,
                                 n_fdm_in_cube=F, sample_period_us=1048.576, cell_l_rad=1e-3, cell_m_rad=1e-3,
                                 l0_rad=0.0, m0_rad=0.0, fine_dm_pc_cc=np.arange(F, dtype=np.float64),
                                 mjd_start=60000.0)
        blob, surv = wire.cube_batch(got, geom, merger=None)
        rows = wire.parse_c1_batch(blob.decode().rstrip("\n").split("\n")).candidates
        assert [(r.l_pix, r.m_pix, r.fine_dm_idx, r.event_specnum, r.width_samples) for r in rows] == \
            [(int(c.l), int(c.m), c.dm_idx, c.event_specnum, c.width_samples) for c in surv]
