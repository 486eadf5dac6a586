"""Pins of the gridder / imager oracle (CPU).

Known answers from the reference's tests: DC-only grid -> uniform 1/N^2
image and the single-side identity (tests/test_image_imager.py:132-175),
point source of amplitude 2.0 on 240 cells imaging to 2.0 * 240 / 32^2 at
pixel (9, 21) (tests/x2/test_search2x_ref.py:97-104), mask geometry
(tests/test_image_imager.py:83-130).  The cell map is pinned by its closed
form for single baselines (F20 negation) and the auto / out-of-grid rules.
"""
import numpy as np

from oracle import image as oimage


def test_dc_only_grid_images_uniform():
    n = 8
    uv = np.zeros((n, n), dtype=np.complex128)
    uv[0, 0] = 1.0
    img = oimage.image(uv, np.ones((n, n)))
    np.testing.assert_allclose(img, np.full((n, n), 1.0 / n ** 2), rtol=1e-12, atol=1e-15)


def test_single_side_identity():
    rng = np.random.default_rng(0)
    n = 8
    truth = rng.standard_normal((n, n))
    full = np.fft.fft2(truth)
    v_pos = np.zeros_like(full)
    v_pos[:, 1:n // 2] = full[:, 1:n // 2]
    v_pos[:, 0] = 0.5 * full[:, 0]
    v_pos[:, n // 2] = 0.5 * full[:, n // 2]
    rec = np.fft.ifftshift(oimage.image(v_pos, np.ones((n, n))))
    np.testing.assert_allclose(2.0 * rec, truth, rtol=1e-10, atol=1e-10)


def _point_source_values(cells, n, l_pix, m_pix, amp):
    h = n // 2
    u = cells // n - h
    w = cells % n - h
    return amp * np.exp(-2j * np.pi * (u * (l_pix - h) + w * (m_pix - h)) / n)


def test_point_source_known_answer():
    N = 32
    rng = np.random.default_rng(2)
    G = np.zeros(N * N, dtype=np.complex128)
    for _ in range(4):
        cells = rng.choice(N * N, size=60, replace=False)
        np.add.at(G, cells, _point_source_values(cells, N, 9, 21, 2.0))
    img = oimage.image(G.reshape(N, N), oimage.edge_mask(N, 1))
    assert np.unravel_index(np.argmax(img), img.shape) == (9, 21)
    assert abs(img[9, 21] - 2.0 * 240 / N ** 2) < 1e-12


def test_edge_mask_geometry():
    N = 16
    m = oimage.edge_mask(N, 1)
    npad = oimage.mask_npad(N, 1)
    assert npad == 2
    assert np.all(m[:npad] == 0) and np.all(m[-npad:] == 0)
    assert np.all(m[:, :npad] == 0) and np.all(m[:, -npad:] == 0)
    assert m[N // 2, N // 2] == 0
    ii, jj = np.indices((N, N))
    inner = (ii >= npad) & (ii < N - npad) & (jj >= npad) & (jj < N - npad) & ~((ii == N // 2) & (jj == N // 2))
    assert np.array_equal(np.abs(m)[inner], np.ones(inner.sum()))
    assert np.array_equal(np.sign(m[inner]), (1 - 2 * ((ii + jj) & 1))[inner])


def test_cell_map_closed_form_single_baseline():
    # two antennas: (0, 0) and (E, N); baseline bls = 1*2/2 + 0 = 1 (a=1, b=0)
    e = np.array([0.0, 300.0, -120.0])
    n = np.array([0.0, 40.0, 10.0])
    f = np.array([1.4])
    N = 64
    rows, cols, cell = oimage.cell_map(e, n, f, N)
    lam = oimage.C_LIGHT / 1.4e9
    # a=1, b=0: u = -(E1 - E0)/lam
    u = -(300.0 / lam)
    v = -(40.0 / lam)
    assert cols[1, 0] == int(np.rint(u / cell)) + N // 2
    assert rows[1, 0] == int(np.rint(v / cell)) + N // 2
    # autos are dropped
    for a in range(3):
        assert rows[a * (a + 1) // 2 + a, 0] == -1
    # auto-fit: the longest |u| or |v| sits at N/2 cells from the centre
    assert np.isclose(cell, max(abs(300.0 + 120.0), 300.0, 120.0, 40, 30, 10) / lam * 2 / N)


def test_cell_map_drops_out_of_grid():
    e = np.array([0.0, 1000.0, 10.0])
    n = np.zeros(3)
    rows, cols, cell = oimage.cell_map(e, n, np.array([1.4]), 16, cell_lambda=1.0)
    assert np.all(rows[[1, 3, 4]] == -1)            # every cross baseline beyond +-8 cells


def test_grid_plane_is_linear():
    rng = np.random.default_rng(3)
    rows = rng.integers(0, 8, size=(10, 3))
    cols = rng.integers(0, 8, size=(10, 3))
    rows[2, 1] = cols[2, 1] = -1
    v1 = rng.standard_normal((10, 3)) + 1j * rng.standard_normal((10, 3))
    v2 = rng.standard_normal((10, 3)) + 1j * rng.standard_normal((10, 3))
    g = lambda v: oimage.grid_plane(v, rows, cols, 8)
    np.testing.assert_allclose(g(2 * v1 - 3j * v2), 2 * g(v1) - 3j * g(v2), atol=1e-12)
    assert np.isclose(g(v1).sum(), v1[rows >= 0].sum())


def test_sparse_twin_equals_add_at_form():
    """The sparse-product / scipy.fft twin used for 96-antenna oracle streams
    equals the np.add.at + numpy ifft2 images (integer visibilities, several
    baselines per cell, dropped cells, a pixel subset)."""
    rng = np.random.default_rng(9)
    n_ant, n_chan, n_sub, N, steps = 9, 8, 2, 16, 3
    e = rng.uniform(-300, 300, n_ant)
    n = rng.uniform(-80, 80, n_ant)
    freqs = np.linspace(1.49, 1.40, n_chan)
    rows, cols, _ = oimage.cell_map(e, n, freqs, N, cell_lambda=None)
    rows[3, 2] = cols[3, 2] = -1
    nb = rows.shape[0]
    vis = rng.integers(-50, 50, size=(steps, n_chan, nb, 2)).astype(np.int64)
    mask = oimage.edge_mask(N, 1)
    want = oimage.subband_images(vis, rows, cols, N, n_sub, mask).reshape(n_sub, steps, N * N)
    got = oimage.subband_images_sparse(vis, rows, cols, N, n_sub, mask)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())
    pix = np.array([0, 17, 100, 255])
    np.testing.assert_allclose(oimage.subband_images_sparse(vis, rows, cols, N, n_sub, mask, pix=pix),
                               want[:, :, pix], rtol=0, atol=1e-12 * np.abs(want).max())
