"""Geometric gridder and textbook imager oracle (test infrastructure only).

Pinned by (``tests/test_oracle_image.py``): the DC-only grid imaging to a
uniform ``1/N^2`` image and the single-side identity
(``tests/test_image_imager.py:132-175`` in the reference), the point-source
known answer ``2.0 * 240 / 32^2`` at pixel (9, 21)
(``tests/x2/test_search2x_ref.py:97-104``), the mask geometry
(``tests/test_image_imager.py:83-130``), and gridder linearity / autos
excluded / F20 cell position closed form.
"""
from __future__ import annotations

import math

import numpy as np

C_LIGHT = 299_792_458.0


def cell_map(antpos_e, antpos_n, freqs_GHz, n_grid, cell_lambda=None):
    """``(row, col, cell)``: nearest cell per (bls, chan), -1 for dropped.

    Restates ``build_pattern`` / ``FastVisGridder.from_pattern``
    (sparsity_pattern.py:1058-1191, grid/kernel.py:373-391): cross baselines
    only, ``u = -(E_a - E_b)/lambda``, auto-fit cell, ``rint`` + N/2, in-grid.
    """
    e = np.asarray(antpos_e)
    n = np.asarray(antpos_n)
    na = e.shape[0]
    a_l, b_l = [], []
    for a in range(na):
        for b in range(a + 1):
            a_l.append(a)
            b_l.append(b)
    a_idx = np.asarray(a_l)
    b_idx = np.asarray(b_l)
    keep = a_idx != b_idx
    du = e[a_idx[keep]].astype(np.float64) - e[b_idx[keep]].astype(np.float64)
    dv = n[a_idx[keep]].astype(np.float64) - n[b_idx[keep]].astype(np.float64)
    lam = C_LIGHT / (np.asarray(freqs_GHz, dtype=np.float64) * 1e9)
    u = -(du[:, None] / lam[None, :])
    v = -(dv[:, None] / lam[None, :])
    if cell_lambda is None:
        cell_lambda = float(np.max(np.maximum(np.abs(u), np.abs(v)))) * 2.0 / n_grid
    half = n_grid // 2
    col = np.rint(u / cell_lambda).astype(np.int64) + half
    row = np.rint(v / cell_lambda).astype(np.int64) + half
    ok = (row >= 0) & (row < n_grid) & (col >= 0) & (col < n_grid)
    nb = a_idx.size
    R = np.full((nb, lam.size), -1, dtype=np.int64)
    Cc = np.full((nb, lam.size), -1, dtype=np.int64)
    R[keep] = np.where(ok, row, -1)
    Cc[keep] = np.where(ok, col, -1)
    return R, Cc, float(cell_lambda)


def grid_plane(vis: np.ndarray, rows: np.ndarray, cols: np.ndarray, n_grid: int) -> np.ndarray:
    """``np.add.at`` of every ``(bls, chan)`` value into its full-plane cell
    (corr2x_ref.py:138-166).  ``vis`` ``(..., nbase, n_chan)`` complex;
    ``rows``/``cols`` ``(nbase, n_chan)``.  Returns ``(..., N, N)`` complex128."""
    vis = np.asarray(vis)
    lead = vis.shape[:-2]
    flat = vis.reshape(-1, vis.shape[-2] * vis.shape[-1]).astype(np.complex128)
    ok = (rows.reshape(-1) >= 0)
    lin = (rows.reshape(-1) * n_grid + cols.reshape(-1))[ok]
    G = np.zeros((flat.shape[0], n_grid * n_grid), dtype=np.complex128)
    for i in range(flat.shape[0]):
        np.add.at(G[i], lin, flat[i, ok])
    return G.reshape(lead + (n_grid, n_grid))


def mask_npad(n_grid: int, kernel_support: int = 1) -> int:
    """``max(K//2 + 2, envelope term)`` with no envelope (imager.py:57-109)."""
    return (kernel_support // 2) + 2


def edge_mask(n_grid: int, kernel_support: int = 1) -> np.ndarray:
    """Signed multiplicative mask: interior 1, outer ``npad`` ring and the DC
    pixel 0, times ``(-1)^(i+j)`` (imager.py:112-175)."""
    npad = mask_npad(n_grid, kernel_support)
    m = np.zeros((n_grid, n_grid), dtype=np.float64)
    if 2 * npad < n_grid:
        m[npad:n_grid - npad, npad:n_grid - npad] = 1.0
    dc = n_grid // 2
    m[dc, dc] = 0.0
    ii, jj = np.indices((n_grid, n_grid))
    return m * (1.0 - 2.0 * ((ii + jj) & 1))


def image(G: np.ndarray, mask: np.ndarray, scale: float = 1.0) -> np.ndarray:
    """``fftshift(Re(ifft2(G))) * mask * scale`` on the last two axes
    (search2x_ref.py:167-170, imager.py:232-245)."""
    img = np.fft.fftshift(np.fft.ifft2(G, axes=(-2, -1)).real, axes=(-2, -1))
    return img * mask * float(scale)


def subband_images(vis_int: np.ndarray, rows: np.ndarray, cols: np.ndarray, n_grid: int,
                   n_sub: int, mask: np.ndarray) -> np.ndarray:
    """Images ``(n_sub, n_steps, N, N)`` float64 from exact visibilities
    ``vis_int`` ``(n_steps, n_chan_out, nbase, 2)``; subband ``s`` grids its
    ``n_chan_out / n_sub`` output channels into one plane."""
    n_steps, nco, nb, _ = vis_int.shape
    per = nco // n_sub
    v = vis_int[..., 0].astype(np.float64) + 1j * vis_int[..., 1].astype(np.float64)
    out = np.zeros((n_sub, n_steps, n_grid, n_grid))
    for s in range(n_sub):
        sl = slice(s * per, (s + 1) * per)
        vs = np.transpose(v[:, sl, :], (0, 2, 1))            # (steps, nbase, per)
        G = grid_plane(vs, rows[:, sl], cols[:, sl], n_grid)
        out[s] = image(G, mask)
    return out


def subband_grids(vis: np.ndarray, rows: np.ndarray, cols: np.ndarray, n_grid: int, n_sub: int) -> np.ndarray:
    """Complex uv grids ``(n_sub, n_steps, N, N)`` of visibilities
    ``(n_steps, n_chan_out, nbase, 2)`` (``np.add.at`` per subband)."""
    n_steps, nco, nb, _ = vis.shape
    per = nco // n_sub
    v = vis[..., 0].astype(np.float64) + 1j * vis[..., 1].astype(np.float64)
    out = np.zeros((n_sub, n_steps, n_grid, n_grid), dtype=np.complex128)
    for s in range(n_sub):
        sl = slice(s * per, (s + 1) * per)
        out[s] = grid_plane(np.transpose(v[:, sl, :], (0, 2, 1)), rows[:, sl], cols[:, sl], n_grid)
    return out


def subband_images_sparse(vis_int: np.ndarray, rows: np.ndarray, cols: np.ndarray, n_grid: int,
                          n_sub: int, mask: np.ndarray, pix=None) -> np.ndarray:
    """Same images as :func:`subband_images`, ``(n_sub, n_steps, n_pix)``
    (optionally only the flat pixels ``pix``), gridded by one sparse product
    per subband instead of ``np.add.at`` -- the integer cell sums are exact
    in float64 whatever the order -- and imaged by ``scipy.fft.ifft2`` on
    all cores.  Equality with :func:`subband_images` is checked in
    ``tests/test_oracle_image.py``; this twin only makes 96-antenna oracle
    streams fast enough for the GPU parity tests."""
    import scipy.fft
    import scipy.sparse as sp
    n_steps, nco, nb, _ = vis_int.shape
    per = nco // n_sub
    NN = n_grid * n_grid
    sel = np.arange(NN) if pix is None else np.asarray(pix, dtype=np.int64)
    out = np.zeros((n_sub, n_steps, sel.size))
    m = mask.reshape(-1)[sel]
    for s in range(n_sub):
        sl = slice(s * per, (s + 1) * per)
        r = rows[:, sl].reshape(-1)
        c = cols[:, sl].reshape(-1)
        ok = r >= 0
        src = np.arange(r.size)[ok]
        S = sp.csr_matrix((np.ones(src.size), (src, (r * n_grid + c)[ok])), shape=(r.size, NN))
        v = vis_int[:, sl]                                            # (steps, per, nbase, 2)
        vr = np.transpose(v[..., 0], (0, 2, 1)).reshape(n_steps, -1).astype(np.float64)
        vi = np.transpose(v[..., 1], (0, 2, 1)).reshape(n_steps, -1).astype(np.float64)
        G = (S.T @ vr.T).T + 1j * (S.T @ vi.T).T                      # (steps, NN)
        img = scipy.fft.ifft2(G.reshape(n_steps, n_grid, n_grid), axes=(-2, -1), workers=-1).real
        img = np.fft.fftshift(img, axes=(-2, -1)).reshape(n_steps, NN)
        out[s] = img[:, sel] * m
    return out
