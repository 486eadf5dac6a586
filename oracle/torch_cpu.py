"""PyTorch-CPU restatement of the reference's CPU path for one node block
(test infrastructure and the CPU baseline of the benchmark only).

Stages, each following the reference's own torch CPU code:

* ``correlate`` -- 4-bit unpack by lookup, then the four real GEMMs
  ``V_re = R^T R + I^T I``, ``V_im = I^T R - R^T I`` per (step, output
  channel) with ``torch.matmul`` (``src/dsart/services/corr_fast_kernel.py:667-755``),
  upper-triangle gather ``bls = a(a+1)/2 + b``;
* ``grid_images`` -- ``index_add_`` of each subband's visibilities into its
  dense uv plane (``src/dsart/grid/kernel.py:590-608``), then
  ``torch.fft.ifft2`` + ``fftshift`` + real part times the signed edge mask
  (``src/dsart/image/imager.py:242-245,304-324``);
* ``search_block`` -- per DM trial the shift-and-add of the subband images
  (``src/dsart/fine_dm/combiner.py:625-653``), cumsum boxcars
  (``src/dsart/detector/forward.py:456-504``), per-chunk maxima against the
  float64 running statistics, and the top-k + merge-radius decode
  (``src/dsart/detector/decoder.py:436-533``) -- the semantics of the fused
  GPU kernel and of ``oracle.search``.

Pinned against the numpy / C oracle in ``tests/test_oracle_torch_cpu.py``.
"""
from __future__ import annotations

from typing import Sequence, Tuple

import numpy as np
import torch

from . import corr as ocorr
from . import image as oimage


def _luts():
    b = torch.arange(256)
    return ((((b & 0xF) ^ 8) - 8).to(torch.float32), ((((b >> 4) & 0xF) ^ 8) - 8).to(torch.float32))


def correlate(raw: np.ndarray, *, n_pkt: int, n_ant: int, n_chan: int, t_int: int, csf: int) -> torch.Tensor:
    """``(n_steps, n_chan // csf, nbase, 2)`` float32 visibilities (exact integers)."""
    lre, lim = _luts()
    r = torch.from_numpy(np.ascontiguousarray(raw).reshape(-1)).long()
    n_steps = 2 * n_pkt // t_int
    nco = n_chan // csf
    K = t_int * csf * 2
    shp = (n_steps, t_int // 2, n_ant, nco, csf, 2, 2)
    R = lre[r].reshape(shp).permute(0, 3, 1, 5, 4, 6, 2).reshape(n_steps, nco, K, n_ant)
    I = lim[r].reshape(shp).permute(0, 3, 1, 5, 4, 6, 2).reshape(n_steps, nco, K, n_ant)
    Rt, It = R.transpose(-1, -2), I.transpose(-1, -2)
    vre = torch.matmul(Rt, R) + torch.matmul(It, I)
    vim = torch.matmul(It, R) - torch.matmul(Rt, I)
    a, b = ocorr.upper_tri(n_ant)
    a, b = torch.from_numpy(a), torch.from_numpy(b)
    return torch.stack([vre[..., a, b], vim[..., a, b]], dim=-1)


def grid_images(vis: torch.Tensor, rows: np.ndarray, cols: np.ndarray, n_grid: int, n_sub: int,
                mask: np.ndarray) -> torch.Tensor:
    """``(n_sub, n_steps, N*N)`` float32 masked dirty images."""
    n_steps, nco, nb, _ = vis.shape
    per = nco // n_sub
    NN = n_grid * n_grid
    m = torch.from_numpy(mask.astype(np.float32)).reshape(-1)
    out = torch.empty((n_sub, n_steps, NN), dtype=torch.float32)
    vc = torch.complex(vis[..., 0], vis[..., 1])                    # (steps, cout, nbase)
    for s in range(n_sub):
        r = torch.from_numpy(rows[:, s * per:(s + 1) * per].T.reshape(-1))
        c = torch.from_numpy(cols[:, s * per:(s + 1) * per].T.reshape(-1))
        ok = r >= 0
        lin = (r * n_grid + c)[ok]
        src = vc[:, s * per:(s + 1) * per, :].reshape(n_steps, -1)[:, ok]
        G = torch.zeros((n_steps, NN), dtype=torch.complex64)
        G.index_add_(1, lin, src)
        img = torch.fft.fftshift(torch.fft.ifft2(G.reshape(n_steps, n_grid, n_grid)), dim=(-2, -1)).real
        out[s] = img.reshape(n_steps, NN) * m
    return out


def search_block(hist: torch.Tensor, tau0: int, shifts: np.ndarray, c_lo: int, n_centres: int,
                 widths: Sequence[int], st_n: np.ndarray, st_mean: torch.Tensor, st_m2: torch.Tensor,
                 min_count: float, chunk: int = 64, dms: Sequence[int] = None) -> Tuple[np.ndarray, ...]:
    """One block of centres against the running statistics, over the float32
    history ``hist`` (n_sub, n_tau, n_pix) holding steps ``[tau0, tau0 + n_tau)``
    (zeros before the stream start).  Returns per DM of ``dms`` (default all)
    the chunk maxima ``(best_snr, best_t, best_k)`` ``(n_dm, n_chunks, n_pix)``
    and the block's float64 ``(sum, sum of squares, count)``."""
    n_sub, n_tau, n_pix = hist.shape
    half = max(widths) // 2
    n_dm = shifts.shape[0]
    dms = range(n_dm) if dms is None else dms
    n_chunks = n_centres // chunk
    bs = np.full((len(dms), n_chunks, n_pix), -np.inf)
    bt = np.full((len(dms), n_chunks, n_pix), -1, dtype=np.int64)
    bk = np.full((len(dms), n_chunks, n_pix), -1, dtype=np.int64)
    s1 = np.zeros((len(dms), n_pix))
    s2 = np.zeros((len(dms), n_pix))
    cnt = np.zeros(len(dms))
    t0 = c_lo - half
    nd = n_centres + 2 * half + 1
    for j, d in enumerate(dms):
        D = torch.zeros((nd, n_pix), dtype=torch.float32)
        for s in range(n_sub):
            lo = t0 - int(shifts[d, s]) - tau0                       # history row of D index 0
            a, b = max(lo, 0), min(lo + nd, n_tau)
            if b > a:
                D[a - lo:b - lo] += hist[s, a:b]
        t_abs = t0 + torch.arange(nd)
        D[t_abs < 0] = 0.0
        vf = int(shifts[d].max())
        tc = c_lo + np.arange(n_centres)
        ok = tc >= vf
        cnt[j] = ok.sum()
        cen = D[half:half + n_centres][torch.from_numpy(ok)].double()
        s1[j] = cen.sum(0).numpy()
        s2[j] = (cen * cen).sum(0).numpy()
        n_a = st_n[d]
        if n_a <= 0 or n_a < min_count:
            continue
        sig = torch.sqrt(st_m2[d] / n_a)
        mu = st_mean[d]
        P = torch.zeros((nd + 1, n_pix), dtype=torch.float64)
        P[1:] = torch.cumsum(D.double(), 0)
        best = torch.full((n_chunks, n_pix), -np.inf, dtype=torch.float64)
        bti = torch.zeros((n_chunks, n_pix), dtype=torch.int64)
        bki = torch.zeros((n_chunks, n_pix), dtype=torch.int64)
        for k, w in enumerate(widths):
            lo = half - w // 2
            B = (P[lo + w:lo + w + n_centres] - P[lo:lo + n_centres]).reshape(n_chunks, chunk, n_pix)
            mb, mt = B.max(dim=1)
            snr = (mb - w * mu) / (sig * np.sqrt(w))
            better = snr > best
            best = torch.where(better, snr, best)
            bti = torch.where(better, mt, bti)
            bki = torch.where(better, torch.full_like(bki, k), bki)
        ch0 = c_lo + np.arange(n_chunks) * chunk
        sc = torch.from_numpy(ch0 - half >= vf)[:, None] & (sig > 0)[None, :]
        bs[j] = torch.where(sc, best, torch.full_like(best, -np.inf)).numpy()
        bt[j] = (torch.from_numpy(ch0)[:, None] + bti).numpy()
        bk[j] = bki.numpy()
    return bs, bt, bk, s1, s2, cnt


def decode(best_snr: np.ndarray, best_t: np.ndarray, best_k: np.ndarray, n_grid: int, *, threshold: float,
           n_top: int, radius_t: int, radius_dm: int, radius_lm: int):
    """``torch.topk`` of the chunk maxima above threshold, emitted in
    (-snr, t, DM, pixel) order, then greedy merge-radius suppression;
    ``(dm, t, k, row, col, snr)`` in emission order."""
    d_i, c_i, p_i = np.nonzero(best_snr > threshold)
    if d_i.size == 0:
        return []
    snr = torch.from_numpy(best_snr[d_i, c_i, p_i])
    sel = torch.topk(snr, k=min(int(n_top), snr.numel())).indices.numpy()
    d_i, c_i, p_i = d_i[sel], c_i[sel], p_i[sel]
    t = best_t[d_i, c_i, p_i].astype(np.int64)
    v = best_snr[d_i, c_i, p_i]
    order = np.lexsort((p_i, d_i, t, -v))
    acc, out = [], []
    for o in order:
        tt, dd, pp = int(t[o]), int(d_i[o]), int(p_i[o])
        r, cc = divmod(pp, n_grid)
        if any(abs(tt - a) <= radius_t and abs(dd - b) <= radius_dm and abs(r - x) <= radius_lm and
               abs(cc - y) <= radius_lm for a, b, x, y in acc):
            continue
        acc.append((tt, dd, r, cc))
        out.append((dd, tt, int(best_k[d_i[o], c_i[o], p_i[o]]), r, cc, float(v[o])))
    return out
