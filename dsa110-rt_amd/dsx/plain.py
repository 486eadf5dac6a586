"""Plain FP32 form of the node pipeline, timed beside :class:`SearchNode`.

The reference's algorithm stage by stage on the GPU with library kernels and
no fusion or batching across DM trials:

1. correlator: complex64 batched GEMM (rocBLAS) of the unpacked 4-bit samples,
   then channel and polarisation sums and the upper-triangle extraction
   (``services/corr_fast_kernel.py:667-755``); exact, since every partial sum
   is an integer below 2^24;
2. gridder: ``index_add_`` of every (baseline, channel) into its dense uv cell
   (``grid/kernel.py:590-608``);
3. imager: ``torch.fft.ifft2`` (rocFFT) + ``fftshift`` + signed edge mask
   (``image/imager.py:232-324``), float32 images;
4. dedispersion: one shift-and-add pass per DM trial over float32 subband
   images (``fine_dm/combiner.py:625-653``), float32 prefix-sum boxcars per
   width (``detector/forward.py:456-504``), chunk maxima and the float64
   running statistics, then the same host decode as the fused path.

Candidate semantics are those of :class:`dsx.pipeline.SearchNode` (detection
window ``[T0 - 64, T0 + L - 64)``, validity gate, burn-in, chunk maxima, ties
to the lowest width then the earliest centre).
"""
from __future__ import annotations

from typing import List, Optional

import numpy as np
import torch

from .config import DETECT_CHUNK_STEPS, DETECT_DELAY_STEPS, NodeConfig
from .corr import unpack_int4_split
from .detector import REC_DTYPE, decode_chunk_maxima
from .imager import compute_edge_mask
from .pattern import build_cell_map, upper_tri_indices
from .pipeline import BlockResult


class PlainNode:
    """Stateful plain-form node; history and statistics stay on the GPU."""

    def __init__(self, cfg: NodeConfig, antpos_e: np.ndarray, antpos_n: np.ndarray, *,
                 device: str = "cuda", max_block_steps: Optional[int] = None, corr_chunk_steps: int = 32,
                 grid_chunk_steps: int = 128):
        cfg.validate()
        if not torch.cuda.is_available():
            raise RuntimeError("PlainNode needs a ROCm GPU")
        self.cfg = cfg
        self.dev = torch.device(device)
        N = cfg.n_grid
        self.n_pix = N * N
        cmap = build_cell_map(antpos_e, antpos_n, cfg.grid_freq_out_GHz, N, cell_lambda=cfg.cell_lambda)
        lin = np.where(cmap.row >= 0, cmap.row.astype(np.int64) * N + cmap.col, -1)   # (nbase, n_chan_out)
        self.lin = [torch.from_numpy(np.ascontiguousarray(lin[:, s * cfg.per_sub:(s + 1) * cfg.per_sub]))
                    .to(self.dev) for s in range(cfg.n_sub)]
        mask = compute_edge_mask(n_grid=N, kernel_support=cfg.kernel_support, dtype=np.float32)
        self.mask = torch.from_numpy(mask).to(self.dev)
        a, b = upper_tri_indices(cfg.n_ant)
        self.ia = torch.from_numpy(np.asarray(a, dtype=np.int64)).to(self.dev)
        self.ib = torch.from_numpy(np.asarray(b, dtype=np.int64)).to(self.dev)
        sh = cfg.shift_table()                                   # [d][s]
        self.shifts = sh
        self.valid_from = sh.max(axis=1).astype(np.int64)
        self.h2 = cfg.w_max // 2
        self.hist = DETECT_DELAY_STEPS + self.h2 + int(sh.max()) + 1
        self.cap = int(max_block_steps or cfg.n_steps)
        self.ring = torch.zeros((cfg.n_sub, self.hist + self.cap, self.n_pix), dtype=torch.float32,
                                device=self.dev)
        self.mean = torch.zeros((cfg.n_dm, self.n_pix), dtype=torch.float64, device=self.dev)
        self.m2 = torch.zeros((cfg.n_dm, self.n_pix), dtype=torch.float64, device=self.dev)
        self.n_prev = np.zeros(cfg.n_dm, dtype=np.float64)
        self.corr_chunk = int(corr_chunk_steps)
        self.grid_chunk = int(grid_chunk_steps)
        self.T0 = 0
        self.prev_len = 0

    # ------------------------------------------------------------------ stages
    def correlate(self, raw: torch.Tensor, n_samp: int) -> torch.Tensor:
        """``(n_steps, n_chan_out, nbase)`` complex64 Stokes-I visibilities,
        ``vis[bls] = sum conj(E_b) E_a``."""
        cfg = self.cfg
        n_pkt = n_samp // 2
        n_steps = n_samp // cfg.t_int
        re, im = unpack_int4_split(raw, device=self.dev, scale=1.0, out_dtype=torch.float32,
                                   shape=(n_pkt, cfg.n_ant, cfg.n_chan))     # (ch, 2t, 2p, pkt, ant)
        E = torch.complex(re, im).permute(0, 2, 3, 1, 4)                       # (ch, p, pkt, 2t, ant)
        E = E.reshape(cfg.n_chan, 2, n_steps, cfg.t_int, cfg.n_ant)          # samples of a step contiguous
        out = torch.empty((n_steps, cfg.n_chan_out, cfg.nbase), dtype=torch.complex64, device=self.dev)
        for t0 in range(0, n_steps, self.corr_chunk):
            t1 = min(n_steps, t0 + self.corr_chunk)
            X = E[:, :, t0:t1].permute(2, 0, 1, 3, 4).contiguous()           # (S, ch, p, t_int, ant)
            V = torch.matmul(X.conj().transpose(-1, -2).contiguous(), X)     # (S, ch, p, b, a)
            V = V.sum(dim=2)                                                 # Stokes I
            V = V.reshape(t1 - t0, cfg.n_chan_out, cfg.csf, cfg.n_ant, cfg.n_ant).sum(dim=2)
            out[t0:t1] = V[..., self.ib, self.ia]
        return out

    def image(self, vis: torch.Tensor, n_steps: int) -> None:
        """Grid, FFT and mask each (subband, step) into the history planes."""
        cfg = self.cfg
        N = cfg.n_grid
        base = self.hist
        for t0 in range(0, n_steps, self.grid_chunk):
            t1 = min(n_steps, t0 + self.grid_chunk)
            S = t1 - t0
            for s in range(cfg.n_sub):
                v = vis[t0:t1, s * cfg.per_sub:(s + 1) * cfg.per_sub, :].transpose(1, 2)   # (S, nbase, per)
                lin = self.lin[s].reshape(-1)
                ok = lin >= 0
                src = torch.view_as_real(v.reshape(S, -1)[:, ok].contiguous())             # (S, K, 2)
                G = torch.zeros((S, N * N, 2), dtype=torch.float32, device=self.dev)
                G.index_add_(1, lin[ok], src)
                img = torch.fft.ifft2(torch.view_as_complex(G).reshape(S, N, N)).real
                img = torch.fft.fftshift(img, dim=(-2, -1)) * self.mask
                self.ring[s, base + t0:base + t1] = img.reshape(S, -1)

    def push(self, n_steps: int) -> None:
        """Keep the last ``hist`` steps as history before the next block."""
        keep = self.hist
        end = self.hist + n_steps
        self.ring[:, :keep] = self.ring[:, end - keep:end].clone()

    def detect(self, n_steps: int) -> BlockResult:
        cfg = self.cfg
        T0 = self.T0
        c_lo = T0 - DETECT_DELAY_STEPS
        L = n_steps
        h2 = self.h2
        chunk = DETECT_CHUNK_STEPS
        n_chunks = L // chunk
        widths = cfg.widths
        # ring position of absolute step tau: hist + tau - T0
        t_abs = torch.arange(c_lo - h2, c_lo + L + h2, device=self.dev)           # D support
        tc = c_lo + np.arange(L)
        recs: List[np.ndarray] = []
        blk_sum = torch.zeros((cfg.n_dm, self.n_pix), dtype=torch.float64, device=self.dev)
        blk_sq = torch.zeros_like(blk_sum)
        counts = np.zeros(cfg.n_dm)
        for d in range(cfg.n_dm):
            D = torch.zeros((L + 2 * h2, self.n_pix), dtype=torch.float32, device=self.dev)
            for s in range(cfg.n_sub):
                lo = self.hist + (c_lo - h2 - int(self.shifts[d, s])) - T0
                D += self.ring[s, lo:lo + L + 2 * h2]
            D[t_abs < 0] = 0.0                                                      # before stream start
            vf = int(self.valid_from[d])
            okc = tc >= vf
            counts[d] = float(okc.sum())
            if counts[d] > 0:
                first = int(np.argmax(okc))
                cen = D[h2 + first:h2 + L].double()
                blk_sum[d] = cen.sum(dim=0)
                blk_sq[d] = (cen * cen).sum(dim=0)
            n_a = self.n_prev[d]
            if n_a <= 0 or n_a < cfg.min_stat_count:
                continue
            sig = torch.sqrt(self.m2[d] / n_a)
            mu = self.mean[d]
            P = torch.zeros((L + 2 * h2 + 1, self.n_pix), dtype=torch.float32, device=self.dev)
            P[1:] = torch.cumsum(D, dim=0)
            best = torch.full((n_chunks, self.n_pix), -np.inf, dtype=torch.float32, device=self.dev)
            best_t = torch.zeros((n_chunks, self.n_pix), dtype=torch.int64, device=self.dev)
            best_k = torch.zeros((n_chunks, self.n_pix), dtype=torch.int64, device=self.dev)
            for k, w in enumerate(widths):
                lo = h2 - w // 2                                                    # centre c -> D index c + h2
                B = (P[lo + w:lo + w + L] - P[lo:lo + L]).reshape(n_chunks, chunk, self.n_pix)
                mb, mt = B.max(dim=1)
                snr = ((mb.double() - w * mu) / (sig * np.sqrt(w))).float()
                better = snr > best
                best = torch.where(better, snr, best)
                best_t = torch.where(better, mt, best_t)
                best_k = torch.where(better, torch.full_like(best_k, k), best_k)
            ch0 = c_lo + np.arange(n_chunks) * chunk
            scored = torch.from_numpy((ch0 - h2 >= vf)).to(self.dev)[:, None] & (sig > 0)[None, :]
            hit = scored & (best > cfg.threshold)
            if bool(hit.any()):
                ci, pi = torch.nonzero(hit, as_tuple=True)
                r = np.zeros(int(ci.numel()), dtype=REC_DTYPE)
                r["t"] = (c_lo + ci * chunk + best_t[ci, pi]).cpu().numpy()
                r["p"] = pi.cpu().numpy()
                r["d"] = d
                r["k"] = best_k[ci, pi].cpu().numpy()
                r["snr"] = best[ci, pi].cpu().numpy()
                recs.append(r)
        # merge this block's statistics (Chan), float64
        nb = torch.from_numpy(counts).to(self.dev)[:, None]
        n_a = torch.from_numpy(self.n_prev).to(self.dev)[:, None]
        upd = nb > 0
        safe_b = torch.where(upd, nb, torch.ones_like(nb))
        mean_b = blk_sum / safe_b
        m2_b = blk_sq - blk_sum * blk_sum / safe_b
        n = n_a + nb
        safe_n = torch.where(upd, n, torch.ones_like(n))
        delta = mean_b - self.mean
        self.mean = torch.where(upd, self.mean + delta * (nb / safe_n), self.mean)
        self.m2 = torch.where(upd, self.m2 + m2_b + delta * delta * (n_a * nb / safe_n), self.m2)
        self.n_prev += counts
        lam = cfg.stats_decay(n_steps)                 # forgetting of the running statistics
        if lam < 1.0:
            self.m2 = self.m2 * lam
            self.n_prev *= lam
        rec = np.concatenate(recs) if recs else np.zeros(0, REC_DTYPE)
        cands = decode_chunk_maxima(rec, widths=widths, n_grid=cfg.n_grid, dm_trials=cfg.dm_trials,
                                    threshold=cfg.threshold, n_top=cfg.n_top, radius_t=cfg.radius_t,
                                    radius_dm=cfg.radius_dm, radius_lm=cfg.radius_lm)
        self.T0 = T0 + n_steps
        return BlockResult(candidates=cands, n_raw=int(rec.size), overflow=False, T0=T0, n_steps=n_steps,
                           records=rec)

    def process_block(self, raw: torch.Tensor, n_samp: Optional[int] = None) -> BlockResult:
        cfg = self.cfg
        n_samp = cfg.n_samp if n_samp is None else int(n_samp)
        n_steps = n_samp // cfg.t_int
        if n_samp % (2 * cfg.t_int) or n_steps % DETECT_CHUNK_STEPS or n_steps > self.cap:
            raise ValueError(f"block of {n_samp} samples: steps must be a multiple of "
                             f"{DETECT_CHUNK_STEPS} and <= {self.cap}")
        if self.T0 > 0:
            self.push(self.prev_len)
        vis = self.correlate(raw, n_samp)
        self.image(vis, n_steps)
        self.prev_len = n_steps
        return self.detect(n_steps)
