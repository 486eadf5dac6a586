"""Block-by-block oracle of the whole fast path (test infrastructure only).

Each block: int64 correlation (``oracle.corr``), np.add.at gridding +
``fftshift(Re(ifft2))`` imaging (``oracle.image``), a float64 image history,
dedispersion + detection against the running per-(DM, pixel) statistics from
previous blocks (``oracle.search``), then the Chan merge of this block's
statistics.  A DM is scored only once it holds ``min_stat_count`` complete
samples of statistics from earlier blocks; before that its blocks only prime
the statistics, like the reference Layer-2 burn-in
(``src/dsart/noise_norm/layer2.py:313-452``).

``cfg`` is duck-typed: ``n_ant, chan_freqs_GHz, t_int, csf, n_sub, n_grid,
grid_freq_GHz, cell_lambda, dm_max, n_dm, widths, threshold, n_top,
radius_t, radius_dm, radius_lm``.
"""
from __future__ import annotations

from typing import List, Optional, Sequence

import numpy as np

from . import corr as ocorr
from . import image as oimage
from . import search as osearch
from . import stage1 as os1

DELAY = 64
CHUNK = 64


class OracleStream:
    def __init__(self, cfg, antpos_e, antpos_n, *, pix: Optional[np.ndarray] = None):
        self.cfg = cfg
        f = np.asarray(cfg.chan_freqs_GHz, dtype=np.float64)
        nco = f.size // cfg.csf
        self.freq_out = f.reshape(nco, cfg.csf).mean(axis=1) if cfg.csf > 1 else f
        gf = self.freq_out if cfg.grid_freq_GHz is None else np.full(nco, float(cfg.grid_freq_GHz))
        self.rows, self.cols, self.cell_lambda = oimage.cell_map(
            antpos_e, antpos_n, gf, cfg.n_grid, cfg.cell_lambda)
        per = nco // cfg.n_sub
        nu_ref = self.freq_out[::per][: cfg.n_sub]
        self.s1_dm = getattr(cfg, "stage1_dm", None)
        if self.s1_dm is not None:
            # stage 1 aligns each subband to its first channel group
            nu_ref = os1.subband_refs(f, n_sub=cfg.n_sub, group=cfg.stage1_csf)
            self.s1_bins = os1.stage1_bins(f, n_sub=cfg.n_sub, group=cfg.stage1_csf, t_int=cfg.t_int,
                                           dm=self.s1_dm)
            self.s1_delay = max(int(getattr(cfg, "stage1_delay_steps", 0)), int(self.s1_bins.max()))
            self.s1_tail = np.zeros((self.s1_delay * cfg.t_int // 2, cfg.n_ant, f.size, 2, 2), np.uint8)
        W = int(getattr(cfg, "static_sky_blocks", 0))
        self.sky = os1.StaticSkyMean(W, int(cfg.static_sky_warmup)) if W > 0 else None
        t_step = cfg.t_int * osearch.NATIVE_US
        # DM trial values are configuration (an explicit list when the config
        # carries one, e.g. one range of an uneven plan)
        self.dm = (np.asarray(cfg.dm_trials, dtype=np.float64) if hasattr(cfg, "dm_trials")
                   else np.linspace(0.0, float(cfg.dm_max), int(cfg.n_dm)))
        self.shifts = osearch.shift_table(self.dm, nu_ref, float(f[-1]), t_step)
        self.mask = oimage.edge_mask(cfg.n_grid, 1)
        n_pix_full = cfg.n_grid * cfg.n_grid
        self.pix = np.arange(n_pix_full) if pix is None else np.asarray(pix, dtype=np.int64)
        npx = self.pix.size
        self.widths = list(cfg.widths)
        self.half = max(self.widths) // 2
        self.st_n = np.zeros(cfg.n_dm)
        self.st_mean = np.zeros((cfg.n_dm, npx))
        self.st_m2 = np.zeros((cfg.n_dm, npx))
        self.hist = np.zeros((cfg.n_sub, npx, 0))        # time-contiguous per pixel
        self.hist_tau0 = 0
        self.t_next = 0                  # first step of the next block
        self.keep = int(self.shifts.max()) + self.half + DELAY + 1

    # -- stages ---------------------------------------------------------------
    def visibilities(self, raw: np.ndarray, n_samp: int) -> np.ndarray:
        c = self.cfg
        return ocorr.correlate(raw, n_pkt=n_samp // 2, n_ant=c.n_ant, n_chan=len(c.chan_freqs_GHz),
                               t_int=c.t_int, csf=c.csf, fuse_pol=True)

    def images(self, vis: np.ndarray) -> np.ndarray:
        """``(n_sub, n_steps, n_pix_sel)`` float64 (the sparse-product twin of
        ``oracle.image.subband_images``, equal to it on integer visibilities).
        With static-sky subtraction the per-cell uv grids pass through
        :class:`oracle.stage1.StaticSkyMean` before imaging."""
        c = self.cfg
        if self.sky is None:
            return oimage.subband_images_sparse(vis, self.rows, self.cols, c.n_grid, c.n_sub, self.mask,
                                                pix=self.pix)
        G = oimage.subband_grids(vis, self.rows, self.cols, c.n_grid, c.n_sub)     # (n_sub, steps, N, N)
        G = np.transpose(self.sky.apply(np.transpose(G, (1, 0, 2, 3))), (1, 0, 2, 3))
        img = oimage.image(G, self.mask)
        return img.reshape(c.n_sub, G.shape[1], -1)[:, :, self.pix]

    def stage1_visibilities(self, raw: np.ndarray, n_samp: int) -> np.ndarray:
        """The stage-1 stream's visibilities of one raw block (steps
        ``[T0 - delay, T0 + L - delay)``); keeps the block's last ``delay`` steps."""
        c = self.cfg
        nch = len(c.chan_freqs_GHz)
        raw = np.asarray(raw, np.uint8).reshape(n_samp // 2, c.n_ant, nch, 2, 2)
        v = os1.stage1_visibilities(self.s1_tail, raw, n_ant=c.n_ant, n_chan=nch, t_int=c.t_int, csf=c.csf,
                                    bins=self.s1_bins, delay=self.s1_delay)
        n_tail = self.s1_tail.shape[0]
        if n_tail:
            self.s1_tail = raw[raw.shape[0] - n_tail:].copy()
        return v

    def start_at(self, t0: int) -> None:
        """Start the stream at step ``t0`` with zeros before it (as a node
        whose history buffers start zeroed)."""
        self.t_next = int(t0)
        self.hist = np.zeros((self.cfg.n_sub, self.pix.size, self.keep))
        self.hist_tau0 = int(t0) - self.keep

    def push_images(self, img: np.ndarray):
        """Append a block of images ``(n_sub, L, n_pix_sel)``; run detection;
        merge statistics."""
        c = self.cfg
        L = img.shape[1]
        T0 = self.t_next
        keep_old = min(self.hist.shape[2], self.keep)
        new = np.empty((c.n_sub, self.pix.size, keep_old + L))
        new[:, :, :keep_old] = self.hist[:, :, self.hist.shape[2] - keep_old:]
        new[:, :, keep_old:] = np.transpose(img, (0, 2, 1))
        self.hist_tau0 += self.hist.shape[2] - keep_old
        self.hist = new
        self.t_next = T0 + L
        c_lo = T0 - DELAY
        bs, bt, bk, s1, s2, cnt = osearch.search_block_c_tmajor(
            self.hist, self.hist_tau0, self.shifts, c_lo, L, CHUNK, self.widths,
            self.st_n, self.st_mean, self.st_m2, getattr(c, "min_stat_count", 512))
        self.last = dict(best_snr=bs, best_t=bt, best_k=bk, blk_sum=s1, blk_sumsq=s2, count=cnt,
                         c_lo=c_lo)
        cands = osearch.decode(bs, bt, bk, self.widths, c.n_grid if self.pix.size == c.n_grid ** 2 else 1,
                               threshold=c.threshold, n_top=c.n_top, radius_t=c.radius_t,
                               radius_dm=c.radius_dm, radius_lm=c.radius_lm) \
            if self.pix.size == c.n_grid ** 2 else []
        self.st_n, self.st_mean, self.st_m2 = osearch.merge_stats(
            self.st_n, self.st_mean, self.st_m2, cnt, s1, s2)
        tau = getattr(c, "stats_tau_s", None)
        if tau is not None:
            self.st_n, self.st_m2 = osearch.decay_stats(
                self.st_n, self.st_m2, osearch.block_decay(L, c.t_int, tau))
        return cands

    def push_raw(self, raw: np.ndarray, n_samp: int):
        if self.s1_dm is not None:
            vis = self.stage1_visibilities(raw, n_samp)
        else:
            vis = self.visibilities_fast(raw, n_samp)
        return self.push_images(self.images(vis))

    def visibilities_fast(self, raw: np.ndarray, n_samp: int) -> np.ndarray:
        """:meth:`visibilities` through the float32-product twin (equal integers)."""
        c = self.cfg
        return ocorr.correlate_blas(raw, n_pkt=n_samp // 2, n_ant=c.n_ant, n_chan=len(c.chan_freqs_GHz),
                                    t_int=c.t_int, csf=c.csf)
