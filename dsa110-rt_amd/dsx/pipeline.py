"""Single-GPU search node: raw 4-bit voltage blocks in, candidate lists out.

Per block (all on one HIP stream, no host round trip until the candidate
copy):

1. ``dsx_corr_i8``       raw bytes -> exact int32 visibilities (step, channel, bls)
2. ``dsx_image_subbands`` visibilities -> masked dirty images (subband, step, N, N)
3. ``dsx_ring_advance``   images -> time-major subband history (double buffered)
4. ``dsx_dedisp_detect``  history -> chunk-maximum candidates + float64 stats
5. host decode           top-k + merge radius -> ``List[Candidate]``

This is the caller the reference spreads over ``corr_fast_integration.
process_block`` (``src/dsart/services/corr_fast_integration.py:2063-2585``)
and ``CubePipeline.process`` (``src/dsart/services/cube_pipeline.py:2641-2728``),
collapsed onto one GPU with no transport in between.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np
import torch

from . import _lib
from .config import DETECT_CHUNK_STEPS, DETECT_DELAY_STEPS, NodeConfig
from .dedisp import plan_dedisp
from .detector import (CAND_DTYPE, N_LEVELS, REC_DTYPE, Candidate, decode_chunk_maxima,
                       device_records, record_levels)
from .imager import mask_npad
from .pattern import build_cell_map, build_imager_taps

#: LDS per dedispersion work-group: two 256-lane work-groups per CU.
LDS_BUDGET_DEDISP = 80 * 1024
LDS_BUDGET_ASYNC = 48 * 1024


@dataclass
class BlockResult:
    candidates: List[Candidate]
    n_raw: int                 # chunk maxima above threshold (before decode)
    overflow: bool
    T0: int
    n_steps: int
    records: Optional[np.ndarray] = None   # chunk-maximum records (REC_DTYPE, absolute int64 time)
    #: threshold the kept records are complete above: the node's threshold, or
    #: after an overflow the lowest record-count level whose records fit
    threshold_used: Optional[float] = None
    #: per DM range of an uneven plan: the threshold its records are complete above
    range_thresholds: Optional[List[float]] = None


class SearchNode:
    """Stateful node pipeline; state (history ring, statistics) stays on the GPU."""

    def __init__(self, cfg: NodeConfig, antpos_e: np.ndarray, antpos_n: np.ndarray, *,
                 device: str = "cuda", ring_dtype: torch.dtype = torch.float32,
                 max_block_steps: Optional[int] = None, cand_cap: int = 1 << 20,
                 stream: Optional[torch.cuda.Stream] = None, lds_budget: Optional[int] = None,
                 dedisp_mode: Optional[int] = None, n_slots: int = 1, with_front: bool = True,
                 with_detect: bool = True, share_front: Optional["SearchNode"] = None):
        cfg.validate()
        if ring_dtype not in (torch.float16, torch.float32):
            raise TypeError("ring_dtype must be float16 or float32")
        if not torch.cuda.is_available():
            raise RuntimeError("SearchNode needs a ROCm GPU")
        _lib.lib()
        self.cfg = cfg
        self.dev = torch.device(device)
        self.ring_dtype = ring_dtype
        self.stream = stream
        N = cfg.n_grid
        self.n_pix = N * N
        # ---- gridder taps
        self.cmap = build_cell_map(antpos_e, antpos_n, cfg.grid_freq_out_GHz, N,
                                   cell_lambda=cfg.cell_lambda)
        taps = build_imager_taps(self.cmap, cfg.n_sub)
        self.n_rows = taps.n_rows
        self.tap_src = torch.from_numpy(taps.src).to(self.dev)
        self.tap_dst = torch.from_numpy(taps.dst).to(self.dev)
        self.tap_off = torch.from_numpy(taps.offsets).to(self.dev)
        self.npad = mask_npad(N, cfg.kernel_support)
        # ---- DM plan
        sh = cfg.shift_table()                                   # [d][s]
        self.shifts_host = sh
        self.shift = torch.from_numpy(np.ascontiguousarray(sh.T)).to(self.dev)   # [s][d]
        self.valid_from_host = sh.max(axis=1).astype(np.int64)
        self.valid_from = torch.from_numpy(self.valid_from_host.astype(np.int32)).to(self.dev)
        self.shmax = sh.max(axis=0).astype(np.int32)             # per subband
        self.n_widths = len(cfg.widths)
        self.w_max = cfg.w_max
        h2 = self.w_max // 2
        # rounded up to 8 steps: 16-byte aligned history rows for the fp16 copy kernels
        self.hist = -(-(DETECT_DELAY_STEPS + h2 + int(sh.max()) + 1) // 8) * 8
        self.cap = int(max_block_steps or cfg.n_steps)
        # ---- buffers
        nb = cfg.nbase
        # per-slot block buffers: a block's front stage (correlate + image) can run
        # while the previous block's back stage (ring + dedispersion) still reads its slot
        if n_slots < 1:
            raise ValueError("n_slots must be >= 1")
        self.n_slots = int(n_slots)
        nf = self.n_slots if with_front else 0
        if share_front is not None:
            # nodes driven in turn on one stream (one per coarse DM) share the block buffers:
            # stream order puts each node's reads of a slot before the next node's writes
            if (share_front.cap < self.cap or share_front.n_slots < nf or share_front.ring_dtype != ring_dtype
                    or tuple(share_front.vis_slots[0].shape[1:]) != (cfg.n_chan_out, nb, 2)):
                raise ValueError("share_front: the other node's block buffers do not fit this node")
            self.vis_slots = share_front.vis_slots[:nf]
            self.img_slots = share_front.img_slots[:nf]
        else:
            self.vis_slots = [torch.empty((self.cap, cfg.n_chan_out, nb, 2), dtype=torch.int32, device=self.dev)
                              for _ in range(nf)]
            self.img_slots = [torch.empty(cfg.n_sub * self.cap * N * N, dtype=ring_dtype, device=self.dev)
                              for _ in range(nf)]
        self.vis = self.vis_slots[0] if nf else None
        row_len = self.hist + self.cap
        self.row_len = row_len
        nd = 2 if with_detect else 0
        self.rings = [torch.zeros((cfg.n_sub, self.n_pix, row_len), dtype=ring_dtype, device=self.dev)
                      for _ in range(nd)]
        self.cur = 0
        ns = self.n_pix if with_detect else 0
        # float64 noise state in n_slots + 1 rotating buffers: block k copies
        # buffer k into buffer k + 1 and merges its sums there in place, so the
        # state a block was detected against stays intact until block
        # k + n_slots + 1 -- long enough to detect a block whose records
        # overflowed again (finish())
        self.n_stat = self.n_slots + 1
        self.mean_bufs = [torch.zeros((ns, cfg.n_dm), dtype=torch.float64, device=self.dev)
                          for _ in range(self.n_stat)]
        self.m2_bufs = [torch.zeros((ns, cfg.n_dm), dtype=torch.float64, device=self.dev)
                        for _ in range(self.n_stat)]
        self.n_prev_bufs = [torch.zeros(cfg.n_dm, dtype=torch.float64, device=self.dev)
                            for _ in range(self.n_stat)]
        self.stat_cur = 0
        self.n_prev_host = np.zeros(cfg.n_dm, dtype=np.float64)
        self.blk_count = torch.zeros(cfg.n_dm, dtype=torch.float64, device=self.dev)
        self.cand_cap = int(cand_cap)
        self.cand_slots = [torch.empty((self.cand_cap if with_detect else 0, 4), dtype=torch.int32,
                                       device=self.dev) for _ in range(self.n_slots)]
        self.n_cand_slots = [torch.zeros(1, dtype=torch.int32, device=self.dev) for _ in range(self.n_slots)]
        self.cand, self.n_cand = self.cand_slots[0], self.n_cand_slots[0]
        # which block last wrote each history / state buffer (re-detection check)
        self.block_seq = 0
        self.ring_writer = [-1, -1]
        self.stat_writer = [-1] * self.n_stat
        # pinned host staging per slot: block counts in, candidate count + records out
        self._rec_prefix = min(self.cand_cap, 1 << 16)
        self._blk_pinned = [torch.zeros(cfg.n_dm, dtype=torch.float64, pin_memory=True) for _ in range(self.n_slots)]
        self._cnt_pinned = [torch.zeros(1, dtype=torch.int32, pin_memory=True) for _ in range(self.n_slots)]
        self._rec_pinned = [torch.zeros((self._rec_prefix, 4), dtype=torch.int32, pin_memory=True)
                            for _ in range(self.n_slots)]
        # default layout: float32 history -> two window buffers filled by LDS-direct
        # loads (tile 128 within 48 KB); fp16 history -> one unpaired buffer (80 KB)
        if dedisp_mode is None:
            dedisp_mode = 3 if ring_dtype == torch.float32 else 2
        if dedisp_mode in (3, 4) and ring_dtype != torch.float32:
            raise ValueError(f"dedisp_mode {dedisp_mode} (LDS-direct window loads) needs a float32 history")
        if lds_budget is None:
            lds_budget = LDS_BUDGET_ASYNC if dedisp_mode == 3 else LDS_BUDGET_DEDISP
        self.lds_budget = int(lds_budget)
        self.dedisp_mode = int(dedisp_mode)
        self.plans: Dict[int, tuple] = {}
        self.cal = None                                  # (gains re, im, time keep, scale) when calibrating
        # ---- stage 1: per-channel coarse-DM shifts on the correlator's operand load;
        # the stage-1 stream trails the raw stream by s1_delay steps (its first block
        # starts at step -s1_delay, the missing past read as zeros)
        self.s1_delay = cfg.stage1_delay
        self.s1_shift = None
        self.s1_tail = None
        if cfg.stage1_dm is not None and with_front:
            self.s1_shift = torch.from_numpy(cfg.stage1_bins).to(self.dev)
            n_tail = self.s1_delay * (cfg.t_int // 2) * cfg.n_ant * cfg.n_chan * 4
            self.s1_tail = torch.zeros(max(n_tail, 4), dtype=torch.uint8, device=self.dev)
        # ---- static sky: per-(subband, pixel) means of the last static_sky_blocks blocks
        self.sky_ring = self.sky_sum = None
        self.sky_head = self.sky_count = self.sky_seen = 0
        if cfg.static_sky_blocks > 0 and with_front:
            self.sky_ring = torch.zeros((cfg.static_sky_blocks, cfg.n_sub, self.n_pix), dtype=torch.float32,
                                        device=self.dev)
            self.sky_sum = torch.zeros((cfg.n_sub, self.n_pix), dtype=torch.float64, device=self.dev)
        self.T0 = -self.s1_delay
        self.prev_len = 0
        self.marks: Optional[Dict[str, list]] = None     # stage -> [start, end] events when timing

    @property
    def mean(self) -> torch.Tensor:
        """float64 running mean ``[pixel, DM]`` the next block is detected against."""
        return self.mean_bufs[self.stat_cur]

    @property
    def m2(self) -> torch.Tensor:
        """float64 running sum of squared deviations ``[pixel, DM]``."""
        return self.m2_bufs[self.stat_cur]

    @property
    def n_prev(self) -> torch.Tensor:
        """Complete samples per DM behind :attr:`mean` / :attr:`m2` (device float64)."""
        return self.n_prev_bufs[self.stat_cur]

    def device_bytes(self) -> int:
        """Device bytes this node holds (histories, block slots, state, records)."""
        ts = [*self.rings, *self.vis_slots, *self.img_slots, *self.cand_slots, *self.mean_bufs,
              *self.m2_bufs]
        return int(sum(t.numel() * t.element_size() for t in ts))

    def enable_timing(self) -> None:
        """Record HIP events around every kernel (read with :meth:`stage_ms`)."""
        names = ("corr", "image", "ring", "dedisp")
        self.marks = {n: [torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)]
                      for n in names}

    def _mark(self, name: str, which: int) -> None:
        if self.marks is not None:
            s = self.stream if self.stream is not None else torch.cuda.current_stream()
            self.marks[name][which].record(s)

    def stage_ms(self) -> Dict[str, float]:
        """Kernel times of the last block (synchronises)."""
        torch.cuda.synchronize()
        return {n: e[0].elapsed_time(e[1]) for n, e in (self.marks or {}).items()}

    # ------------------------------------------------------------------ helpers
    def _plan_for(self, n_centres: int):
        """Device copy of the dedispersion LDS plan for blocks of ``n_centres`` steps."""
        if n_centres not in self.plans:
            kw = dict(n_widths=self.n_widths, n_centres=n_centres, ring_dtype=self.ring_dtype)
            # the node's layout within its budget; wide plans: one work-group per CU with
            # the whole LDS; two window buffers that do not fit even then: one buffer
            tries = [(self.dedisp_mode, self.lds_budget), (self.dedisp_mode, 160 * 1024)]
            if self.dedisp_mode == 3:
                tries += [(2, LDS_BUDGET_DEDISP), (2, 160 * 1024)]
            plan, err = None, None
            for mode, budget in tries:
                try:
                    plan = plan_dedisp(self.shifts_host.T, lds_budget=budget, mode=mode, **kw)
                    break
                except RuntimeError as e:
                    err = e
            if plan is None:
                raise err
            dev = {k: torch.from_numpy(np.ascontiguousarray(getattr(plan, k))).to(self.dev)
                   for k in ("sub_info", "rows", "tbase")}
            self.plans[n_centres] = (plan, dev)
        return self.plans[n_centres]

    def _block_counts(self, c_lo: int, L: int) -> np.ndarray:
        start = np.maximum(self.valid_from_host, c_lo)
        return np.clip(c_lo + L - start, 0, L).astype(np.float64)

    # ------------------------------------------------------------------ stages
    def _stream(self, stream: Optional[torch.cuda.Stream]) -> torch.cuda.Stream:
        if stream is not None:
            return stream
        return self.stream if self.stream is not None else torch.cuda.current_stream()

    def correlate(self, raw: torch.Tensor, n_samp: int, slot: int = 0,
                  stream: Optional[torch.cuda.Stream] = None) -> torch.Tensor:
        cfg = self.cfg
        n_pkt = n_samp // 2
        n_steps = n_samp // cfg.t_int
        _lib.require_gpu(raw)
        if raw.dtype != torch.uint8 or raw.numel() != n_pkt * cfg.n_ant * cfg.n_chan * 4:
            raise ValueError("raw must be uint8 of shape (n_pkt, n_ant, n_chan, 2, 2)")
        if n_steps > self.cap:
            raise ValueError(f"{n_steps} steps > max_block_steps {self.cap}")
        vis = self.vis_slots[slot][:n_steps]
        self._mark("corr", 0)
        if self.s1_shift is not None:
            if self.cal is not None:
                raise ValueError("stage 1 and calibration do not combine in this node")
            if n_steps < self.s1_delay:
                raise ValueError(f"a {n_steps}-step block is shorter than the stage-1 delay ({self.s1_delay})")
            s = self._stream(stream)
            _lib.call("dsx_corr_i8_stage1", _lib.ptr(raw), _lib.ptr(self.s1_tail), _lib.ptr(self.s1_shift),
                      self.s1_delay, _lib.ptr(vis), n_pkt, cfg.n_ant, cfg.n_chan, cfg.t_int, cfg.csf, 1, 1,
                      _lib.stream(s))
            n_tail = self.s1_delay * (cfg.t_int // 2) * cfg.n_ant * cfg.n_chan * 4
            if n_tail:
                with torch.cuda.stream(s):
                    self.s1_tail[:n_tail].copy_(raw.reshape(-1)[raw.numel() - n_tail:])
        elif self.cal is None:
            _lib.call("dsx_corr_i8", _lib.ptr(raw), None, _lib.ptr(vis), n_pkt, cfg.n_ant, cfg.n_chan,
                      cfg.t_int, cfg.csf, 1, 0, 1, _lib.stream(self._stream(stream)))
        else:
            # calibrated: float (re, im) pairs in the same slot
            vis = vis.view(torch.float32)
            cr, ci, tk, scale = self.cal
            n_acc = int(tk.shape[0]) if tk is not None else 1
            if tk is not None and n_pkt % n_acc:
                raise ValueError(f"time_keep has {n_acc} accumulations; {n_pkt} packets do not divide")
            _lib.call("dsx_corr_cal", _lib.ptr(raw), None, _lib.ptr(cr), _lib.ptr(ci), _lib.ptr(tk), n_acc,
                      _lib.ptr(vis), n_pkt, cfg.n_ant, cfg.n_chan, cfg.t_int, cfg.csf, 1, 0, 1, float(scale),
                      _lib.stream(self._stream(stream)))
        self._mark("corr", 1)
        return vis

    def set_calibration(self, cal_real=None, cal_imag=None, *, cube_mask=None, time_keep=None,
                        scale: float = 1.0) -> None:
        """Apply complex gains on the correlator's operand load from now on.

        ``cal_real``/``cal_imag``: ``(NCHAN or 1, 1, 2, 1, NANTS)``
        (``make_cal_broadcast_tensors``) or ``(NCHAN, 2, NANTS)``; ``cube_mask``
        ``(ANT, CH, POL)`` True = flagged (zero gain); ``time_keep`` ``(n_acc,
        NCHAN, 2)`` per-accumulation keep factors of every block; voltages are
        ``fp16(code * scale)``.  Visibilities become float32 (fp16 MFMA, float32
        accumulation) and the imager sums them with float32 LDS atomics.
        ``set_calibration()`` with no gains returns to the exact integer path."""
        from .corr import _compact_gains
        cfg = self.cfg
        if cal_real is None:
            self.cal = None
            return
        cr, ci = _compact_gains((cal_real, cal_imag), cfg.n_chan, cfg.n_ant, self.dev)
        if cube_mask is not None:
            keep = (~torch.as_tensor(cube_mask, device=self.dev, dtype=torch.bool)).permute(1, 2, 0)
            cr = (cr * keep).contiguous()
            ci = (ci * keep).contiguous()
        tk = None
        if time_keep is not None:
            tk = torch.as_tensor(time_keep, device=self.dev).to(torch.float32).contiguous()
            if tk.dim() != 3 or tuple(tk.shape[1:]) != (cfg.n_chan, 2):
                raise ValueError(f"time_keep must be (n_acc, {cfg.n_chan}, 2), got {tuple(tk.shape)}")
        self.cal = (cr, ci, tk, float(scale))

    def image(self, vis: torch.Tensor, n_steps: int, slot: int = 0,
              stream: Optional[torch.cuda.Stream] = None) -> torch.Tensor:
        cfg = self.cfg
        N = cfg.n_grid
        img = self.img_slots[slot][:cfg.n_sub * n_steps * N * N].view(cfg.n_sub, n_steps, N, N)
        self._mark("image", 0)
        _lib.call("dsx_image_subbands" if vis.dtype == torch.int32 else "dsx_image_subbands_f32", _lib.ptr(vis), _lib.ptr(self.tap_src), _lib.ptr(self.tap_dst),
                  _lib.ptr(self.tap_off), _lib.ptr(img), n_steps, cfg.n_sub, cfg.per_sub, cfg.nbase,
                  cfg.n_grid, self.n_rows, self.npad, 1.0,
                  _lib.DSX_F16 if self.ring_dtype == torch.float16 else _lib.DSX_F32,
                  _lib.stream(self._stream(stream)))
        if self.sky_ring is not None:
            self._static_sky(img, n_steps, stream)
        self._mark("image", 1)
        return img

    def _static_sky(self, img: torch.Tensor, n_steps: int, stream: Optional[torch.cuda.Stream]) -> None:
        """Subtract the mean of the previous ``static_sky_blocks`` blocks' mean
        images (after ``static_sky_warmup`` blocks) and push this block's mean
        (``StaticSkyMean.apply``, corr_fast_integration.py:1049-1246, in the
        image domain)."""
        cfg = self.cfg
        W = cfg.static_sky_blocks
        subtract = int(self.sky_seen >= cfg.static_sky_warmup)
        _lib.call("dsx_static_sky", _lib.ptr(img), _lib.DSX_F16 if self.ring_dtype == torch.float16 else _lib.DSX_F32,
                  cfg.n_sub, n_steps, self.n_pix, _lib.ptr(self.sky_ring), _lib.ptr(self.sky_sum), W, self.sky_head,
                  self.sky_count, subtract, _lib.stream(self._stream(stream)))
        self.sky_count = min(self.sky_count + 1, W)
        self.sky_head = (self.sky_head + 1) % W
        self.sky_seen += 1

    def push_images(self, img: torch.Tensor, n_steps: int, stream: Optional[torch.cuda.Stream] = None) -> None:
        cfg = self.cfg
        if n_steps > self.cap:
            raise ValueError(f"{n_steps} steps > max_block_steps {self.cap}")
        _lib.require_gpu(img)
        if img.dtype != self.ring_dtype:
            raise TypeError(f"images are {img.dtype}; this history holds {self.ring_dtype}")
        if img.numel() < cfg.n_sub * n_steps * self.n_pix or not img.is_contiguous():
            raise ValueError(f"images must be contiguous [n_sub, >= {n_steps}, n_grid, n_grid]")
        old, new = self.rings[self.cur], self.rings[1 - self.cur]
        self.ring_writer[1 - self.cur] = self.block_seq
        self._mark("ring", 0)
        _lib.call("dsx_ring_advance", _lib.ptr(old), _lib.ptr(new), _lib.ptr(img), cfg.n_sub,
                  self.n_pix, self.hist, self.cap, self.prev_len, n_steps,
                  _lib.DSX_F16 if self.ring_dtype == torch.float16 else _lib.DSX_F32,
                  _lib.stream(self._stream(stream)))
        self._mark("ring", 1)
        self.cur = 1 - self.cur
        self.prev_len = n_steps

    def detect_async(self, n_steps: int, slot: int = 0,
                     stream: Optional[torch.cuda.Stream] = None) -> "PendingBlock":
        """Enqueue dedispersion + detection of the block just pushed (centres
        [T0 - 64, T0 + L - 64)) and the copy of its candidate records to host
        memory; :meth:`finish` waits for them and decodes."""
        s = self._stream(stream)
        T0 = self.T0
        c_lo = T0 - DETECT_DELAY_STEPS
        plan, pd = self._plan_for(n_steps)
        counts = self._block_counts(c_lo, n_steps)
        cand, n_cand = self.cand_slots[slot], self.n_cand_slots[slot]
        si, so = self.stat_cur, (self.stat_cur + 1) % self.n_stat
        seq = self.block_seq
        cfg = self.cfg
        lam = cfg.stats_decay(n_steps)
        with torch.cuda.stream(s):
            self._blk_pinned[slot].copy_(torch.from_numpy(counts))
            self.blk_count.copy_(self._blk_pinned[slot], non_blocking=True)
            self.mean_bufs[so].copy_(self.mean_bufs[si])
            self.m2_bufs[so].copy_(self.m2_bufs[si])
            if bool((self.n_prev_host == 0).any()):
                # DMs without state: centre this block's sums on a few of its samples
                _lib.call("dsx_dedisp_prime", _lib.ptr(self.rings[self.cur]),
                          _lib.DSX_F16 if self.ring_dtype == torch.float16 else _lib.DSX_F32,
                          cfg.n_sub, self.n_pix, self.hist, self.row_len, T0, _lib.ptr(self.shift), cfg.n_dm,
                          _lib.ptr(self.valid_from), _lib.ptr(self.n_prev_bufs[si]), c_lo, n_steps,
                          _lib.ptr(self.mean_bufs[so]), _lib.stream(s))
            n_cand.zero_()
            self._mark("dedisp", 0)
            self._launch_dedisp(plan, pd, self.n_prev_bufs[si], self.mean_bufs[so], self.m2_bufs[so], True,
                                cand, n_cand, T0, n_steps, s, float(self.cfg.threshold), self.cur)
            self._mark("dedisp", 1)
            torch.add(self.n_prev_bufs[si], self.blk_count, out=self.n_prev_bufs[so])
            if lam < 1.0:
                _lib.call("dsx_stats_decay", _lib.ptr(self.m2_bufs[so]), self.n_pix * cfg.n_dm,
                          _lib.ptr(self.n_prev_bufs[so]), cfg.n_dm, lam, _lib.stream(s))
            self._cnt_pinned[slot].copy_(n_cand, non_blocking=True)
            self._rec_pinned[slot].copy_(cand[:self._rec_prefix], non_blocking=True)
            done = torch.cuda.Event()
            done.record(s)
        self.stat_writer[so] = seq
        self.stat_cur = so
        self.n_prev_host += counts
        self.n_prev_host *= lam
        self.T0 = T0 + n_steps
        self.block_seq = seq + 1
        return PendingBlock(slot=slot, T0=T0, n_steps=n_steps, done=done, stream=s, seq=seq, stat_in=si,
                            ring=self.cur, plan_key=n_steps)

    def _launch_dedisp(self, plan, pd, n_prev, mean, m2, merge, cand, n_cand, T0, n_steps, s,
                       threshold, ring, cand_cap=None) -> None:
        """One ``dsx_dedisp_detect`` launch on history buffer ``ring`` against
        the state ``mean`` / ``m2`` behind ``n_prev`` samples, merged in place
        when ``merge``; ``cand_cap`` 0 only counts records."""
        cfg = self.cfg
        _lib.call("dsx_dedisp_detect", _lib.ptr(self.rings[ring]),
                  _lib.DSX_F16 if self.ring_dtype == torch.float16 else _lib.DSX_F32,
                  cfg.n_sub, self.n_pix, self.hist, self.row_len, T0,
                  _lib.ptr(pd["sub_info"]), _lib.ptr(pd["rows"]) if plan.n_rows else None, plan.n_rows,
                  _lib.ptr(pd["tbase"]), plan.n_terms, cfg.n_dm, _lib.ptr(self.valid_from),
                  _lib.ptr(n_prev), _lib.ptr(self.blk_count), _lib.ptr(mean), _lib.ptr(m2), int(bool(merge)),
                  T0 - DETECT_DELAY_STEPS, n_steps, self.n_widths, plan.tile, plan.r_total, plan.q_total,
                  int(plan.flags), float(threshold), float(cfg.min_stat_count), _lib.ptr(cand),
                  self.cand_cap if cand_cap is None else int(cand_cap), _lib.ptr(n_cand), _lib.stream(s))

    AUTOTUNE_BUDGETS = (16 << 10, 24 << 10, 32 << 10, 48 << 10, 64 << 10, 80 << 10, 112 << 10, 160 << 10)

    def layouts(self, n_steps: int, modes: Optional[tuple] = None, budgets=None) -> list:
        """Every distinct dedispersion LDS layout (time tile, pair rows, two
        window buffers) the planner offers for blocks of ``n_steps`` over the
        autotuner's LDS budgets."""
        if modes is None:
            modes = (2, 1, 3, 4) if self.ring_dtype == torch.float32 else (2, 1)
        kw = dict(n_widths=self.n_widths, n_centres=n_steps, ring_dtype=self.ring_dtype)
        seen, out = set(), []
        for mode in modes:
            for b in (budgets or self.AUTOTUNE_BUDGETS):
                try:
                    plan = plan_dedisp(self.shifts_host.T, lds_budget=b, mode=mode, **kw)
                except RuntimeError:
                    continue
                key = (plan.tile, plan.paired, plan.flags)
                if key not in seen:
                    seen.add(key)
                    out.append(plan)
        return out

    def use_plan(self, n_steps: int, plan) -> None:
        """Detect blocks of ``n_steps`` with ``plan`` (a :class:`DedispPlan`)."""
        pd = {k: torch.from_numpy(np.ascontiguousarray(getattr(plan, k))).to(self.dev)
              for k in ("sub_info", "rows", "tbase")}
        self.plans[n_steps] = (plan, pd)

    def autotune(self, n_steps: int, *, budgets=AUTOTUNE_BUDGETS,
                 modes: Optional[tuple] = None, reps: int = 2) -> List[dict]:
        """Time the dedispersion kernel under each LDS layout the planner
        offers (time tile from the LDS budget; unpaired or subband-pair rows)
        on the current history and keep the fastest for blocks of ``n_steps``.
        Runs on scratch copies of the statistics and candidate buffers, so the
        stream's state is untouched.  Returns the timings."""
        if modes is None:
            modes = (2, 1, 3, 4) if self.ring_dtype == torch.float32 else (2, 1)
        s = self._stream(None)
        mean, m2 = torch.empty_like(self.mean), torch.empty_like(self.m2)
        cand = torch.empty_like(self.cand_slots[0])
        n_cand = torch.zeros(1, dtype=torch.int32, device=self.dev)
        seen, rows, best = set(), [], None
        kw = dict(n_widths=self.n_widths, n_centres=n_steps, ring_dtype=self.ring_dtype)
        for mode in modes:
            for b in budgets:
                try:
                    plan = plan_dedisp(self.shifts_host.T, lds_budget=b, mode=mode, **kw)
                except RuntimeError:
                    continue
                key = (plan.tile, plan.paired, plan.flags)
                if key in seen:
                    continue
                seen.add(key)
                pd = {k: torch.from_numpy(np.ascontiguousarray(getattr(plan, k))).to(self.dev)
                      for k in ("sub_info", "rows", "tbase")}
                ms = []
                with torch.cuda.stream(s):
                    for _ in range(reps):
                        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                        mean.copy_(self.mean)
                        m2.copy_(self.m2)
                        n_cand.zero_()
                        e0.record(s)
                        self._launch_dedisp(plan, pd, self.n_prev, mean, m2, True, cand, n_cand, self.T0, n_steps,
                                            s, float(self.cfg.threshold), self.cur)
                        e1.record(s)
                        e1.synchronize()
                        ms.append(e0.elapsed_time(e1))
                t = min(ms)
                rows.append({"tile": plan.tile, "paired": plan.paired, "async": bool(plan.dma),
                             "lds_bytes": plan.lds_bytes, "ms": t})
                if best is None or t < best[0]:
                    best = (t, plan, pd)
        if best is None:
            raise RuntimeError("no dedispersion layout fits")
        self.plans[n_steps] = (best[1], best[2])
        return rows

    def finish(self, pend: "PendingBlock") -> BlockResult:
        """Wait for a block's detection and decode its candidates.

        Overflow policy: when more chunk maxima beat the threshold than the
        record buffer holds, the kept records would depend on the order of
        the kernel's atomic appends.  The block is then detected again
        (records only, no statistics merge) against the same history and
        state at the lowest threshold of the ladder ``threshold *
        sqrt(2)**j`` whose records all fit, so the result is deterministic and
        keeps every strongest record; ``threshold_used`` reports that
        threshold and ``n_raw`` the uncapped count at the node's threshold."""
        cfg = self.cfg
        pend.done.synchronize()
        n = int(self._cnt_pinned[pend.slot][0])
        thr = float(cfg.threshold)
        c_lo = pend.T0 - DETECT_DELAY_STEPS
        k = min(n, self.cand_cap)
        if k <= self._rec_prefix:
            rec_dev = self._rec_pinned[pend.slot][:k].numpy().copy().view(CAND_DTYPE).reshape(-1)
        else:
            with torch.cuda.stream(pend.stream):
                rec_dev = self.cand_slots[pend.slot][:k].cpu().numpy().view(CAND_DTYPE).reshape(-1)
        if n > self.cand_cap and thr > 0 and self._can_redetect(pend):
            thr, rec_dev = self._redetect(pend, n, rec_dev)
        rec = device_records(rec_dev, c_lo)
        cands = decode_chunk_maxima(rec, widths=cfg.widths, n_grid=cfg.n_grid, dm_trials=cfg.dm_trials,
                                    threshold=thr, n_top=cfg.n_top, radius_t=cfg.radius_t,
                                    radius_dm=cfg.radius_dm, radius_lm=cfg.radius_lm)
        return BlockResult(candidates=cands, n_raw=n, overflow=n > self.cand_cap, T0=pend.T0,
                           n_steps=pend.n_steps, records=rec, threshold_used=thr)

    def _redetect(self, pend: "PendingBlock", n: int, kept: np.ndarray):
        """Lowest ladder threshold (j >= 1) whose records fit, and its records.

        The kept records are a sample of the block's records, so the fraction
        above each threshold estimates where to probe first; count-only passes
        then walk the ladder until level j fits and level j - 1 does not (or
        j = 1), which makes the result independent of the sample, and a last
        pass keeps the records of level j."""
        levels = record_levels(self.cfg.threshold)
        frac = np.array([(kept["snr"] > lv).mean() if kept.size else 1.0 for lv in levels])
        est = n * frac
        fits = [j for j in range(1, N_LEVELS) if est[j] <= 0.9 * self.cand_cap]
        j = fits[0] if fits else N_LEVELS - 1
        plan, pd = self._plan_for(pend.plan_key)
        cand, n_cand = self.cand_slots[pend.slot], self.n_cand_slots[pend.slot]

        def probe(jj, keep):
            with torch.cuda.stream(pend.stream):
                n_cand.zero_()
                si = pend.stat_in
                self._launch_dedisp(plan, pd, self.n_prev_bufs[si], self.mean_bufs[si], self.m2_bufs[si], False,
                                    cand, n_cand, pend.T0, pend.n_steps, pend.stream, float(levels[jj]), pend.ring,
                                    None if keep else 0)
                return int(n_cand[0].item())

        counts = {}

        def count(jj):
            if jj not in counts:
                counts[jj] = probe(jj, False)
            return counts[jj]

        while j < N_LEVELS - 1 and count(j) > self.cand_cap:
            j += 1
        while j > 1 and count(j - 1) <= self.cand_cap:
            j -= 1
        counts[j] = probe(j, True)
        m = min(counts[j], self.cand_cap)
        with torch.cuda.stream(pend.stream):
            rec_dev = cand[:m].cpu().numpy().view(CAND_DTYPE).reshape(-1)
        return float(levels[j]), rec_dev

    def _can_redetect(self, pend: "PendingBlock") -> bool:
        """The history and state buffers block ``pend`` read are unchanged."""
        return self.ring_writer[pend.ring] == pend.seq and self.stat_writer[pend.stat_in] < pend.seq

    def detect(self, n_steps: int) -> BlockResult:
        """Dedisperse + detect the block just pushed and wait for its candidates."""
        return self.finish(self.detect_async(n_steps))

    def _check_block(self, n_samp: int) -> int:
        cfg = self.cfg
        n_steps = n_samp // cfg.t_int
        if n_samp % (2 * cfg.t_int) or n_steps % DETECT_CHUNK_STEPS or n_steps > self.cap or n_steps <= 0:
            raise ValueError(f"block of {n_samp} samples: steps must be a positive multiple of "
                             f"{DETECT_CHUNK_STEPS} and <= {self.cap}")
        return n_steps

    def front(self, raw: torch.Tensor, n_samp: int, slot: int = 0,
              stream: Optional[torch.cuda.Stream] = None) -> torch.Tensor:
        """Correlate and image a block into ``slot``; returns its images."""
        n_steps = self._check_block(n_samp)
        vis = self.correlate(raw, n_samp, slot, stream)
        return self.image(vis, n_steps, slot, stream)

    def back(self, img: torch.Tensor, n_steps: int, slot: int = 0,
             stream: Optional[torch.cuda.Stream] = None) -> "PendingBlock":
        """Push a block's images into the history and enqueue its detection.
        ``pend.img_free`` marks when the slot's images have been consumed."""
        s = self._stream(stream)
        self.push_images(img, n_steps, s)
        img_free = torch.cuda.Event()
        img_free.record(s)
        pend = self.detect_async(n_steps, slot, s)
        pend.img_free = img_free
        return pend

    def process_block(self, raw: torch.Tensor, n_samp: Optional[int] = None) -> BlockResult:
        n_samp = self.cfg.n_samp if n_samp is None else int(n_samp)
        n_steps = self._check_block(n_samp)
        img = self.front(raw, n_samp)
        return self.finish(self.back(img, n_steps))


class CoarseDMNode:
    """Node stream with stage 1 (``cfg.coarse_dms``): per coarse DM, the
    correlator reads every channel group ``stage1_bins`` steps later (its delay
    inside the subband at that DM removed before gridding,
    ``Stage1MultiDMCoarseDM``, corr_fast_integration.py:495-1040), images the
    stage-1 stream (static sky subtracted when configured) into that coarse
    DM's history, and searches the fine trials whose nearest coarse DM it is.
    Per coarse DM this is a :class:`SearchNode`; the nodes share one set of
    block buffers and run in turn on one stream.  Candidates of all groups are
    decoded together (DM index offset by the earlier groups' trials); every
    group's stream trails the raw stream by the same stage-1 delay, so block
    times agree."""

    def __init__(self, cfg: NodeConfig, antpos_e: np.ndarray, antpos_n: np.ndarray, *,
                 device: str = "cuda", ring_dtype: torch.dtype = torch.float32,
                 max_block_steps: Optional[int] = None, stream: Optional[torch.cuda.Stream] = None,
                 cand_cap: int = 1 << 20):
        if not cfg.coarse_dms:
            raise ValueError("CoarseDMNode needs cfg.coarse_dms")
        cfg.validate()
        self.cfg = cfg
        cap = int(max_block_steps or cfg.n_steps)
        self.groups = []
        first = None
        off = 0
        for gc in cfg.coarse_groups():
            nd = SearchNode(gc, antpos_e, antpos_n, device=device, ring_dtype=ring_dtype, max_block_steps=cap,
                            stream=stream, cand_cap=cand_cap, share_front=first)
            first = first or nd
            self.groups.append((nd, off))
            off += gc.n_dm

    @property
    def stage1_delay(self) -> int:
        return self.groups[0][0].s1_delay

    def autotune(self, n_samp: Optional[int] = None) -> List[List[dict]]:
        n_samp = self.cfg.n_samp if n_samp is None else int(n_samp)
        n_steps = self.groups[0][0]._check_block(n_samp)
        return [nd.autotune(n_steps) for nd, _ in self.groups]

    def enable_timing(self) -> None:
        for nd, _ in self.groups:
            nd.enable_timing()

    def stage_ms(self) -> Dict[str, float]:
        out: Dict[str, float] = {}
        for nd, _ in self.groups:
            for k, v in nd.stage_ms().items():
                out[k] = out.get(k, 0.0) + v
        return out

    def process_block(self, raw: torch.Tensor, n_samp: Optional[int] = None) -> BlockResult:
        cfg = self.cfg
        n_samp = cfg.n_samp if n_samp is None else int(n_samp)
        pends = []
        for nd, _ in self.groups:
            n_steps = nd._check_block(n_samp)
            img = nd.front(raw, n_samp)
            pends.append(nd.back(img, n_steps))
        recs = []
        n_raw = 0
        overflow = False
        thr = []
        for (nd, off), pend in zip(self.groups, pends):
            res = nd.finish(pend)
            n_raw += res.n_raw
            overflow = overflow or res.overflow
            thr.append(res.threshold_used)
            r = res.records.copy()
            r["d"] = r["d"] + off
            recs.append(r)
        rec = np.concatenate(recs) if recs else np.zeros(0, REC_DTYPE)
        cands = decode_chunk_maxima(rec, widths=list(cfg.widths), n_grid=cfg.n_grid, dm_trials=cfg.dm_trials,
                                    threshold=cfg.threshold, n_top=cfg.n_top, radius_t=cfg.radius_t,
                                    radius_dm=cfg.radius_dm, radius_lm=cfg.radius_lm)
        return BlockResult(candidates=cands, n_raw=int(n_raw), overflow=overflow, T0=int(pends[0].T0),
                           n_steps=pends[0].n_steps, records=rec, threshold_used=max(thr), range_thresholds=thr)


@dataclass
class PendingBlock:
    """A block whose detection is enqueued (see :meth:`SearchNode.finish`)."""
    slot: int
    T0: int
    n_steps: int
    done: torch.cuda.Event
    stream: torch.cuda.Stream
    img_free: Optional[torch.cuda.Event] = None
    seq: int = 0                 # block number on its node
    stat_in: int = 0             # state buffer the block was detected against
    ring: int = 0                # history buffer it read
    plan_key: int = 0


class MultiRangeNode:
    """Node stream with an uneven DM plan (``cfg.dm_ranges``): one front stage
    (correlate + image at full time resolution) feeding one image history and
    detector per DM range, each at its own time resolution (images summed over
    ``f`` consecutive steps by ``dsx_downsample_time``) with its own trials,
    widths in samples of that resolution, float64 statistics and LDS plan.
    Candidates of all ranges are decoded together in full-resolution steps:
    time ``t * f``, width ``w * f``, DM index offset by the earlier ranges'
    trials.  ``cand_cap`` bounds each range's chunk-maximum records per block;
    the block's ``overflow`` is set when any range exceeded it and ``n_raw``
    counts every range's records, kept or not."""

    def __init__(self, cfg: NodeConfig, antpos_e: np.ndarray, antpos_n: np.ndarray, *,
                 device: str = "cuda", ring_dtype: torch.dtype = torch.float32,
                 max_block_steps: Optional[int] = None, stream: Optional[torch.cuda.Stream] = None,
                 cand_cap: int = 1 << 20):
        if not cfg.dm_ranges:
            raise ValueError("MultiRangeNode needs cfg.dm_ranges")
        cfg.validate()
        self.cfg = cfg
        self.stream = stream
        cap = int(max_block_steps or cfg.n_steps)
        self.front_node = SearchNode(cfg.with_(dm_ranges=()), antpos_e, antpos_n, device=device,
                                     ring_dtype=ring_dtype, max_block_steps=cap, stream=stream, with_detect=False)
        self.ranges = []
        off = 0
        for rc, (lo, hi, n, f, _) in zip(cfg.range_configs(), cfg.dm_ranges):
            nd = SearchNode(rc, antpos_e, antpos_n, device=device, ring_dtype=ring_dtype,
                            max_block_steps=cap // int(f), stream=stream, with_front=False,
                            cand_cap=cand_cap)
            buf = None if int(f) == 1 else torch.empty(cfg.n_sub * (cap // int(f)) * cfg.n_grid ** 2,
                                                       dtype=ring_dtype, device=torch.device(device))
            self.ranges.append((nd, int(f), off, buf))
            off += int(n)
        self.widths_all = [w * f for _, f, _, _ in self.ranges for w in cfg.widths]
        self.ring_dtype = ring_dtype

    def autotune(self, n_samp: Optional[int] = None) -> List[List[dict]]:
        """Pick each range's dedispersion LDS layout for blocks of ``n_samp``
        samples by timing (:meth:`SearchNode.autotune`); returns the timings
        per range."""
        n_samp = self.cfg.n_samp if n_samp is None else int(n_samp)
        n_steps = self.front_node._check_block(n_samp)
        return [nd.autotune(n_steps // f) for nd, f, _, _ in self.ranges]

    def process_block(self, raw: torch.Tensor, n_samp: Optional[int] = None) -> BlockResult:
        cfg = self.cfg
        n_samp = cfg.n_samp if n_samp is None else int(n_samp)
        n_steps = self.front_node._check_block(n_samp)
        s = self.front_node._stream(None)
        img = self.front_node.front(raw, n_samp, 0, s)
        N2 = cfg.n_grid * cfg.n_grid
        pends = []
        done = {1: img}                     # images at each resolution made so far
        for nd, f, off, buf in self.ranges:
            if n_steps % (DETECT_CHUNK_STEPS * f):
                raise ValueError(f"{n_steps} steps do not downsample by {f} into whole 64-sample chunks")
            if f == 1:
                img_f = img
            else:
                # from the coarsest resolution already made whose factor divides f
                # (a 4x range sums pairs of the 2x range's images)
                fs = max(k for k in done if f % k == 0)
                if fs == f:
                    img_f = done[f]          # an earlier range at the same resolution
                else:
                    img_f = buf[:cfg.n_sub * (n_steps // f) * N2].view(cfg.n_sub, n_steps // f, cfg.n_grid,
                                                                       cfg.n_grid)
                    _lib.call("dsx_downsample_time", _lib.ptr(done[fs]), _lib.ptr(img_f), cfg.n_sub, n_steps // fs,
                              N2, f // fs, _lib.DSX_F16 if self.ring_dtype == torch.float16 else _lib.DSX_F32,
                              _lib.stream(s))
                    done[f] = img_f
            nd.push_images(img_f, n_steps // f, s)
            pends.append(nd.detect_async(n_steps // f, 0, s))
        recs, T0 = [], None
        k_off = n_raw = 0
        overflow = False
        thr_used = []
        for (nd, f, off, _), pend in zip(self.ranges, pends):
            res = nd.finish(pend)
            thr_used.append(res.threshold_used)
            n_raw += res.n_raw
            overflow = overflow or res.overflow
            T0 = pend.T0 * f if T0 is None else T0
            r = res.records.copy()
            r["t"] = r["t"] * f                 # int64
            r["d"] = r["d"] + off
            r["k"] = r["k"] + k_off
            recs.append(r)
            k_off += len(cfg.widths)
        rec = np.concatenate(recs) if recs else np.zeros(0, REC_DTYPE)
        cands = decode_chunk_maxima(rec, widths=self.widths_all, n_grid=cfg.n_grid, dm_trials=cfg.dm_trials,
                                    threshold=cfg.threshold, n_top=cfg.n_top, radius_t=cfg.radius_t,
                                    radius_dm=cfg.radius_dm, radius_lm=cfg.radius_lm)
        return BlockResult(candidates=cands, n_raw=int(n_raw), overflow=overflow, T0=int(T0), n_steps=n_steps,
                           records=rec, threshold_used=max(thr_used), range_thresholds=thr_used)
