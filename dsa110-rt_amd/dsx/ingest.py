"""Host ingest: raw voltage pages in pinned host memory, copied to the device
block by block on a copy stream while the previous block computes.

The reference copies each 288 MiB page synchronously before its correlator
(``src/dsart/x2/corr/pipeline1x.py:1-40``; the pinned staging of
``src/dsart/services/slow_corr_kernel.py:281-301``).  Here a block's H2D copy
(on its own HIP stream, default priority) runs while the GPU is still on the
previous block's correlate / image / dedisperse chain: the compute stream
waits on the copy's event before the correlator reads the slot, and the copy
into a slot waits for the correlator that last read it.

``PinnedPagePool`` holds the host pages a receiver (NIC, file reader, test
generator) fills in place -- on-wire layout ``(n_pkt, n_ant, n_chan, 2, 2)``
bytes.  ``HostIngest.submit(page, n_samp)`` enqueues the copy and returns a
``StagedBlock`` whose ``raw`` view the node processes after ``wait()``.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional

import torch

from .config import NodeConfig


def page_bytes(cfg: NodeConfig, n_samp: int) -> int:
    """Bytes of one raw block of ``n_samp`` samples (one byte per complex sample)."""
    return (n_samp // 2) * cfg.n_ant * cfg.n_chan * 4


class PinnedPagePool:
    """``n_pages`` page-locked host pages of ``max_bytes`` (``hipHostMalloc``
    through torch's pinned allocator), handed out round robin."""

    def __init__(self, n_pages: int, max_bytes: int):
        if n_pages < 1 or max_bytes < 1:
            raise ValueError("n_pages and max_bytes must be positive")
        self.pages: List[torch.Tensor] = [torch.empty(int(max_bytes), dtype=torch.uint8, pin_memory=True)
                                          for _ in range(int(n_pages))]
        self._next = 0

    def next_page(self) -> torch.Tensor:
        p = self.pages[self._next]
        self._next = (self._next + 1) % len(self.pages)
        return p


@dataclass
class StagedBlock:
    """A block whose H2D copy is enqueued into device slot ``slot``."""
    raw: torch.Tensor               # device view (n_pkt, n_ant, n_chan, 2, 2) uint8
    n_samp: int
    slot: int
    ready: torch.cuda.Event         # recorded on the copy stream after the copy
    t_submit: float = 0.0

    def wait(self, stream: Optional[torch.cuda.Stream] = None) -> None:
        """Make ``stream`` (default: the current one) wait for the copy."""
        (stream or torch.cuda.current_stream()).wait_event(self.ready)


class HostIngest:
    """Double-buffered (``n_slots`` >= 2) H2D of raw pages into device block
    slots on a dedicated copy stream."""

    def __init__(self, cfg: NodeConfig, *, max_samp: Optional[int] = None, n_slots: int = 2,
                 device: str = "cuda"):
        if n_slots < 2:
            raise ValueError("n_slots must be >= 2 (one block copies while the previous computes)")
        self.cfg = cfg
        self.max_samp = int(max_samp or cfg.n_samp)
        self.max_bytes = page_bytes(cfg, self.max_samp)
        dev = torch.device(device)
        self.slots = [torch.empty(self.max_bytes, dtype=torch.uint8, device=dev) for _ in range(n_slots)]
        self.copy_stream = torch.cuda.Stream(device=dev)          # default priority
        self.freed: List[Optional[torch.cuda.Event]] = [None] * n_slots
        self._next = 0
        self.copy_events: List[tuple] = []

    def submit(self, page: torch.Tensor, n_samp: int, *, timed: bool = False) -> StagedBlock:
        """Enqueue the copy of ``page[:bytes(n_samp)]`` into the next slot."""
        cfg = self.cfg
        nbytes = page_bytes(cfg, n_samp)
        if n_samp > self.max_samp or n_samp % (2 * cfg.t_int):
            raise ValueError(f"block of {n_samp} samples: at most {self.max_samp}, a multiple of {2 * cfg.t_int}")
        if page.dtype != torch.uint8 or page.device.type != "cpu" or page.numel() < nbytes:
            raise ValueError("page must be a host uint8 tensor holding the block's on-wire bytes")
        if not page.is_pinned():
            raise ValueError("page must be pinned host memory (PinnedPagePool) for an asynchronous copy")
        slot = self._next
        self._next = (slot + 1) % len(self.slots)
        dst = self.slots[slot][:nbytes]
        with torch.cuda.stream(self.copy_stream):
            if self.freed[slot] is not None:
                self.copy_stream.wait_event(self.freed[slot])     # its previous block was read
            e0 = torch.cuda.Event(enable_timing=True) if timed else None
            if e0 is not None:
                e0.record(self.copy_stream)
            dst.copy_(page[:nbytes], non_blocking=True)
            ready = torch.cuda.Event(enable_timing=timed)
            ready.record(self.copy_stream)
        if timed:
            self.copy_events.append((e0, ready, nbytes))
        raw = dst.view(n_samp // 2, cfg.n_ant, cfg.n_chan, 2, 2)
        return StagedBlock(raw=raw, n_samp=n_samp, slot=slot, ready=ready)

    def release(self, blk: StagedBlock, stream: Optional[torch.cuda.Stream] = None) -> None:
        """Mark ``blk``'s slot reusable once ``stream``'s work so far (the
        correlator that read it) is done."""
        ev = torch.cuda.Event()
        ev.record(stream or torch.cuda.current_stream())
        self.freed[blk.slot] = ev

    def copy_rate(self) -> dict:
        """Mean H2D time and rate of the timed copies so far (synchronises)."""
        if not self.copy_events:
            return {}
        self.copy_events[-1][1].synchronize()
        ms = [a.elapsed_time(b) for a, b, _ in self.copy_events]
        nb = [n for _, _, n in self.copy_events]
        mean = sum(ms) / len(ms)
        return {"h2d_ms": mean, "h2d_GBps": (sum(nb) / len(nb)) / (mean * 1e-3) / 1e9, "copies": len(ms)}


def process_host_stream(node, ingest: HostIngest, pages, n_samps, on_result=None):
    """Run ``node`` (``SearchNode``) over host pages: block i + 1's copy is
    enqueued before block i is processed, so the copy overlaps its compute.
    Returns the ``BlockResult`` list (``on_result(i, res)`` is called as each
    completes)."""
    out = []
    staged = ingest.submit(pages[0], n_samps[0]) if pages else None
    for i in range(len(pages)):
        cur = staged
        staged = ingest.submit(pages[i + 1], n_samps[i + 1]) if i + 1 < len(pages) else None
        cur.wait()
        res = node.process_block(cur.raw, cur.n_samp)
        ingest.release(cur)
        out.append(res)
        if on_result is not None:
            on_result(i, res)
    return out
