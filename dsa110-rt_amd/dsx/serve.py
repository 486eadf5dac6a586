"""Streaming server: voltage blocks of uneven length from one or more node
streams, admitted by HBM budget and processed with overlapping stages.

Each node stream keeps its image history, float64 noise state and
dedispersion plan resident on the device (:class:`SearchNode`).  A block's
front stage (correlate + image, on the ``front`` HIP stream) runs while the
previous block's back stage (history + fused dedispersion/detection, on the
``back`` stream) is still on the GPU; blocks from different node streams share
the two HIP streams, so their work overlaps too.  A block is admitted when a
slot of its node is free and the bytes it brings fit the HBM budget; blocks
that arrive while others are on the GPU wait for that.  Within a node, blocks
are processed in arrival order (the history and the statistics carry from one
block to the next), so a block's result does not depend on what else is in
the mix.

This is the caller the reference spreads over its corr block loop and search
cube loop (``src/dsart/services/corr_fast_integration.py:2229-2367``,
``src/dsart/services/cube_pipeline.py:2641-2728``) on one GPU.
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np
import torch

from .config import NodeConfig
from .pipeline import BlockResult, PendingBlock, SearchNode


@dataclass
class BlockRequest:
    """One voltage block of a node stream, arriving at ``t_arrival`` (seconds
    on the server clock).  ``make`` enqueues the block's bytes on the given
    HIP stream and returns the device tensor (on-wire layout)."""
    stream_id: int
    seq: int
    n_samp: int
    t_arrival: float
    make: Callable[[torch.cuda.Stream], torch.Tensor]


@dataclass
class BlockDone:
    stream_id: int
    seq: int
    n_samp: int
    t_arrival: float
    t_admit: float
    t_done: float
    result: BlockResult

    @property
    def latency(self) -> float:
        """Arrival to candidate list on the host (seconds)."""
        return self.t_done - self.t_arrival


@dataclass
class _Inflight:
    req: BlockRequest
    t_admit: float
    pend: PendingBlock
    nbytes: int


class NodeServer:
    """Serve block requests for ``len(nodes)`` node streams on one GPU."""

    def __init__(self, nodes: Sequence[SearchNode], *, hbm_budget_bytes: int = 200 << 30):
        if not nodes:
            raise ValueError("at least one node stream")
        self.nodes = list(nodes)
        for nd in self.nodes:
            if nd.n_slots < 2:
                raise ValueError("served nodes need n_slots >= 2 (overlapping front and back stages)")
        self.budget = int(hbm_budget_bytes)
        self.front = torch.cuda.Stream()          # default priority
        self.back = torch.cuda.Stream()
        self.inflight: collections.deque = collections.deque()
        self.next_slot: Dict[int, int] = {i: 0 for i in range(len(self.nodes))}
        self.slot_free: Dict[tuple, Optional[torch.cuda.Event]] = {}
        self.owner: Dict[tuple, Optional[_Inflight]] = {}
        self.bytes_inflight = 0
        self.budget_waits = 0                      # admissions that waited for HBM, not for a slot
        self.done: List[BlockDone] = []
        self.t0 = time.perf_counter()

    # ------------------------------------------------------------------ clock
    def now(self) -> float:
        return time.perf_counter() - self.t0

    @staticmethod
    def block_bytes(nd: SearchNode, n_samp: int) -> int:
        """Device bytes a block brings while in flight (its on-wire voltages)."""
        return (n_samp // 2) * nd.cfg.n_ant * nd.cfg.n_chan * 4

    def resident_bytes(self) -> int:
        """Device bytes held by the node streams (histories, slots, state)."""
        return sum(nd.device_bytes() for nd in self.nodes)

    # ------------------------------------------------------------------ stages
    def _retire(self, block: bool) -> bool:
        """Finish the oldest in-flight block if it is done (or wait when
        ``block``); returns True when one was retired."""
        if not self.inflight:
            return False
        f: _Inflight = self.inflight[0]
        if not block and not f.pend.done.query():
            return False
        self.inflight.popleft()
        nd = self.nodes[f.req.stream_id]
        res = nd.finish(f.pend)
        self.bytes_inflight -= f.nbytes
        self.owner[(f.req.stream_id, f.pend.slot)] = None
        self.done.append(BlockDone(f.req.stream_id, f.req.seq, f.req.n_samp, f.req.t_arrival, f.t_admit,
                                   self.now(), res))
        return True

    def _admit(self, req: BlockRequest) -> None:
        nd = self.nodes[req.stream_id]
        nbytes = self.block_bytes(nd, req.n_samp)
        if self.resident_bytes() + nbytes > self.budget:
            raise ValueError(f"block of {req.n_samp} samples cannot fit the HBM budget "
                             f"({self.budget} B) next to {self.resident_bytes()} B resident")
        slot = self.next_slot[req.stream_id]
        key = (req.stream_id, slot)
        # wait until the slot is free and the block fits the budget
        waited_for_memory = False
        while self.owner.get(key) is not None or \
                self.resident_bytes() + self.bytes_inflight + nbytes > self.budget:
            if self.owner.get(key) is None:
                waited_for_memory = True
            self._retire(block=True)
        self.budget_waits += int(waited_for_memory)
        t_admit = self.now()
        n_steps = nd._check_block(req.n_samp)
        free = self.slot_free.get(key)
        if free is not None:
            self.front.wait_event(free)               # the slot's images were consumed
        with torch.cuda.stream(self.front):
            raw = req.make(self.front)
        img = nd.front(raw, req.n_samp, slot, self.front)
        raw.record_stream(self.front)
        front_done = torch.cuda.Event()
        front_done.record(self.front)
        self.back.wait_event(front_done)
        pend = nd.back(img, n_steps, slot, self.back)
        self.slot_free[key] = pend.img_free
        f = _Inflight(req, t_admit, pend, nbytes)
        self.owner[key] = f
        self.inflight.append(f)
        self.bytes_inflight += nbytes
        self.next_slot[req.stream_id] = (slot + 1) % nd.n_slots
        del raw

    # ------------------------------------------------------------------ driver
    def serve(self, requests: Sequence[BlockRequest], *, realtime: bool = True) -> List[BlockDone]:
        """Admit ``requests`` (sorted by arrival) as they arrive and slots and
        memory free up; returns every block's result in completion order.
        With ``realtime`` False arrivals are ignored (as fast as possible)."""
        reqs = sorted(requests, key=lambda r: (r.t_arrival, r.stream_id, r.seq))
        last_seq: Dict[int, int] = {}
        for r in reqs:
            if r.seq <= last_seq.get(r.stream_id, -1):
                raise ValueError(f"stream {r.stream_id}: blocks must arrive in order")
            last_seq[r.stream_id] = r.seq
        self.t0 = time.perf_counter()
        for r in reqs:
            if realtime:
                while self.now() < r.t_arrival:
                    if not self._retire(block=False):
                        time.sleep(min(0.0005, max(0.0, r.t_arrival - self.now())))
            self._admit(r)
            while self._retire(block=False):
                pass
        while self._retire(block=True):
            pass
        return self.done


def poisson_arrivals(durations_s: Sequence[float], rate: float, seed: int) -> np.ndarray:
    """Arrival times of consecutive blocks whose inter-arrival gaps are
    exponential with mean ``duration / rate`` (data offered at ``rate`` x real
    time on average)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(np.asarray(durations_s, dtype=np.float64) / float(rate))
    return np.concatenate([[0.0], np.cumsum(gaps[:-1])])


def summarize(done: Sequence[BlockDone], cfg: NodeConfig) -> dict:
    lat = np.array([d.latency for d in done])
    data_s = sum(d.n_samp for d in done) * cfg.t_step_us / cfg.t_int * 1e-6
    span = max(d.t_done for d in done) - min(d.t_arrival for d in done)
    block_s = np.array([d.n_samp * cfg.t_step_us / cfg.t_int * 1e-6 for d in done])
    return {"blocks": len(done), "data_seconds": data_s, "wall_seconds": span,
            "sustained_rtf": data_s / span,
            "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3), "p99": float(np.percentile(lat, 99) * 1e3),
                           "max": float(lat.max() * 1e3)},
            "latency_over_block_duration_p99": float(np.percentile(lat / block_s, 99)),
            "candidates": int(sum(len(d.result.candidates) for d in done))}
