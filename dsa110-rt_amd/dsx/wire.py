"""Candidate wire format: C1 -> C2 ASCII batches and the C2 -> C1 trigger packet.

Byte-level contract of ``docs/c1c2/C1C2_WIRE_SCHEMA.md`` (schema version 1),
owned in the reference by ``src/dsart/coinc/wire.py:85-415``; the mapping of
a detector ``Candidate`` onto a row follows
``src/dsart/services/c1_emit.py:380-469``.

* Batch: a header line ``# C1 <schema> <cube_id> <event_specnum_start>
  <mjd_start %.11f> <sample_period_specnum> <sample_period_us %.6f> <n_grid>
  <n_fdm_in_cube> <search_node_id> <gpu_half> <n_candidates>``, one row per
  candidate ``<snr %.6e> <l_rad %.9e> <m_rad %.9e> <l_pix> <m_pix>
  <dm_pc_cc %.6f> <dm_idx_global> <fine_dm_idx> <event_specnum>
  <width_samples> <kernel_id> <flags>``, then ``# END``; sent as one write.
* Trigger: 64 bytes little endian ``<IHH16sqdII16s``: magic ``b"DSRT"``,
  version, class id, NUL-padded event name, event specnum, MJD, flags,
  reserved, padding.

``cube_batch`` turns one cube's (or search block's) detector output into an
encoded batch: C1 cross-kernel merge, row projection, header.  The TCP
transport, its queue and reconnect policy are a service concern and not
part of this module.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field
from typing import Iterable, List, Optional, Sequence, Tuple

import numpy as np

from .merger import MergerConfig, merge_across_kernels_c1

SCHEMA_VERSION = 1
C2_TRIGGER_MAGIC = 0x54525344            # b"DSRT" read as a little-endian u32
C2_TRIGGER_PACKET_SIZE = 64
C2_TRIGGER_FLAG_DUMP_CUBE = 1 << 0
C2_TRIGGER_FLAG_DUMP_VOLTAGE = 1 << 1
_TRIGGER = struct.Struct("<IHH16sqdII16s")
assert _TRIGGER.size == C2_TRIGGER_PACKET_SIZE


class BadBatch(ValueError):
    """Framing or schema error in a C1 batch or trigger packet."""


@dataclass(frozen=True)
class C1CandidateRow:
    """One candidate row; ``width_samples`` is the boxcar width in search samples."""
    snr: float
    l_rad: float
    m_rad: float
    l_pix: int
    m_pix: int
    dm_pc_cc: float
    dm_idx_global: int
    fine_dm_idx: int
    event_specnum: int
    width_samples: int
    kernel_id: str
    flags: int


@dataclass(frozen=True)
class C1BatchHeader:
    schema_version: int
    cube_id: int
    event_specnum_start: int
    mjd_start: float
    sample_period_specnum: int
    sample_period_us: float
    n_grid: int
    n_fdm_in_cube: int
    search_node_id: int
    gpu_half: int
    n_candidates: int

    def candidate_mjd(self, event_specnum: int) -> float:
        """MJD of a row: both specnums count search samples, so their
        difference times ``sample_period_us`` is the offset
        (``C1C2_WIRE_SCHEMA.md`` section 1.1; ``sample_period_specnum`` does
        not enter)."""
        return self.mjd_start + (event_specnum - self.event_specnum_start) * self.sample_period_us / 1e6 / 86400.0


@dataclass(frozen=True)
class C1Batch:
    header: C1BatchHeader
    candidates: Tuple[C1CandidateRow, ...] = field(default_factory=tuple)


def build_header(*, cube_id: int, event_specnum_start: int, mjd_start: float, sample_period_specnum: int,
                 sample_period_us: float, n_grid: int, n_fdm_in_cube: int, search_node_id: int, gpu_half: int,
                 n_candidates: int) -> C1BatchHeader:
    return C1BatchHeader(SCHEMA_VERSION, int(cube_id), int(event_specnum_start), float(mjd_start),
                         int(sample_period_specnum), float(sample_period_us), int(n_grid), int(n_fdm_in_cube),
                         int(search_node_id), int(gpu_half), int(n_candidates))


_HEADER_FMT = "# C1 {} {} {} {:.11f} {} {:.6f} {} {} {} {} {}\n"
_ROW_FMT = "{:.6e} {:.9e} {:.9e} {:d} {:d} {:.6f} {:d} {:d} {:d} {:d} {} {:d}\n"


def encode_c1_batch(header: C1BatchHeader, rows: Sequence[C1CandidateRow]) -> bytes:
    if header.n_candidates != len(rows):
        raise ValueError(f"header.n_candidates={header.n_candidates} != len(rows)={len(rows)}")
    h = header
    out = [_HEADER_FMT.format(h.schema_version, h.cube_id, h.event_specnum_start, h.mjd_start,
                              h.sample_period_specnum, h.sample_period_us, h.n_grid, h.n_fdm_in_cube,
                              h.search_node_id, h.gpu_half, h.n_candidates)]
    for r in rows:
        if not r.kernel_id or any(ch.isspace() for ch in r.kernel_id):
            raise ValueError(f"kernel_id {r.kernel_id!r} must be a non-empty token without spaces")
        out.append(_ROW_FMT.format(r.snr, r.l_rad, r.m_rad, r.l_pix, r.m_pix, r.dm_pc_cc, r.dm_idx_global,
                                   r.fine_dm_idx, r.event_specnum, r.width_samples, r.kernel_id, r.flags))
    out.append("# END\n")
    return "".join(out).encode("utf-8")


class C1BatchEncoder:
    """Name-compatible holder of ``encode_c1_batch``."""
    encode = staticmethod(encode_c1_batch)


def _parse_header(line: str) -> C1BatchHeader:
    s = line.strip()
    if not s.startswith("# C1 "):
        raise BadBatch(f"header must start with '# C1 '; got {line!r}")
    tok = s[5:].split()
    if len(tok) != 11:
        raise BadBatch(f"header expects 11 fields after '# C1 '; got {len(tok)}: {tok}")
    kinds = (int, int, int, float, int, float, int, int, int, int, int)
    try:
        v = [k(t) for k, t in zip(kinds, tok)]
    except ValueError as exc:
        raise BadBatch(f"header type error: {exc}; line={line!r}") from None
    hdr = C1BatchHeader(*v)
    if hdr.schema_version != SCHEMA_VERSION:
        raise BadBatch(f"unsupported schema_version={hdr.schema_version} (this build accepts only {SCHEMA_VERSION})")
    if hdr.n_candidates < 0:
        raise BadBatch(f"n_candidates={hdr.n_candidates} must be >= 0")
    if hdr.sample_period_specnum <= 0:
        raise BadBatch(f"sample_period_specnum={hdr.sample_period_specnum} must be > 0")
    if hdr.sample_period_us <= 0.0:
        raise BadBatch(f"sample_period_us={hdr.sample_period_us} must be > 0")
    return hdr


def _parse_row(line: str) -> C1CandidateRow:
    tok = line.split()
    if len(tok) != 12:
        raise BadBatch(f"row expects 12 fields; got {len(tok)}: {line!r}")
    kinds = (float, float, float, int, int, float, int, int, int, int, str, int)
    try:
        return C1CandidateRow(*[k(t) for k, t in zip(kinds, tok)])
    except ValueError as exc:
        raise BadBatch(f"row type error: {exc}; line={line!r}") from None


def parse_c1_batch(lines: Iterable[str]) -> C1Batch:
    """One batch from its lines (without newlines); ``BadBatch`` on any
    framing error, including a ``#`` line where a row is due (a torn batch)."""
    it = iter(lines)
    first = next(it, None)
    if first is None:
        raise BadBatch("empty batch")
    hdr = _parse_header(first)
    rows: List[C1CandidateRow] = []
    for _ in range(hdr.n_candidates):
        ln = next(it, None)
        if ln is None or ln.lstrip().startswith("#"):
            seen = "" if ln is None else f" (saw '#' line {ln!r})"
            raise BadBatch(f"truncated batch: expected {hdr.n_candidates} rows, got {len(rows)}{seen}")
        rows.append(_parse_row(ln))
    end = next(it, None)
    if end is None:
        raise BadBatch("missing END marker")
    if end.strip() != "# END":
        raise BadBatch(f"expected '# END', got {end!r}")
    return C1Batch(header=hdr, candidates=tuple(rows))


# ---------------------------------------------------------------------------
# C2 -> C1 trigger packet
# ---------------------------------------------------------------------------


@dataclass(frozen=True)
class C2TriggerPacket:
    event_name: str
    event_specnum: int
    mjd_target: float
    trigger_class_id: int = 0
    flags: int = C2_TRIGGER_FLAG_DUMP_CUBE


def encode_c2_trigger(pkt: C2TriggerPacket) -> bytes:
    name = pkt.event_name.encode("ascii")
    if len(name) > 16:
        raise ValueError(f"event_name {pkt.event_name!r} > 16 bytes; truncate or change schema")
    return _TRIGGER.pack(C2_TRIGGER_MAGIC, SCHEMA_VERSION, int(pkt.trigger_class_id), name.ljust(16, b"\0"),
                         int(pkt.event_specnum), float(pkt.mjd_target), int(pkt.flags), 0, bytes(16))


def decode_c2_trigger(buf: bytes) -> C2TriggerPacket:
    if len(buf) != C2_TRIGGER_PACKET_SIZE:
        raise BadBatch(f"trigger packet wrong size: {len(buf)} bytes (expected {C2_TRIGGER_PACKET_SIZE})")
    magic, ver, cls, name, spn, mjd, flags, _res, _pad = _TRIGGER.unpack(buf)
    if magic != C2_TRIGGER_MAGIC:
        raise BadBatch(f"trigger magic mismatch: 0x{magic:08x} (expected 0x{C2_TRIGGER_MAGIC:08x})")
    if ver != SCHEMA_VERSION:
        raise BadBatch(f"trigger schema_version={ver} unsupported (build accepts {SCHEMA_VERSION})")
    return C2TriggerPacket(event_name=name.rstrip(b"\0").decode("ascii"), event_specnum=int(spn),
                           mjd_target=float(mjd), trigger_class_id=int(cls), flags=int(flags))


# ---------------------------------------------------------------------------
# Candidate -> row
# ---------------------------------------------------------------------------


@dataclass(frozen=True)
class CubeGeometry:
    """Per-cube geometry sidecar (``src/dsart/common/contracts.py:524-581``):
    pixel pitch and origin in radians, the cube's fine-DM grid (pc cm^-3),
    sample periods and the MJD of sample 0."""
    cube_id: int
    specnum_start: int
    sample_period_specnum: int
    t_det: int
    n_grid: int
    n_fdm_in_cube: int
    sample_period_us: float
    cell_l_rad: float
    cell_m_rad: float
    l0_rad: float
    m0_rad: float
    fine_dm_pc_cc: np.ndarray
    mjd_start: float


def centred_pix_offset(pix: float, n_grid: int) -> float:
    """Signed offset of a pixel from the image centre ``n_grid // 2``
    (``src/dsart/cluster/features.py:63-88``); unchanged for ``n_grid <= 0``."""
    return float(pix) - float(int(n_grid) // 2) if int(n_grid) > 0 else float(pix)


def candidate_to_c1_row(cand, *, geom: CubeGeometry, fine_global_idx: Optional[Sequence[int]] = None) -> C1CandidateRow:
    """Row for one candidate.  Pixels are rounded and clamped to the grid;
    the image is centred and its row axis is sky m, its column axis sky l,
    so ``l_rad`` comes from the column (``cand.m``) and ``m_rad`` from the
    row (``cand.l``); ``fine_dm_idx`` is the nearest entry of the cube's
    fine-DM grid, and ``dm_idx_global`` maps through ``fine_global_idx``
    when given (``c1_emit.py:380-469``)."""
    n = int(geom.n_grid)
    l_pix = min(max(int(round(float(cand.l))), 0), n - 1)
    m_pix = min(max(int(round(float(cand.m))), 0), n - 1)
    l_rad = float(geom.l0_rad) + float(geom.cell_l_rad) * centred_pix_offset(float(cand.m), n)
    m_rad = float(geom.m0_rad) + float(geom.cell_m_rad) * centred_pix_offset(float(cand.l), n)
    fdm = geom.fine_dm_pc_cc
    fine_idx = int(np.argmin(np.abs(np.asarray(fdm, dtype=np.float64) - float(cand.dm_fine)))) \
        if fdm is not None and np.size(fdm) > 0 else 0
    dm_global = int(cand.dm_idx)
    if fine_global_idx is not None and 0 <= fine_idx < len(fine_global_idx):
        dm_global = int(fine_global_idx[fine_idx])
    return C1CandidateRow(snr=float(cand.snr), l_rad=l_rad, m_rad=m_rad, l_pix=l_pix, m_pix=m_pix,
                          dm_pc_cc=float(cand.dm_fine), dm_idx_global=dm_global, fine_dm_idx=fine_idx,
                          event_specnum=int(cand.event_specnum), width_samples=int(cand.width_samples),
                          kernel_id=str(cand.kernel_id), flags=int(cand.flags))


def block_geometry(node, result, *, cube_id: int = 0, mjd_start: float = 0.0) -> CubeGeometry:
    """Geometry of one ``SearchNode`` block: the block's first step as
    ``specnum_start`` (candidate ``event_specnum`` counts absolute steps),
    one step per search sample, the centred image's pitch ``1 / (N *
    cell)`` radians for a uv cell of ``cell`` wavelengths, the node's DM
    trials as the fine-DM grid."""
    cfg = node.cfg
    pitch = 1.0 / (cfg.n_grid * float(node.cmap.cell_lambda))
    dms = np.asarray(cfg.dm_trials, dtype=np.float64)
    return CubeGeometry(cube_id=int(cube_id), specnum_start=int(result.T0), sample_period_specnum=cfg.t_int // 2,
                        t_det=int(result.n_steps), n_grid=cfg.n_grid, n_fdm_in_cube=int(dms.size),
                        sample_period_us=float(cfg.t_step_us), cell_l_rad=pitch, cell_m_rad=pitch, l0_rad=0.0,
                        m0_rad=0.0, fine_dm_pc_cc=dms, mjd_start=float(mjd_start))


def cube_batch(candidates: Sequence, geom: CubeGeometry, *, search_node_id: int = 0, gpu_half: int = 0,
               merger: Optional[MergerConfig] = MergerConfig(),
               fine_global_idx: Optional[Sequence[int]] = None) -> Tuple[bytes, List]:
    """Encoded batch of one cube's candidates and the survivors it carries:
    the C1 cross-kernel merge (skipped when ``merger`` is None), one row per
    survivor, the header from ``geom``."""
    surv = merge_across_kernels_c1(candidates, merger) if merger is not None else list(candidates)
    rows = [candidate_to_c1_row(c, geom=geom, fine_global_idx=fine_global_idx) for c in surv]
    hdr = build_header(cube_id=geom.cube_id, event_specnum_start=geom.specnum_start, mjd_start=geom.mjd_start,
                       sample_period_specnum=geom.sample_period_specnum, sample_period_us=geom.sample_period_us,
                       n_grid=geom.n_grid, n_fdm_in_cube=geom.n_fdm_in_cube, search_node_id=search_node_id,
                       gpu_half=gpu_half, n_candidates=len(rows))
    return encode_c1_batch(hdr, rows), surv
