"""C1 -> C2 batch encoding and the C2 trigger packet (``dsx.wire``).

Known answers come from the schema document (field formats, the MJD
recipe) and from the reference's ``tests/test_c1c2_wire.py`` cases (round
trips, framing errors, trigger magic); the expected bytes are written out
by hand from the documented formats, not produced by the code under test.
"""
from __future__ import annotations

import dataclasses
import struct
from types import SimpleNamespace

import numpy as np
import pytest

from dsx import wire
from dsx.detector import Candidate, CandidateFlags
from dsx.merger import MergerConfig


def rows(n):
    return [wire.C1CandidateRow(snr=10.0 + 0.5 * i, l_rad=1.25e-3 * (i - 2), m_rad=-2.5e-4 * i, l_pix=128 + i,
                                m_pix=100 - i, dm_pc_cc=375.25 + i, dm_idx_global=50 + i, fine_dm_idx=i,
                                event_specnum=100_000 + 16 * i, width_samples=1 << (i % 6),
                                kernel_id="unit:d1:b4", flags=0) for i in range(n)]


def header(n):
    return wire.build_header(cube_id=7, event_specnum_start=100_000, mjd_start=60781.123456789012,
                             sample_period_specnum=16, sample_period_us=1048.576, n_grid=256, n_fdm_in_cube=34,
                             search_node_id=1, gpu_half=0, n_candidates=n)


def test_golden_bytes():
    blob = wire.encode_c1_batch(header(1), rows(1))
    want = (b"# C1 1 7 100000 60781.12345678901 16 1048.576000 256 34 1 0 1\n"
            b"1.000000e+01 -2.500000000e-03 -0.000000000e+00 128 100 375.250000 50 0 100000 1 unit:d1:b4 0\n"
            b"# END\n")
    assert blob == want
    assert wire.encode_c1_batch(header(0), []) == \
        b"# C1 1 7 100000 60781.12345678901 16 1048.576000 256 34 1 0 0\n# END\n"


def test_candidate_mjd_counts_search_samples():
    h = header(0)
    spn = h.event_specnum_start + 100
    assert h.candidate_mjd(spn) == pytest.approx(h.mjd_start + 100 * 1048.576e-6 / 86400.0, abs=1e-15)
    assert h.candidate_mjd(h.event_specnum_start) == h.mjd_start
    h1 = dataclasses.replace(h, sample_period_specnum=1)
    assert h1.candidate_mjd(spn + 18) == h.candidate_mjd(spn + 18)


@pytest.mark.parametrize("n", [0, 1, 5, 25])
def test_roundtrip(n):
    r, h = rows(n), header(n)
    blob = wire.encode_c1_batch(h, r)
    got = wire.parse_c1_batch(blob.decode().rstrip("\n").split("\n"))
    assert got.header == h and len(got.candidates) == n
    for a, b in zip(r, got.candidates):
        for f in ("l_pix", "m_pix", "event_specnum", "width_samples", "kernel_id", "dm_idx_global", "fine_dm_idx",
                  "flags"):
            assert getattr(a, f) == getattr(b, f)
        assert abs(a.snr - b.snr) < 1e-5 * abs(a.snr)
        assert abs(a.dm_pc_cc - b.dm_pc_cc) < 1e-5
        assert abs(a.l_rad - b.l_rad) <= 1e-12 and abs(a.m_rad - b.m_rad) <= 1e-12


def test_framing_errors():
    with pytest.raises(ValueError):
        wire.encode_c1_batch(header(3), rows(2))
    with pytest.raises(wire.BadBatch, match="schema_version"):
        wire.parse_c1_batch(["# C1 9 1 0 60781.00000000000 16 1048.576000 256 34 1 0 0", "# END"])
    lines = wire.encode_c1_batch(header(3), rows(3)).decode().rstrip("\n").split("\n")
    with pytest.raises(wire.BadBatch, match="truncated"):
        wire.parse_c1_batch(lines[:2] + ["# END"])
    with pytest.raises(wire.BadBatch, match="missing END"):
        wire.parse_c1_batch(lines[:-1])
    bad = lines[:]
    bad[-1] = "# NOT_END"
    with pytest.raises(wire.BadBatch, match="END"):
        wire.parse_c1_batch(bad)
    bad = lines[:]
    bad[1] = " ".join(bad[1].split()[:-1])
    with pytest.raises(wire.BadBatch, match="12 fields"):
        wire.parse_c1_batch(bad)
    with pytest.raises(wire.BadBatch, match="empty"):
        wire.parse_c1_batch([])
    with pytest.raises(ValueError, match="kernel_id"):
        wire.encode_c1_batch(header(1), [dataclasses.replace(rows(1)[0], kernel_id="a b")])


def test_trigger_packet():
    pkt = wire.C2TriggerPacket(event_name="260521abcd", event_specnum=987_654_321, mjd_target=60781.5,
                               trigger_class_id=1, flags=wire.C2_TRIGGER_FLAG_DUMP_CUBE)
    blob = wire.encode_c2_trigger(pkt)
    assert len(blob) == 64 and blob[:4] == b"DSRT"
    # field offsets of the documented struct: version @4, class @6, name @8, specnum @24, mjd @32, flags @40
    assert struct.unpack_from("<H", blob, 4)[0] == 1 and struct.unpack_from("<H", blob, 6)[0] == 1
    assert blob[8:24] == b"260521abcd" + bytes(6)
    assert struct.unpack_from("<q", blob, 24)[0] == 987_654_321
    assert struct.unpack_from("<d", blob, 32)[0] == 60781.5
    assert struct.unpack_from("<I", blob, 40)[0] == 1 and blob[44:] == bytes(20)
    assert wire.decode_c2_trigger(blob) == pkt
    with pytest.raises(ValueError):
        wire.encode_c2_trigger(wire.C2TriggerPacket(event_name="x" * 17, event_specnum=0, mjd_target=0.0))
    with pytest.raises(wire.BadBatch, match="wrong size"):
        wire.decode_c2_trigger(bytes(32))
    b = bytearray(blob)
    b[0] ^= 0xFF
    with pytest.raises(wire.BadBatch, match="magic"):
        wire.decode_c2_trigger(bytes(b))


def _cand(**kw):
    base = dict(l=64.0, m=200.5, dm_fine=500.0, dm_idx=3, event_specnum=1234, width_samples=4,
                kernel_id="unit:d1:b4", snr=12.5, detector_version="dsx.v1", flags=int(CandidateFlags.NONE),
                search_node_id=0, gpu_half=0)
    base.update(kw)
    return Candidate(**base)


def _geom(**kw):
    base = dict(cube_id=3, specnum_start=1000, sample_period_specnum=16, t_det=256, n_grid=256, n_fdm_in_cube=4,
                sample_period_us=1048.576, cell_l_rad=1.24039e-4, cell_m_rad=1.24039e-4, l0_rad=0.0, m0_rad=0.0,
                fine_dm_pc_cc=np.array([0.0, 250.0, 499.7, 750.0]), mjd_start=60000.0)
    base.update(kw)
    return wire.CubeGeometry(**base)


def test_candidate_row_geometry():
    """Centred image, row axis = sky m, column axis = sky l: a column of
    200.5 on a 256 grid at 1.24039e-4 rad per cell is l = 72.5 cells =
    0.0089928 rad (the reference's l = +0.009 injection landing at column
    200.5); the row 64 is m = -64 cells."""
    r = wire.candidate_to_c1_row(_cand(), geom=_geom())
    assert r.l_rad == pytest.approx(72.5 * 1.24039e-4, rel=1e-15)
    assert r.m_rad == pytest.approx(-64 * 1.24039e-4, rel=1e-15)
    assert (r.l_pix, r.m_pix) == (64, 200)            # round half to even, as Python's round
    assert r.fine_dm_idx == 2 and r.dm_idx_global == 3 and r.dm_pc_cc == 500.0
    assert wire.candidate_to_c1_row(_cand(), geom=_geom(), fine_global_idx=[10, 11, 12, 13]).dm_idx_global == 12
    c = wire.candidate_to_c1_row(_cand(l=-3.0, m=300.0), geom=_geom())
    assert (c.l_pix, c.m_pix) == (0, 255)
    assert wire.centred_pix_offset(5.0, 0) == 5.0


def test_cube_batch_and_block_geometry():
    """Merged survivors become rows under a header taken from the block."""
    cfg = SimpleNamespace(n_grid=256, t_int=8, dm_trials=np.array([0.0, 10.0, 20.0]), t_step_us=262.144)
    node = SimpleNamespace(cfg=cfg, cmap=SimpleNamespace(cell_lambda=31.5))
    res = SimpleNamespace(T0=4096, n_steps=2048)
    geom = wire.block_geometry(node, res, cube_id=9, mjd_start=60100.25)
    assert geom.cell_l_rad == pytest.approx(1.0 / (256 * 31.5))
    assert (geom.specnum_start, geom.sample_period_specnum, geom.t_det, geom.n_fdm_in_cube) == (4096, 4, 2048, 3)
    cands = [_cand(snr=20.0, event_specnum=5000, dm_fine=10.0, dm_idx=1),
             _cand(snr=15.0, event_specnum=5002, dm_fine=10.0, dm_idx=1, l=66.0),
             _cand(snr=9.0, event_specnum=6000, dm_fine=20.0, dm_idx=2)]
    blob, surv = wire.cube_batch(cands, geom, search_node_id=2, gpu_half=1,
                                 merger=MergerConfig(sample_period_specnum=geom.sample_period_specnum))
    assert [c.snr for c in surv] == [20.0, 9.0]
    got = wire.parse_c1_batch(blob.decode().rstrip("\n").split("\n"))
    h = got.header
    assert (h.cube_id, h.event_specnum_start, h.sample_period_specnum, h.n_grid, h.n_fdm_in_cube,
            h.search_node_id, h.gpu_half, h.n_candidates) == (9, 4096, 4, 256, 3, 2, 1, 2)
    assert h.sample_period_us == pytest.approx(262.144)
    assert [r.fine_dm_idx for r in got.candidates] == [1, 2]
    assert h.candidate_mjd(got.candidates[0].event_specnum) == pytest.approx(
        60100.25 + (5000 - 4096) * 262.144e-6 / 86400.0, abs=1e-12)
