"""dsx_cube_candidates_noalloc (static device scratch, no allocation) gives
exactly the records, counts and order of dsx_cube_candidates (workspace
argument) on the same cube prefix -- two calls in a row, the second on a
smaller cube, so stale scratch contents from the first cannot leak -- and
rejects n_top beyond its maximum."""
import ctypes

import numpy as np
import pytest
import torch

from dsx import _lib

pytestmark = pytest.mark.gpu

WIDTHS = np.array([1, 2, 4, 8], dtype=np.int32)
H = 4


def _prefix(rng, T, n_cell):
    x = torch.from_numpy(rng.standard_normal((T, n_cell)).astype(np.float32)).cuda()
    cs = torch.empty((T + 2 * H + 1, n_cell), dtype=torch.float32, device="cuda")
    _lib.call("dsx_cube_cumsum", _lib.ptr(x), _lib.DSX_F32, _lib.ptr(cs), T, n_cell, H, None)
    return x, cs


def _run(name, cs, T, n_cell, n_top, thr, ws=None):
    out = torch.full((n_top, 4), -7, dtype=torch.int32, device="cuda")
    counts = torch.full((2,), -7, dtype=torch.int32, device="cuda")
    inv = np.array([1.0, 0.7, 0.5, 0.35], dtype=np.float32)
    P = ctypes.c_void_p
    args = [_lib.ptr(cs), T, n_cell, H, WIDTHS.ctypes.data_as(P), inv.ctypes.data_as(P), len(WIDTHS),
            ctypes.c_float(thr), H, T - 2 * H, n_top, _lib.ptr(out), _lib.ptr(counts[:1]), _lib.ptr(counts[1:])]
    if ws is not None:
        args += [_lib.ptr(ws), int(ws.numel())]
    _lib.call(name, *args, None)
    torch.cuda.synchronize()
    c = counts.cpu().numpy()
    return out[:int(c[0])].cpu().numpy(), c


def test_noalloc_matches_workspace_entry_point():
    rng = np.random.default_rng(12)
    cap = int(_lib.lib().dsx_cube_candidates_noalloc_capacity())
    assert cap == 1 << 20
    ws = torch.empty(256 + 16 * cap, dtype=torch.uint8, device="cuda")
    for T, n_cell, n_top, thr in ((200, 40000, 512, 3.0), (64, 3000, 100, 2.5)):
        print(f"T={T} n_cell={n_cell} n_top={n_top}", flush=True)
        x, cs = _prefix(rng, T, n_cell)
        want, wc = _run("dsx_cube_candidates", cs, T, n_cell, n_top, thr, ws)
        got, gc = _run("dsx_cube_candidates_noalloc", cs, T, n_cell, n_top, thr)
        assert np.array_equal(gc, wc) and wc[0] == min(n_top, wc[1]) and wc[1] > n_top
        assert np.array_equal(got, want)
        snr = got[:, 3].view(np.float32)
        assert np.all(snr[:-1] >= snr[1:]) and np.all(snr > thr)


def test_noalloc_rejects_n_top_beyond_maximum():
    rng = np.random.default_rng(13)
    x, cs = _prefix(rng, 32, 100)
    with pytest.raises(RuntimeError, match="beyond the maximum"):
        _run("dsx_cube_candidates_noalloc", cs, 32, 100, 5000, 3.0)
