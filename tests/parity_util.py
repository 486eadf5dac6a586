"""Record comparison helpers shared by the GPU parity tests (not collected)."""
import numpy as np


def oracle_records(o, thr):
    """Chunk maxima of the oracle's last block above ``thr``:
    {(DM, time, pixel): (width index, snr)}."""
    L = o.last
    d, c, p = np.nonzero(L["best_snr"] > thr)
    return {(int(dd), int(L["best_t"][dd, cc, pp]), int(o.pix[pp])): (int(L["best_k"][dd, cc, pp]),
                                                                       float(L["best_snr"][dd, cc, pp]))
            for dd, cc, pp in zip(d, c, p)}


def gpu_records(rec, pix=None):
    """Host records -> {(DM, time, pixel): (width index, snr)}, optionally
    restricted to the pixels ``pix``."""
    keep = None if pix is None else set(int(x) for x in pix)
    return {(int(r["d"]), int(r["t"]), int(r["p"])): (int(r["k"]), float(r["snr"])) for r in rec
            if keep is None or int(r["p"]) in keep}


def compare_records(g, o, thr, rtol, what=""):
    """Identical record sets (ignoring S/N within 1e-3 of the threshold, where
    float32 and float64 may fall on different sides), identical widths, S/N
    within ``rtol`` relative."""
    near = lambda v: abs(v - thr) < 1e-3 * max(1.0, abs(thr))
    gk = {k for k, v in g.items() if not near(v[1])}
    ok = {k for k, v in o.items() if not near(v[1])}
    assert gk == ok, (f"{what}: record sets differ: gpu-only {sorted(gk - ok)[:5]} "
                      f"oracle-only {sorted(ok - gk)[:5]}")
    for k in gk:
        assert g[k][0] == o[k][0], f"{what} {k}: width {g[k][0]} vs {o[k][0]}"
        assert abs(g[k][1] - o[k][1]) <= rtol * abs(o[k][1]), f"{what} {k}: snr {g[k][1]} vs {o[k][1]}"
    return len(gk)
