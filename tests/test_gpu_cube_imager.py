"""GpuImager.build(cfg).process_cube against the float64 oracle: random cint8
uv streams of 16 chgroups, per-trial shifts (some reaching past the stream
start and end, so chgroups drop out), unit-scale (exact int32 sums) and
per-chgroup / per-(chgroup, row) calibrated combines, t_lo > 0, per-trial
masks, fp32 and fp16 cubes, N = 64 (radix-2 rows) and N = 256 (in-register
16-point rows, four uv-row passes).

Oracle: sum_g scale * z_g[t - shift] + offset in float64, then
``fftshift(Re(ifft2(G))) * signed edge mask`` (oracle.image), i.e. the
reference's ``Re(fftshift(ifft2(uv)))`` times ``compute_edge_mask``.
Bars: relative to the cube's peak, 1e-5 for float32 cubes and 1e-3 for
float16 cubes; rows before t_lo untouched."""
import numpy as np
import pytest
import torch

from dsx.imager import GpuImager, GpuImagerConfig
from oracle import image as oimage

pytestmark = pytest.mark.gpu


def _oracle(streams, shifts, t_det, t_lo, scale=None, off_re=None, off_im=None, mask=None):
    n_chg, T, _, N, _ = streams.shape
    n_fdm = shifts.shape[0]
    out = np.zeros((t_det, n_fdm, N, N))
    z = streams[:, :, 0].astype(np.float64) + 1j * streams[:, :, 1].astype(np.float64)
    for f in range(n_fdm):
        for t in range(t_lo, t_det):
            G = np.zeros((N, N), dtype=np.complex128)
            for g in range(n_chg):
                tau = t - int(shifts[f, g])
                if 0 <= tau < T:
                    sc = 1.0 if scale is None else (scale[g, tau] if scale.ndim == 2 else scale[g])
                    ore = 0.0 if off_re is None else (off_re[g, tau] if off_re.ndim == 2 else off_re[g])
                    oim = 0.0 if off_im is None else (off_im[g, tau] if off_im.ndim == 2 else off_im[g])
                    G += sc * z[g, tau] + (ore + 1j * oim)
            m = mask if mask.ndim == 2 else mask[f]
            out[t, f] = oimage.image(G, m)
    return out


@pytest.mark.parametrize("n_grid,cube_dtype,cal,t_lo", [
    (64, torch.float32, None, 0), (64, torch.float16, "1d", 3), (256, torch.float32, "2d", 2),
    (256, torch.float16, None, 0)])
def test_process_cube_vs_float64(n_grid, cube_dtype, cal, t_lo):
    rng = np.random.default_rng(n_grid + (7 if cal else 0))
    n_chg, T, t_det, n_fdm = 16, 14, 10, 3
    cfg = GpuImagerConfig(n_grid=n_grid, t_det=t_det, n_fdm=n_fdm, n_chgroup=n_chg, cube_dtype=cube_dtype)
    im = GpuImager.build(cfg)
    streams = rng.integers(-127, 128, size=(n_chg, T, 2, n_grid, n_grid), dtype=np.int8)
    shifts = rng.integers(-2, 7, size=(n_fdm, n_chg)).astype(np.int32)
    shifts[0] = 0
    kw, want_kw = {}, {}
    if cal == "1d":
        sc, ore, oim = (rng.uniform(0.5, 2.0, n_chg).astype(np.float32),
                        rng.normal(0, 3, n_chg).astype(np.float32), rng.normal(0, 3, n_chg).astype(np.float32))
        kw = dict(chgroup_scales=torch.from_numpy(sc).cuda(), chgroup_offsets_re=torch.from_numpy(ore).cuda(),
                  chgroup_offsets_im=torch.from_numpy(oim).cuda())
        want_kw = dict(scale=sc.astype(np.float64), off_re=ore.astype(np.float64), off_im=oim.astype(np.float64))
    elif cal == "2d":
        sc = rng.uniform(0.5, 2.0, (n_chg, T)).astype(np.float32)
        ore = rng.normal(0, 3, n_chg).astype(np.float32)
        kw = dict(chgroup_scales=torch.from_numpy(sc).cuda(), chgroup_offsets_re=torch.from_numpy(ore).cuda())
        want_kw = dict(scale=sc.astype(np.float64), off_re=ore.astype(np.float64))
    mask = oimage.edge_mask(n_grid, 5)
    if n_grid == 64:            # per-trial masks with a scale folded in
        per = np.stack([mask * (1.0 + 0.5 * f) for f in range(n_fdm)]).astype(np.float32)
        im.set_edge_mask_per_fdm(torch.from_numpy(per).cuda())
        mask = per.astype(np.float64)
    s_t = torch.from_numpy(streams).cuda()
    sh_t = torch.from_numpy(shifts).cuda()
    sentinel = 12288.0          # exact in float16
    im.output_cube.fill_(sentinel)
    out = im.process_cube(streams_cint8=s_t, time_shifts_gpu=sh_t, t_lo=t_lo, **kw)
    torch.cuda.synchronize()
    assert out.data_ptr() == im.output_cube.data_ptr() and out.dtype == cube_dtype
    got = out.float().cpu().numpy()
    want = _oracle(streams, shifts, t_det, t_lo, mask=mask, **want_kw)
    peak = np.abs(want).max()
    err = np.abs(got[t_lo:] - want[t_lo:]).max() / peak
    print(f"N={n_grid} {cube_dtype} cal={cal} t_lo={t_lo}: max err {err:.2e} of peak {peak:.3g}", flush=True)
    assert err <= (1e-5 if cube_dtype == torch.float32 else 1e-3), err
    assert np.all(got[:t_lo] == sentinel)
    # the next cube goes to the other buffer
    out2 = im.process_cube(streams_cint8=s_t, time_shifts_gpu=sh_t, t_lo=t_lo, **kw)
    assert out2.data_ptr() == im.output_cube_alt.data_ptr()


def test_process_cube_rejects_bad_shapes():
    cfg = GpuImagerConfig(n_grid=64, t_det=8, n_fdm=2, n_chgroup=4, cube_dtype=torch.float32)
    im = GpuImager.build(cfg)
    s = torch.zeros((4, 8, 2, 64, 64), dtype=torch.int8, device="cuda")
    sh = torch.zeros((2, 4), dtype=torch.int32, device="cuda")
    with pytest.raises(ValueError, match="int8"):
        im.process_cube(streams_cint8=s.float(), time_shifts_gpu=sh)
    with pytest.raises(ValueError, match="N_chgroup"):
        im.process_cube(streams_cint8=s[:3], time_shifts_gpu=sh)
    with pytest.raises(ValueError, match="time_shifts_gpu.shape"):
        im.process_cube(streams_cint8=s, time_shifts_gpu=sh[:1])
    with pytest.raises(ValueError, match="T_stream"):
        im.process_cube(streams_cint8=s[:, :4], time_shifts_gpu=sh)
    with pytest.raises(ValueError, match="t_lo"):
        im.process_cube(streams_cint8=s, time_shifts_gpu=sh, t_lo=8)
    with pytest.raises(ValueError, match="chgroup_scales.shape"):
        im.process_cube(streams_cint8=s, time_shifts_gpu=sh, chgroup_scales=torch.ones(5, device="cuda"))
