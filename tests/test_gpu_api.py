"""GPU tests of the reference-named stage APIs against the oracle / numpy.

Bars: integer paths bit-exact; float32 paths within 1e-5 of the float64
result relative to its peak (1e-3 where the inputs are float16)."""
import numpy as np
import pytest
import torch

from dsx import config as dcfg
from dsx import synth
from dsx.corr import FastCorrKernel, stokes_i_pol_sum, unpack_int4_split
from dsx.dedisp import combine_chgroups
from dsx.detector import boxcar_via_cumsum, multi_boxcar_argmax
from dsx.grid import FastVisGridder, build_pattern
from dsx.imager import dirty_image_from_uv_grid
from oracle import corr as ocorr
from oracle import image as oimage
from oracle import search as osearch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small():
    cfg = dcfg.small_node_config().with_(n_samp=256)
    sc = synth.make_scenario(cfg)
    return cfg, sc, sc.block(0, 256)


@pytest.mark.parametrize("dtype", [torch.int8, torch.float16, torch.float32])
def test_unpack_int4_split(small, dtype):
    cfg, sc, raw = small
    re, im = unpack_int4_split(raw, device="cuda", scale=0.05, out_dtype=dtype, shape=(128, 16, 64))
    assert re.shape == (64, 2, 2, 128, 16)
    lo, hi = ocorr.unpack_bytes(raw)                      # (pkt, ant, ch, t, p)
    want_re = np.transpose(lo, (2, 3, 4, 0, 1))
    want_im = np.transpose(hi, (2, 3, 4, 0, 1))
    if dtype == torch.int8:
        assert np.array_equal(re.cpu().numpy(), want_re) and np.array_equal(im.cpu().numpy(), want_im)
    else:
        np.testing.assert_allclose(re.float().cpu().numpy(), want_re * np.float32(0.05), rtol=1e-3)
        np.testing.assert_allclose(im.float().cpu().numpy(), want_im * np.float32(0.05), rtol=1e-3)


def test_unpack_rejects_wrong_size(small):
    cfg, sc, raw = small
    with pytest.raises(ValueError):
        unpack_int4_split(raw[:10], shape=(128, 16, 64))


@pytest.mark.parametrize("fuse,csf", [(False, 1), (True, 1), (True, 4)])
def test_fast_corr_compute_split_exact(small, fuse, csf):
    cfg, sc, raw = small
    re, im = unpack_int4_split(raw, out_dtype=torch.int8, shape=(128, 16, 64))
    k = FastCorrKernel(device="cuda", t_int_fast_native=8, nants=16, nchan=64, accum_dtype=torch.int32)
    got = k.compute_split(re, im, fuse_stokes_i=fuse, chan_sum_factor=csf).cpu().numpy()
    want = ocorr.correlate(raw, n_pkt=128, n_ant=16, n_chan=64, t_int=8, csf=csf, fuse_pol=fuse)
    want = np.transpose(want, (0, 2, 1, 3) if fuse else (0, 3, 1, 2, 4))   # -> (fv, nbase, nch, [pol,] 2)
    assert np.array_equal(got.astype(np.int64), want)
    raw_t = torch.from_numpy(raw).cuda()
    got_raw = k.compute_raw(raw_t, fuse_stokes_i=fuse, chan_sum_factor=csf).cpu().numpy()
    assert np.array_equal(got_raw, got)


def test_fast_corr_float_path_and_stokes(small):
    cfg, sc, raw = small
    re, im = unpack_int4_split(raw, out_dtype=torch.float32, scale=0.05, shape=(128, 16, 64))
    k = FastCorrKernel(device="cuda", t_int_fast_native=8, nants=16, nchan=64)
    v2 = k.compute_split(re, im)                            # (fv, nbase, nch, 2) complex64
    vi = k.compute_split(re, im, fuse_stokes_i=True)
    # float planes: fp16 operands fp16(0.05 k), float32 accumulation
    xr, xi = ocorr.cal_operands(raw, n_pkt=128, n_ant=16, n_chan=64, scale=0.05)
    w, bound = ocorr.correlate_float(xr, xi, t_int=8, fuse_pol=False)
    w, bound = w.transpose(0, 3, 1, 2), bound.transpose(0, 3, 1, 2)
    assert np.all(np.abs(v2.cpu().numpy() - w) <= 1e-5 * bound)
    np.testing.assert_allclose(stokes_i_pol_sum(v2).cpu().numpy(), vi.cpu().numpy(), rtol=1e-6, atol=1e-6)


def test_pattern_and_gridder(small):
    cfg, sc, raw = small
    f = np.asarray(cfg.chan_freqs_GHz)
    pat = build_pattern(sc.antpos_e, sc.antpos_n, chgroup=0, dec_deg=71.6, n_grid=64, chan_freqs_GHz=f)
    g = FastVisGridder.from_pattern(pat, sc.antpos_e, sc.antpos_n, device="cuda")
    vis = ocorr.correlate(raw, n_pkt=128, n_ant=16, n_chan=64, t_int=8)            # (fv, ch, nbase, 2)
    v_i32 = torch.from_numpy(np.ascontiguousarray(vis.transpose(0, 2, 1, 3)).astype(np.int32)).cuda()
    got = g.compute(v_i32).cpu().numpy()
    rows, cols, cell = oimage.cell_map(sc.antpos_e, sc.antpos_n, f, 64)
    assert cell == pat.cell_lambda
    vc = vis[..., 0] + 1j * vis[..., 1]
    G = oimage.grid_plane(np.transpose(vc, (0, 2, 1)), rows, cols, 64).reshape(vis.shape[0], -1)
    lin = pat.ix_row.astype(np.int64) * 64 + pat.ix_col
    assert np.array_equal(got[..., 0], G[:, lin].real.astype(np.int64))
    assert np.array_equal(got[..., 1], G[:, lin].imag.astype(np.int64))
    assert np.abs(G).sum() == np.abs(G[:, lin]).sum()          # no value outside the pattern
    gc = g.compute(torch.complex(v_i32[..., 0].float(), v_i32[..., 1].float())).cpu().numpy()
    np.testing.assert_allclose(gc, G[:, lin], rtol=1e-6, atol=1e-3)
    w = g.cell_weights.cpu().numpy()
    assert w.sum() == (rows >= 0).sum()


@pytest.mark.parametrize("N,B", [(64, 5), (256, 3)])
def test_dirty_image_from_uv_grid(N, B):
    rng = np.random.default_rng(N)
    G = (rng.standard_normal((B, N, N)) + 1j * rng.standard_normal((B, N, N))).astype(np.complex64)
    got = dirty_image_from_uv_grid(torch.from_numpy(G).cuda(), prescale=0.5).cpu().numpy()
    want = np.fft.fftshift(np.fft.ifft2(0.5 * G.astype(np.complex128), axes=(-2, -1)).real, axes=(-2, -1))
    assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()
    # numpy in, numpy out (the reference's numpy path), same HIP imager
    got_np = dirty_image_from_uv_grid(G, prescale=0.5)
    assert isinstance(got_np, np.ndarray) and got_np.dtype == np.float32 and np.array_equal(got_np, got)
    assert dirty_image_from_uv_grid(G, out_dtype=np.float64).dtype == np.float64


def test_combine_chgroups_vs_numpy():
    rng = np.random.default_rng(4)
    T, N, G = 40, 16, 5
    streams = {g: (rng.standard_normal((T, N, N)) + 1j * rng.standard_normal((T, N, N))).astype(np.complex64)
               for g in (0, 1, 3, 4)}                           # stream 2 missing
    shifts = np.array([7, 5, 3, 2, 0], dtype=np.int32)
    got = combine_chgroups(per_chgroup_streams={g: torch.from_numpy(s).cuda() for g, s in streams.items()},
                           time_shift_per_chgroup=shifts, t_window=(3, 38)).cpu().numpy()
    want = np.zeros((35, N, N), dtype=np.complex128)
    for g, s in streams.items():
        for i, t in enumerate(range(3, 38)):
            if 0 <= t - shifts[g] < T:
                want[i] += s[t - shifts[g]]
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("w", [1, 2, 5, 8, 32])
def test_boxcar_via_cumsum(w):
    rng = np.random.default_rng(20260505)
    x = rng.standard_normal((64, 3, 4, 4)).astype(np.float32)
    got = boxcar_via_cumsum(torch.from_numpy(x).cuda(), axis=0, width=w)
    got = got.cpu().numpy()
    np.testing.assert_allclose(got, osearch.boxcar_np(x.astype(np.float64), w), rtol=1e-5, atol=1e-5)
    y = rng.standard_normal((4, 32, 5)).astype(np.float32)
    got1 = boxcar_via_cumsum(torch.from_numpy(y).cuda(), axis=1, width=min(w, 32)).cpu().numpy()
    np.testing.assert_allclose(got1, osearch.boxcar_np(y.astype(np.float64), min(w, 32), axis=1),
                               rtol=1e-5, atol=1e-5)


def test_multi_boxcar_argmax():
    rng = np.random.default_rng(7)
    T, F, H, W, maxw = 96, 2, 4, 4, 16
    x = rng.standard_normal((T, F, H, W))
    pl = maxw // 2
    cs = np.zeros((T + maxw, F, H, W))
    cs[1 + pl:1 + pl + T] = np.cumsum(x, axis=0)
    cs[1 + pl + T:] = cs[pl + T]
    widths = np.array([1, 2, 4, 8, 16], dtype=np.int32)
    offs = (pl - widths // 2).astype(np.int32)
    sinv = (1.0 / np.sqrt(widths)).astype(np.float32)
    cs32 = cs.astype(np.float32)
    mx, win = multi_boxcar_argmax(torch.from_numpy(cs32).cuda(), widths=torch.from_numpy(widths).cuda(),
                                  offsets=torch.from_numpy(offs).cuda(), sigma_inv=torch.from_numpy(sinv).cuda(),
                                  n_out=T - maxw)
    scores = np.stack([(cs32[np.arange(T - maxw)[:, None] + o + w].reshape(T - maxw, -1)
                        - cs32[np.arange(T - maxw)[:, None] + o].reshape(T - maxw, -1)) * s
                       for w, o, s in zip(widths, offs, sinv)])             # (K, n_out, inner)
    np.testing.assert_allclose(mx.cpu().numpy().reshape(T - maxw, -1), scores.max(axis=0), rtol=1e-6, atol=1e-6)
    assert np.array_equal(win.cpu().numpy().reshape(T - maxw, -1), scores.argmax(axis=0))


@pytest.mark.parametrize("dtype,hist", [(torch.float16, 24), (torch.float16, 21), (torch.float32, 24), (torch.float32, 22)])
def test_ring_advance_matches_numpy(dtype, hist):
    """dsx_ring_advance over four blocks (64, 256, 128, 64 steps): the history
    rows hold the last `hist` steps of the previous contents followed by the
    new block, time-major per (subband, pixel).  hist = 24 takes the 16-byte
    paths (fp16 and fp32), hist = 21 / 22 the element paths; all are exact."""
    from dsx import _lib
    n_sub, n_pix, cap = 3, 128, 256
    row_len = hist + cap
    code = _lib.DSX_F16 if dtype == torch.float16 else _lib.DSX_F32
    rings = [torch.zeros((n_sub, n_pix, row_len), dtype=dtype, device="cuda") for _ in range(2)]
    want = np.zeros((n_sub, n_pix, row_len))
    rng = np.random.default_rng(5)
    cur, prev_len = 0, 0
    for n_new in (64, 256, 128, 64):
        img = rng.standard_normal((n_sub, n_new, n_pix)).astype(np.float32)
        img_t = torch.from_numpy(img).to(dtype).cuda().contiguous()
        img = img_t.float().cpu().numpy()
        _lib.call("dsx_ring_advance", _lib.ptr(rings[cur]), _lib.ptr(rings[1 - cur]), _lib.ptr(img_t), n_sub,
                  n_pix, hist, cap, prev_len, n_new, code, None)
        new = np.zeros_like(want)
        new[:, :, :hist] = want[:, :, prev_len:prev_len + hist]
        new[:, :, hist:hist + n_new] = np.transpose(img, (0, 2, 1))
        want = new
        cur, prev_len = 1 - cur, n_new
        torch.cuda.synchronize()
        got = rings[cur].float().cpu().numpy()
        assert np.array_equal(got[:, :, :hist + n_new], want[:, :, :hist + n_new]), f"block of {n_new} steps"


@pytest.mark.parametrize("dtype,n_inner,f", [(torch.float32, 256, 2), (torch.float32, 256, 4), (torch.float32, 6, 2),
                                             (torch.float16, 256, 4)])
def test_downsample_time_matches_numpy(dtype, n_inner, f):
    """dsx_downsample_time: out[s][t][k] = sum_j in[s][f t + j][k] accumulated in
    float32 in j order, stored in the input dtype -- bit-exact against the
    same arithmetic in numpy.  n_inner 256 (float32) takes the 16-byte path,
    6 the element path."""
    from dsx import _lib
    n_planes, n_in = 3, 64
    rng = np.random.default_rng(9)
    x = torch.from_numpy(rng.standard_normal((n_planes, n_in, n_inner)).astype(np.float32)).to(dtype).cuda()
    out = torch.empty((n_planes, n_in // f, n_inner), dtype=dtype, device="cuda")
    code = _lib.DSX_F16 if dtype == torch.float16 else _lib.DSX_F32
    _lib.call("dsx_downsample_time", _lib.ptr(x), _lib.ptr(out), n_planes, n_in, n_inner, f, code, None)
    torch.cuda.synchronize()
    xs = x.float().cpu().numpy().reshape(n_planes, n_in // f, f, n_inner)
    acc = np.zeros((n_planes, n_in // f, n_inner), np.float32)
    for j in range(f):
        acc = (acc + xs[:, :, j]).astype(np.float32)
    want = acc.astype(np.float16).astype(np.float32) if dtype == torch.float16 else acc
    assert np.array_equal(out.float().cpu().numpy(), want)


@pytest.mark.parametrize("dtype,row_len", [(torch.float32, 2300), (torch.float16, 4500)])
def test_dedisp_prime_first_block_means(dtype, row_len):
    """dsx_dedisp_prime: for DMs without state the written mean is the float64
    average of the DM's complete samples of the block (summed here by brute
    force); DMs with state keep their mean.  Rows longer than one LDS chunk
    (4500 > 2048) and shifts reaching back to position 0 are covered."""
    from dsx import _lib
    rng = np.random.default_rng(8)
    n_sub, n_pix, n_dm, hist = 5, 6, 300, 200
    T0 = 10_000
    L = row_len - hist
    c_lo, n_c = T0 - 64, L
    sh = rng.integers(0, hist - 64 + 1, size=(n_sub, n_dm)).astype(np.int32)
    sh[:, 0] = 0
    sh[0, 1] = hist - 64
    vf = sh.max(axis=0).astype(np.int32)
    vf[5] = T0 + 100                     # complete only part way into the block
    vf[6] = T0 + L                       # no complete sample in this block
    n_prev = np.zeros(n_dm)
    n_prev[::7] = 50.0                   # these DMs have state
    ring = (rng.standard_normal((n_sub, n_pix, row_len)) + 3.0).astype(np.float32)
    ring_t = torch.from_numpy(ring).to(dtype).cuda().contiguous()
    ring = ring_t.float().cpu().numpy().astype(np.float64)
    mean0 = rng.standard_normal((n_pix, n_dm))
    mean_t = torch.from_numpy(mean0.copy()).cuda()
    sh_t, vf_t, np_t = (torch.from_numpy(x).cuda() for x in (sh, vf, n_prev))     # alive until the kernel ran
    _lib.call("dsx_dedisp_prime", _lib.ptr(ring_t), _lib.DSX_F16 if dtype == torch.float16 else _lib.DSX_F32,
              n_sub, n_pix, hist, row_len, T0, _lib.ptr(sh_t), n_dm, _lib.ptr(vf_t), _lib.ptr(np_t), c_lo, n_c,
              _lib.ptr(mean_t), None)
    torch.cuda.synchronize()
    got = mean_t.cpu().numpy()
    want = mean0.copy()
    for d in range(n_dm):
        if n_prev[d] != 0:
            continue
        t_b = max(int(vf[d]), c_lo)
        if t_b >= c_lo + n_c:
            continue
        t = np.arange(t_b, c_lo + n_c)
        D = sum(ring[s][:, hist + t - sh[s, d] - T0] for s in range(n_sub))      # (n_pix, n_t)
        want[:, d] = D.mean(axis=1)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)
    assert np.array_equal(got[:, 6], mean0[:, 6]) and np.array_equal(got[:, ::7], mean0[:, ::7])
