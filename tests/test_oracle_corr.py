"""Pins of the correlator oracle (CPU).

Known answers from the reference's own tests: the exhaustive 256-byte nibble
table and pack/unpack round trip (tests/test_slow_corr_synth.py:172-211),
the sign convention E_1 = (1, 0), E_2 = (0, 1) -> V = +iK
(tests/test_slow_corr_synth.py:219-277), zero-in/zero-out and real
non-negative autos (tests/test_fast_corr_kernel.py:242-285), t_int = block
equals the block-integrated correlator (:182-234).  Plus an independent
complex-arithmetic brute force on tiny inputs.
"""
import numpy as np

from oracle import corr as ocorr


def _pack(re, im):
    """(S, A, C, P) int -> on-wire bytes (S/2, A, C, 2t, P)."""
    S, A, C, P = re.shape
    byte = ((re & 0xF) | ((im & 0xF) << 4)).astype(np.uint8)
    return np.ascontiguousarray(byte.reshape(S // 2, 2, A, C, P).transpose(0, 2, 3, 1, 4))


def test_nibble_table_exhaustive():
    raw = np.arange(256, dtype=np.uint8)
    lo, hi = ocorr.unpack_bytes(raw)
    for byte in range(256):
        l, h = byte & 0xF, (byte >> 4) & 0xF
        assert lo[byte] == (l - 16 if l >= 8 else l)
        assert hi[byte] == (h - 16 if h >= 8 else h)


def test_pack_unpack_roundtrip():
    rng = np.random.default_rng(0xCAFE)
    re = rng.integers(-8, 8, size=1024)
    im = rng.integers(-8, 8, size=1024)
    packed = ((re & 0xF) | ((im & 0xF) << 4)).astype(np.uint8)
    lo, hi = ocorr.unpack_bytes(packed)
    assert np.array_equal(lo, re) and np.array_equal(hi, im)


def test_upper_tri_order():
    a, b = ocorr.upper_tri(7)
    assert np.all(b <= a)
    assert np.array_equal(a * (a + 1) // 2 + b, np.arange(a.size))


def test_sign_known_answer():
    S, A, C = 32, 8, 3
    re = np.zeros((S, A, C, 2), dtype=np.int64)
    im = np.zeros_like(re)
    re[:, 3] = 1          # E_3 = (1, 0)
    im[:, 5] = 1          # E_5 = (0, 1)
    v = ocorr.correlate(_pack(re, im), n_pkt=S // 2, n_ant=A, n_chan=C, t_int=8)
    bls = 5 * 6 // 2 + 3
    assert np.all(v[:, :, bls, 0] == 0)
    assert np.all(v[:, :, bls, 1] == 8 * 2)     # +K, K = t_int samples x 2 pols


def test_zero_in_zero_out_and_autos():
    rng = np.random.default_rng(1)
    S, A, C = 16, 6, 4
    z = np.zeros((S, A, C, 2), dtype=np.int64)
    assert np.all(ocorr.correlate(_pack(z, z), n_pkt=S // 2, n_ant=A, n_chan=C, t_int=8) == 0)
    re = rng.integers(-7, 8, size=(S, A, C, 2))
    im = rng.integers(-7, 8, size=(S, A, C, 2))
    v = ocorr.correlate(_pack(re, im), n_pkt=S // 2, n_ant=A, n_chan=C, t_int=8)
    autos = [a * (a + 1) // 2 + a for a in range(A)]
    assert np.all(v[:, :, autos, 1] == 0)
    assert np.all(v[:, :, autos, 0] >= 0)


def test_brute_force_complex_loops():
    rng = np.random.default_rng(2)
    S, A, C, t_int, csf = 8, 4, 4, 4, 2
    re = rng.integers(-8, 8, size=(S, A, C, 2))
    im = rng.integers(-8, 8, size=(S, A, C, 2))
    v = ocorr.correlate(_pack(re, im), n_pkt=S // 2, n_ant=A, n_chan=C, t_int=t_int, csf=csf)
    E = re + 1j * im
    for st in range(S // t_int):
        for co in range(C // csf):
            k = 0
            for a in range(A):
                for b in range(a + 1):
                    acc = 0j
                    for t in range(st * t_int, (st + 1) * t_int):
                        for c in range(co * csf, (co + 1) * csf):
                            for p in range(2):
                                acc += np.conj(E[t, b, c, p]) * E[t, a, c, p]
                    assert v[st, co, k, 0] == acc.real and v[st, co, k, 1] == acc.imag
                    k += 1


def test_block_integration_is_sum_of_steps():
    rng = np.random.default_rng(3)
    S, A, C = 64, 5, 2
    re = rng.integers(-7, 8, size=(S, A, C, 2))
    im = rng.integers(-7, 8, size=(S, A, C, 2))
    raw = _pack(re, im)
    fine = ocorr.correlate(raw, n_pkt=S // 2, n_ant=A, n_chan=C, t_int=8)
    whole = ocorr.correlate(raw, n_pkt=S // 2, n_ant=A, n_chan=C, t_int=S)
    assert np.array_equal(fine.sum(axis=0), whole[0])


def test_pol_split_sums_to_stokes_i():
    rng = np.random.default_rng(4)
    S, A, C = 16, 4, 4
    re = rng.integers(-7, 8, size=(S, A, C, 2))
    im = rng.integers(-7, 8, size=(S, A, C, 2))
    raw = _pack(re, im)
    per = ocorr.correlate(raw, n_pkt=S // 2, n_ant=A, n_chan=C, t_int=8, fuse_pol=False)
    fused = ocorr.correlate(raw, n_pkt=S // 2, n_ant=A, n_chan=C, t_int=8)
    assert np.array_equal(per.sum(axis=2), fused)


def test_blas_twin_equals_einsum_form():
    """The float64-product twin used for 96-antenna oracle streams returns
    the einsum form's integers exactly (channel sums and both pols fused)."""
    rng = np.random.default_rng(5)
    n_pkt, n_ant, n_chan = 16, 7, 12
    raw = rng.integers(0, 256, size=(n_pkt, n_ant, n_chan, 2, 2), dtype=np.uint8)
    for t_int, csf in ((2, 1), (8, 3), (4, 12)):
        want = ocorr.correlate(raw, n_pkt=n_pkt, n_ant=n_ant, n_chan=n_chan, t_int=t_int, csf=csf)
        got = ocorr.correlate_blas(raw, n_pkt=n_pkt, n_ant=n_ant, n_chan=n_chan, t_int=t_int, csf=csf)
        assert got.dtype == np.int64 and np.array_equal(got, want), (t_int, csf)
    # extreme codes (-8 everywhere) reach the largest partial sums
    raw8 = np.full((n_pkt, n_ant, n_chan, 2, 2), 0x88, dtype=np.uint8)
    assert np.array_equal(ocorr.correlate_blas(raw8, n_pkt=n_pkt, n_ant=n_ant, n_chan=n_chan, t_int=8, csf=4),
                          ocorr.correlate(raw8, n_pkt=n_pkt, n_ant=n_ant, n_chan=n_chan, t_int=8, csf=4))


# ---- calibrated correlator (cal + RFI mask on operand load) ----

def _rand_raw(rng, S, A, C):
    re = np.clip(np.rint(rng.normal(0, 2.5, size=(S, A, C, 2))), -7, 7).astype(np.int64)
    im = np.clip(np.rint(rng.normal(0, 2.5, size=(S, A, C, 2))), -7, 7).astype(np.int64)
    return re, im, _pack(re, im)


def test_cal_operands_without_gains_are_fp16_voltages():
    rng = np.random.default_rng(11)
    re, im, raw = _rand_raw(rng, 16, 4, 3)
    xr, xi = ocorr.cal_operands(raw, n_pkt=8, n_ant=4, n_chan=3, scale=0.05)
    assert np.array_equal(xr, (re * np.float32(0.05)).astype(np.float32).astype(np.float16).astype(np.float64))
    xr1, xi1 = ocorr.cal_operands(raw, n_pkt=8, n_ant=4, n_chan=3, scale=1.0)
    assert np.array_equal(xr1, re) and np.array_equal(xi1, im)


def test_cal_known_answer_gain_product():
    """Gains exactly representable in fp16 (and their products with 4-bit
    codes): V_cal[a, b] = conj(g_b) g_a V[a, b] exactly (the CASA-style G
    application of slow_corr_kernel.py:459-502)."""
    rng = np.random.default_rng(12)
    S, A, C, t_int = 16, 5, 2, 8
    re, im, raw = _rand_raw(rng, S, A, C)
    gvals = np.array([1.0, 0.5, -0.25 + 0.5j, 0.75j, 1.5 - 0.25j])
    g = np.broadcast_to(gvals[:, None, None], (A, C, 2))           # (ant, chan, pol)
    cal_re = np.ascontiguousarray(g.real.transpose(1, 2, 0))       # (chan, pol, ant)
    cal_im = np.ascontiguousarray(g.imag.transpose(1, 2, 0))
    xr, xi = ocorr.cal_operands(raw, n_pkt=S // 2, n_ant=A, n_chan=C, scale=1.0, cal_re=cal_re, cal_im=cal_im)
    # operands are exact: x = g * E
    assert np.array_equal(xr + 1j * xi, (re + 1j * im) * gvals[None, :, None, None])
    v, bound = ocorr.correlate_float(xr, xi, t_int=t_int)
    v_int = ocorr.correlate(raw, n_pkt=S // 2, n_ant=A, n_chan=C, t_int=t_int)
    v_int = v_int[..., 0] + 1j * v_int[..., 1]
    a, b = ocorr.upper_tri(A)
    want = np.conj(gvals[b]) * gvals[a] * v_int
    assert np.array_equal(v, want)
    assert np.all(bound >= np.abs(v))


def test_cal_rfi_mask_and_time_keep_zero_their_samples():
    """A zero gain (masked (ant, chan, pol)) equals zeroing those voltages;
    a zero keep factor removes every packet of its accumulation."""
    rng = np.random.default_rng(13)
    S, A, C, t_int = 32, 4, 3, 8
    re, im, raw = _rand_raw(rng, S, A, C)
    cal_re = np.ones((C, 2, A), np.float32)
    cal_im = np.zeros((C, 2, A), np.float32)
    cal_re[1, 0, 2] = 0.0                         # mask ant 2, chan 1, pol 0
    keep = np.ones((4, C, 2), np.float32)
    keep[2, 0, 1] = 0.0                           # accumulation 2 (packets 8..11), chan 0, pol 1
    xr, xi = ocorr.cal_operands(raw, n_pkt=S // 2, n_ant=A, n_chan=C, scale=1.0, cal_re=cal_re,
                                cal_im=cal_im, time_keep=keep)
    re2, im2 = re.copy(), im.copy()
    re2[:, 2, 1, 0] = 0
    im2[:, 2, 1, 0] = 0
    re2[16:24, :, 0, 1] = 0                       # samples of packets 8..11
    im2[16:24, :, 0, 1] = 0
    assert np.array_equal(xr, re2) and np.array_equal(xi, im2)
    v, _ = ocorr.correlate_float(xr, xi, t_int=t_int, fuse_pol=False)
    w = ocorr.correlate(_pack(re2, im2), n_pkt=S // 2, n_ant=A, n_chan=C, t_int=t_int, fuse_pol=False)
    assert np.array_equal(v, w[..., 0] + 1j * w[..., 1])

