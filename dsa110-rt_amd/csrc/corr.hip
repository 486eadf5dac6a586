// 4-bit unpack and the int8-MFMA correlator (gfx950).
//
// One work item = (step, output channel).  Its samples (t_int native samples
// x csf channels x 2 pols per antenna) are unpacked once into LDS as int8
// rows [R_p0 | R_p1 | I_p0 | I_p1] (each part kq bytes, zero padded), then
// every upper-triangle 16x16 antenna tile pair is reduced with
// v_mfma_i32_16x16x64_i8:
//   re  += R_b . R_a + I_b . I_a
//   imp += R_b . I_a,   imn += I_b . R_a,   im = imp - imn
// so V[b][a] = sum conj(E_b) E_a exactly in int32.  The K order inside a
// part is irrelevant to the sum; A and B operands use the same per-lane
// byte ranges, so the MFMA's internal K mapping cancels.
#include "dsx_internal.h"

namespace dsx {

constexpr int CORR_THREADS = 256;

// Sign-extend the low nibble of each byte of x (bits 0-3 per byte) to int8.
__device__ __forceinline__ uint32_t sext_nib4(uint32_t x) {
    const uint32_t s = x & 0x08080808u;
    return x | (s * 0x1Eu);   // copies bit 3 into bits 4..7 of the same byte
}

struct CorrParams {
    const uint8_t *raw;
    const int8_t *re8;
    const int8_t *im8;
    int32_t *vis;
    int n_pkt, n_ant, n_chan, t_int, csf, fuse_pol, in_layout, out_layout;
    int vec4;   // raw input read 4 channels (16 bytes) at a time: csf, n_chan multiples of 4
    int n_steps, n_cout, nbase, kq, row_stride, n_tiles, n_pairs, n_items;
    // stage 1: output step t reads channel c at step t + chan_shift[c] - delay of the
    // raw block, steps before the block from `tail` (the previous block's last
    // `delay` steps); NULL chan_shift: no stage 1
    const int32_t *chan_shift;
    const uint8_t *tail;
    int delay;
};

constexpr int CORR_PF = 12;   // prefetched 16-byte raw words per thread and item

__device__ __forceinline__ void store_vis(const CorrParams &P, int step, int co, int q, int bls,
                                          int re, int im) {
    const int npol_out = P.fuse_pol ? 1 : 2;
    int64_t idx;
    if (P.out_layout == 1)
        idx = (((int64_t)step * P.n_cout + co) * npol_out + q) * P.nbase + bls;
    else
        idx = (((int64_t)step * P.nbase + bls) * P.n_cout + co) * npol_out + q;
    int2 v;
    v.x = re;
    v.y = im;
    *reinterpret_cast<int2 *>(P.vis + 2 * idx) = v;
}

// Stage 1: the 4-byte word of (channel ch, antenna, packet pkt_l of output step
// `step`) read `shift` steps later in the stream delayed by P.delay steps: from
// the raw block, from the carried tail (the previous block's last P.delay
// steps) before it, zero outside both (the host keeps shifts in [0, delay]).
__device__ __forceinline__ uint32_t stage1_word(const CorrParams &P, int step, int pkt_l, int ant, int ch,
                                                int shift) {
    const int ppf = P.t_int / 2;
    const int64_t src = (int64_t)(step + shift - P.delay) * ppf + pkt_l;
    const int64_t n_tail = (int64_t)P.delay * ppf;
    if (src >= 0 && src < P.n_pkt)
        return *reinterpret_cast<const uint32_t *>(P.raw + (((src * P.n_ant + ant) * P.n_chan + ch) << 2));
    if (src < 0 && src >= -n_tail)
        return *reinterpret_cast<const uint32_t *>(
            P.tail + ((((src + n_tail) * P.n_ant + ant) * P.n_chan + ch) << 2));
    return 0u;
}

// The same for 4 channels ch0 .. ch0 + 3 sharing one shift: one 16-byte load.
__device__ __forceinline__ uint4 stage1_word4(const CorrParams &P, int step, int pkt_l, int ant, int ch0,
                                              int shift) {
    const int ppf = P.t_int / 2;
    const int64_t src = (int64_t)(step + shift - P.delay) * ppf + pkt_l;
    const int64_t n_tail = (int64_t)P.delay * ppf;
    if (src >= 0 && src < P.n_pkt)
        return *reinterpret_cast<const uint4 *>(P.raw + (((src * P.n_ant + ant) * P.n_chan + ch0) << 2));
    if (src < 0 && src >= -n_tail)
        return *reinterpret_cast<const uint4 *>(
            P.tail + ((((src + n_tail) * P.n_ant + ant) * P.n_chan + ch0) << 2));
    return make_uint4(0u, 0u, 0u, 0u);
}

__global__ __launch_bounds__(CORR_THREADS) void corr_i8_kernel(CorrParams P) {
    extern __shared__ __align__(16) int8_t lds[];
    const int tid = threadIdx.x;
    const int lds_bytes = P.n_tiles * 16 * P.row_stride;
    for (int i = tid * 16; i < lds_bytes; i += CORR_THREADS * 16)
        *reinterpret_cast<v4i *>(lds + i) = v4i{0, 0, 0, 0};
    __syncthreads();

    const int ppf = P.t_int / 2;            // packets per step
    const int kq = P.kq;
    const int wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, li = lane & 15;

    // vec4 unpack: thread -> (packet, antenna, 4-channel group) slots, stepped by
    // CORR_THREADS slots with carries instead of divisions
    const int q4 = P.csf >> 2;
    const int d_grp = P.vec4 ? CORR_THREADS % q4 : 0, d_row = P.vec4 ? CORR_THREADS / q4 : 0;
    const int grp0 = P.vec4 ? tid % q4 : 0, row0 = P.vec4 ? tid / q4 : 0;
    const int ant0 = P.vec4 ? row0 % P.n_ant : 0, pkl0 = P.vec4 ? row0 / P.n_ant : 0;

    // the 16 raw bytes (4 channels x 2 steps x 2 pols) of one slot of output step `step`
    auto load_slot = [&](int step, int co, int grp, int ant, int pkt_l) -> uint4 {
        const int ch0 = co * P.csf + 4 * grp;
        if (P.chan_shift == nullptr) {
            const int pkt = step * ppf + pkt_l;
            return *reinterpret_cast<const uint4 *>(P.raw + ((((int64_t)pkt * P.n_ant + ant) * P.n_chan + ch0) << 2));
        }
        // stage 1: one 16-byte load when the 4 channels share their shift
        // (channel groups of 4 or more), else one word per channel
        const int s0 = P.chan_shift[ch0], s1 = P.chan_shift[ch0 + 1];
        const int s2 = P.chan_shift[ch0 + 2], s3 = P.chan_shift[ch0 + 3];
        if (s0 == s1 && s0 == s2 && s0 == s3) return stage1_word4(P, step, pkt_l, ant, ch0, s0);
        return make_uint4(stage1_word(P, step, pkt_l, ant, ch0, s0), stage1_word(P, step, pkt_l, ant, ch0 + 1, s1),
                          stage1_word(P, step, pkt_l, ant, ch0 + 2, s2), stage1_word(P, step, pkt_l, ant, ch0 + 3, s3));
    };
    // sign-extend the nibbles of one slot and store them into the antenna's LDS
    // rows; per part 8 bytes (4 channels x 2 steps) per LDS store
    auto unpack_slot = [&](const uint4 w4, int grp, int ant, int pkt_l) {
        const uint32_t wv[4] = {w4.x, w4.y, w4.z, w4.w};
        uint32_t r0[2], r1[2], m0[2], m1[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            uint32_t pr0[2], pr1[2], pm0[2], pm1[2];
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const uint32_t word = wv[2 * h + c];      // bytes: (t0 p0, t0 p1, t1 p0, t1 p1)
                const uint32_t r = sext_nib4(word & 0x0F0F0F0Fu);
                const uint32_t m = sext_nib4((word >> 4) & 0x0F0F0F0Fu);
                pr0[c] = (r & 0xFFu) | ((r >> 8) & 0xFF00u);
                pr1[c] = ((r >> 8) & 0xFFu) | ((r >> 16) & 0xFF00u);
                pm0[c] = (m & 0xFFu) | ((m >> 8) & 0xFF00u);
                pm1[c] = ((m >> 8) & 0xFFu) | ((m >> 16) & 0xFF00u);
            }
            r0[h] = pr0[0] | (pr0[1] << 16);
            r1[h] = pr1[0] | (pr1[1] << 16);
            m0[h] = pm0[0] | (pm0[1] << 16);
            m1[h] = pm1[0] | (pm1[1] << 16);
        }
        const int k = (pkt_l * P.csf + 4 * grp) * 2;           // 8-byte aligned
        int8_t *row = lds + ant * P.row_stride;
        *reinterpret_cast<uint2 *>(row + k) = make_uint2(r0[0], r0[1]);
        *reinterpret_cast<uint2 *>(row + kq + k) = make_uint2(r1[0], r1[1]);
        *reinterpret_cast<uint2 *>(row + 2 * kq + k) = make_uint2(m0[0], m0[1]);
        *reinterpret_cast<uint2 *>(row + 3 * kq + k) = make_uint2(m1[0], m1[1]);
    };
    // prefetch: when an item's slots fit CORR_PF per thread, the next item's raw
    // words are loaded into registers while the current item's MFMAs run
    const int n_slot = P.vec4 ? ppf * P.n_ant * q4 : 0;
    const bool pf = P.vec4 && n_slot <= CORR_PF * CORR_THREADS;
    uint4 wbuf[CORR_PF];
    auto fetch = [&](int it_f) {
        const int st_f = it_f / P.n_cout;
        const int co_f = it_f - st_f * P.n_cout;
#pragma unroll
        for (int k = 0; k < CORR_PF; ++k) {
            const int sl = tid + k * CORR_THREADS;
            if (sl < n_slot) {
                const int row = sl / q4;
                wbuf[k] = load_slot(st_f, co_f, sl - row * q4, row % P.n_ant, row / P.n_ant);
            }
        }
    };
    if (pf && (int)blockIdx.x < P.n_items) fetch(blockIdx.x);

    for (int item = blockIdx.x; item < P.n_items; item += gridDim.x) {
        const int step = item / P.n_cout;
        const int co = item - step * P.n_cout;
        // ---- unpack this (step, channel group) into LDS rows
        if (pf) {
#pragma unroll
            for (int k = 0; k < CORR_PF; ++k) {
                const int sl = tid + k * CORR_THREADS;
                if (sl < n_slot) {
                    const int row = sl / q4;
                    unpack_slot(wbuf[k], sl - row * q4, row % P.n_ant, row / P.n_ant);
                }
            }
        } else if (P.vec4) {
            int grp = grp0, ant = ant0, pkt_l = pkl0;
            while (pkt_l < ppf) {
                unpack_slot(load_slot(step, co, grp, ant, pkt_l), grp, ant, pkt_l);
                grp += d_grp;
                const int carry = grp >= q4 ? 1 : 0;
                grp -= carry * q4;
                ant += d_row + carry;
                while (ant >= P.n_ant) {
                    ant -= P.n_ant;
                    ++pkt_l;
                }
            }
        } else if (P.in_layout == 0) {
            const int n_words = ppf * P.n_ant * P.csf;
            for (int w = tid; w < n_words; w += CORR_THREADS) {
                const int ch_l = w % P.csf;
                const int rest = w / P.csf;
                const int ant = rest % P.n_ant;
                const int pkt_l = rest / P.n_ant;
                const int pkt = step * ppf + pkt_l;
                const int ch = co * P.csf + ch_l;
                uint32_t word;
                if (P.chan_shift == nullptr) {
                    word = *reinterpret_cast<const uint32_t *>(
                        P.raw + ((((int64_t)pkt * P.n_ant + ant) * P.n_chan + ch) << 2));
                } else {
                    // stage 1: this channel's samples of step t + shift - delay
                    word = stage1_word(P, step, pkt_l, ant, ch, P.chan_shift[ch]);
                }
                // bytes: (t0 p0, t0 p1, t1 p0, t1 p1)
                const uint32_t r = sext_nib4(word & 0x0F0F0F0Fu);
                const uint32_t m = sext_nib4((word >> 4) & 0x0F0F0F0Fu);
                const int k = (pkt_l * P.csf + ch_l) * 2;           // t0 at k, t1 at k + 1
                int8_t *row = lds + ant * P.row_stride;
                const uint16_t r0 = (uint16_t)((r & 0xFFu) | ((r >> 8) & 0xFF00u));
                const uint16_t r1 = (uint16_t)(((r >> 8) & 0xFFu) | ((r >> 16) & 0xFF00u));
                const uint16_t m0 = (uint16_t)((m & 0xFFu) | ((m >> 8) & 0xFF00u));
                const uint16_t m1 = (uint16_t)(((m >> 8) & 0xFFu) | ((m >> 16) & 0xFF00u));
                *reinterpret_cast<uint16_t *>(row + k) = r0;
                *reinterpret_cast<uint16_t *>(row + kq + k) = r1;
                *reinterpret_cast<uint16_t *>(row + 2 * kq + k) = m0;
                *reinterpret_cast<uint16_t *>(row + 3 * kq + k) = m1;
            }
        } else {
            // split int8 planes (n_chan, 2t, 2p, n_pkt, n_ant), antenna fastest
            const int n_el = P.csf * 4 * ppf * P.n_ant;
            for (int e = tid; e < n_el; e += CORR_THREADS) {
                const int ant = e % P.n_ant;
                int rest = e / P.n_ant;
                const int pkt_l = rest % ppf;
                rest /= ppf;
                const int tp = rest & 3;                 // t * 2 + p
                const int ch_l = rest >> 2;
                const int t = tp >> 1, p = tp & 1;
                const int ch = co * P.csf + ch_l;
                const int pkt = step * ppf + pkt_l;
                const int64_t src = ((((int64_t)ch * 4 + tp) * P.n_pkt) + pkt) * P.n_ant + ant;
                const int k = (pkt_l * P.csf + ch_l) * 2 + t;
                int8_t *row = lds + ant * P.row_stride;
                row[p * kq + k] = P.re8[src];
                row[2 * kq + p * kq + k] = P.im8[src];
            }
        }
        __syncthreads();
        if (pf && item + (int)gridDim.x < P.n_items) fetch(item + gridDim.x);

        // ---- MFMA over upper-triangle tile pairs (b tile <= a tile)
        for (int pair = wave; pair < P.n_pairs; pair += CORR_THREADS / 64) {
            int at = 0;
            while ((at + 1) * (at + 2) / 2 <= pair) ++at;
            const int bt = pair - at * (at + 1) / 2;
            const int8_t *rowA = lds + (bt * 16 + li) * P.row_stride + g * 16;   // b antennas
            const int8_t *rowB = lds + (at * 16 + li) * P.row_stride + g * 16;   // a antennas
            v4i re[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
            v4i imp[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
            v4i imn[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
            // each polarisation in its own accumulators (a compile-time index: a
            // runtime one makes every MFMA select its accumulator); fused Stokes I
            // adds the two at the store, the same integer sums
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                for (int kc = 0; kc < kq; kc += 64) {
                    const int off = q * kq + kc;
                    const v4i aR = *reinterpret_cast<const v4i *>(rowA + off);
                    const v4i aI = *reinterpret_cast<const v4i *>(rowA + 2 * kq + off);
                    const v4i bR = *reinterpret_cast<const v4i *>(rowB + off);
                    const v4i bI = *reinterpret_cast<const v4i *>(rowB + 2 * kq + off);
                    re[q] = __builtin_amdgcn_mfma_i32_16x16x64_i8(aR, bR, re[q], 0, 0, 0);
                    re[q] = __builtin_amdgcn_mfma_i32_16x16x64_i8(aI, bI, re[q], 0, 0, 0);
                    imp[q] = __builtin_amdgcn_mfma_i32_16x16x64_i8(aR, bI, imp[q], 0, 0, 0);
                    imn[q] = __builtin_amdgcn_mfma_i32_16x16x64_i8(aI, bR, imn[q], 0, 0, 0);
                }
            }
            const int a = at * 16 + li;
            if (a < P.n_ant) {
                for (int v = 0; v < 4; ++v) {
                    const int b = bt * 16 + 4 * g + v;
                    if (b > a) continue;
                    const int bls = a * (a + 1) / 2 + b;
                    if (P.fuse_pol) {
                        store_vis(P, step, co, 0, bls, re[0][v] + re[1][v],
                                  (imp[0][v] + imp[1][v]) - (imn[0][v] + imn[1][v]));
                    } else {
                        store_vis(P, step, co, 0, bls, re[0][v], imp[0][v] - imn[0][v]);
                        store_vis(P, step, co, 1, bls, re[1][v], imp[1][v] - imn[1][v]);
                    }
                }
            }
        }
        __syncthreads();   // LDS rows are rewritten by the next item
    }
}

// On-wire bytes -> GEMM-layout planes (n_chan, 2t, 2p, n_pkt, n_ant).
template <typename T>
__global__ __launch_bounds__(256) void unpack_kernel(const uint8_t *raw, T *re, T *im, int n_pkt,
                                                     int n_ant, int n_chan, float scale) {
    // one thread per (pkt, chan, ant) word; ant fastest so writes along ant coalesce
    const int64_t n = (int64_t)n_pkt * n_chan * n_ant;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int ant = (int)(i % n_ant);
        const int64_t r = i / n_ant;
        const int ch = (int)(r % n_chan);
        const int pkt = (int)(r / n_chan);
        const uint32_t word = *reinterpret_cast<const uint32_t *>(
            raw + ((((int64_t)pkt * n_ant + ant) * n_chan + ch) << 2));
        for (int tp = 0; tp < 4; ++tp) {
            const int b = (word >> (8 * tp)) & 0xFF;
            int lo = b & 0xF, hi = b >> 4;
            lo = lo >= 8 ? lo - 16 : lo;
            hi = hi >= 8 ? hi - 16 : hi;
            const int64_t o = (((int64_t)ch * 4 + tp) * n_pkt + pkt) * n_ant + ant;
            if constexpr (sizeof(T) == 1) {
                re[o] = (T)lo;
                im[o] = (T)hi;
            } else {
                re[o] = (T)((float)lo * scale);
                im[o] = (T)((float)hi * scale);
            }
        }
    }
}

}  // namespace dsx

using namespace dsx;

extern "C" int dsx_unpack_i4(const uint8_t *raw, void *re, void *im, int n_pkt, int n_ant,
                             int n_chan, int out_dtype, float scale, dsx_stream_t stream) {
    DSX_REQUIRE(raw && re && im, "dsx_unpack_i4: null pointer");
    DSX_REQUIRE(n_pkt > 0 && n_ant > 0 && n_chan > 0, "dsx_unpack_i4: bad shape");
    const int64_t n = (int64_t)n_pkt * n_chan * n_ant;
    const int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
    hipStream_t s = as_stream(stream);
    switch (out_dtype) {
        case DSX_I8:
            unpack_kernel<int8_t><<<blocks, 256, 0, s>>>(raw, (int8_t *)re, (int8_t *)im, n_pkt,
                                                         n_ant, n_chan, scale);
            break;
        case DSX_F16:
            unpack_kernel<_Float16><<<blocks, 256, 0, s>>>(raw, (_Float16 *)re, (_Float16 *)im,
                                                           n_pkt, n_ant, n_chan, scale);
            break;
        case DSX_F32:
            unpack_kernel<float><<<blocks, 256, 0, s>>>(raw, (float *)re, (float *)im, n_pkt,
                                                        n_ant, n_chan, scale);
            break;
        default:
            return set_error(DSX_EINVAL, "dsx_unpack_i4: out_dtype %d not supported", out_dtype);
    }
    return check_launch("dsx_unpack_i4");
}

static int corr_i8(const void *raw, const void *im, const uint8_t *tail, const int32_t *chan_shift, int delay,
                   int32_t *vis, int n_pkt, int n_ant, int n_chan, int t_int, int csf, int fuse_pol,
                   int input_layout, int out_layout, dsx_stream_t stream) {
    DSX_REQUIRE(raw && vis, "dsx_corr_i8: null pointer");
    DSX_REQUIRE(input_layout == 0 || input_layout == 1, "dsx_corr_i8: input_layout must be 0 or 1");
    DSX_REQUIRE(input_layout == 0 ? im == nullptr : im != nullptr,
                "dsx_corr_i8: im must be NULL for raw input and set for split input");
    DSX_REQUIRE(out_layout == 0 || out_layout == 1, "dsx_corr_i8: out_layout must be 0 or 1");
    DSX_REQUIRE(n_pkt > 0 && n_ant >= 2 && n_chan > 0, "dsx_corr_i8: bad shape");
    DSX_REQUIRE(t_int >= 2 && t_int % 2 == 0, "dsx_corr_i8: t_int must be even and >= 2");
    DSX_REQUIRE((2 * n_pkt) % t_int == 0, "dsx_corr_i8: 2*n_pkt must be a multiple of t_int");
    DSX_REQUIRE(csf >= 1 && n_chan % csf == 0, "dsx_corr_i8: csf must divide n_chan");
    if (n_ant > 512) return set_error(DSX_ERANGE, "dsx_corr_i8: n_ant=%d > 512", n_ant);
    CorrParams P{};
    P.raw = (const uint8_t *)raw;
    P.re8 = (const int8_t *)raw;
    P.im8 = (const int8_t *)im;
    P.vis = vis;
    P.n_pkt = n_pkt;
    P.n_ant = n_ant;
    P.n_chan = n_chan;
    P.t_int = t_int;
    P.csf = csf;
    P.fuse_pol = fuse_pol ? 1 : 0;
    P.in_layout = input_layout;
    P.out_layout = out_layout;
    P.vec4 = input_layout == 0 && csf % 4 == 0 && n_chan % 4 == 0 && ((uintptr_t)raw & 15) == 0 &&
             ((uintptr_t)tail & 15) == 0;
    P.chan_shift = chan_shift;
    P.tail = tail;
    P.delay = delay;
    P.n_steps = 2 * n_pkt / t_int;
    P.n_cout = n_chan / csf;
    P.nbase = n_ant * (n_ant + 1) / 2;
    P.kq = ((t_int * csf + 63) / 64) * 64;
    P.row_stride = 4 * P.kq + 16;
    P.n_tiles = (n_ant + 15) / 16;
    P.n_pairs = P.n_tiles * (P.n_tiles + 1) / 2;
    P.n_items = P.n_steps * P.n_cout;
    const size_t lds = (size_t)P.n_tiles * 16 * P.row_stride;
    if (lds > 160 * 1024)
        return set_error(DSX_ERANGE, "dsx_corr_i8: %zu B of LDS needed (t_int*csf too large)", lds);
    const int grid = std::min(P.n_items, 4096);
    (void)hipFuncSetAttribute((const void *)corr_i8_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
    corr_i8_kernel<<<grid, CORR_THREADS, lds, as_stream(stream)>>>(P);
    return check_launch("dsx_corr_i8");
}

extern "C" int dsx_corr_i8(const void *raw, const void *im, int32_t *vis, int n_pkt, int n_ant,
                           int n_chan, int t_int, int csf, int fuse_pol, int input_layout,
                           int out_layout, dsx_stream_t stream) {
    return corr_i8(raw, im, nullptr, nullptr, 0, vis, n_pkt, n_ant, n_chan, t_int, csf, fuse_pol,
                   input_layout, out_layout, stream);
}

extern "C" int dsx_corr_i8_stage1(const uint8_t *raw, const uint8_t *tail, const int32_t *chan_shift,
                                  int delay, int32_t *vis, int n_pkt, int n_ant, int n_chan, int t_int,
                                  int csf, int fuse_pol, int out_layout, dsx_stream_t stream) {
    DSX_REQUIRE(raw && chan_shift && vis, "dsx_corr_i8_stage1: null pointer");
    DSX_REQUIRE(delay >= 0 && (delay == 0 || tail != nullptr), "dsx_corr_i8_stage1: delay %d needs a tail",
                delay);
    DSX_REQUIRE(t_int >= 2 && t_int % 2 == 0 && (int64_t)delay * (t_int / 2) <= n_pkt,
                "dsx_corr_i8_stage1: the %d-step delay exceeds the block", delay);
    return corr_i8(raw, nullptr, tail, chan_shift, delay, vis, n_pkt, n_ant, n_chan, t_int, csf, fuse_pol, 0,
                   out_layout, stream);
}
