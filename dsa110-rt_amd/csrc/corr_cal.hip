// Calibrated correlator (gfx950): per-(channel, pol, antenna) complex gains,
// with the RFI mask folded in as zero gains and an optional per-accumulation
// time mask, applied to every voltage as it is loaded; fp16 MFMA with
// float32 accumulation.
//
// Operand of one sample (the arithmetic of src/dsart/x2/corr/gemm2x.py:158-168
// and its CPU twin gemm2x_reference, gemm2x.py:346-390):
//   v = fp16(code * scale)                (raw bytes; or the fp16 planes given)
//   g = gain[ch][p][ant] * keep[acc][ch][p]
//   x = fp16(v_re g_re - v_im g_im, v_re g_im + v_im g_re)
// in float32 with every product and sum rounded on its own (no fused
// multiply-add), rounded once to fp16.  Products of fp16 values are exact in
// float32, so V[bls] = sum conj(x_b) x_a differs from a float64 sum only by
// the float32 accumulation.  With unit gains and integer codes (scale 1)
// every partial sum is an integer below 2^24: the result equals the int8
// correlator's bit for bit.
//
// One work item = (step, output channel); its samples are staged in LDS in
// channel chunks as fp16 rows [R_p0 | R_p1 | I_p0 | I_p1] per antenna (each
// part kc values, zero padded to a multiple of 32), and every wave keeps the
// accumulators of its upper-triangle 16x16 tile pairs across the chunks:
//   re += R_b.R_a + I_b.I_a,  imp += R_b.I_a,  imn += I_b.R_a,  im = imp - imn
// with v_mfma_f32_16x16x32_f16.
#include "dsx_internal.h"

namespace dsx {

typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));

constexpr int CCAL_THREADS = 256;
constexpr int CCAL_WAVES = CCAL_THREADS / 64;
constexpr int CCAL_LDS_BUDGET = 80 * 1024;   // two work-groups per CU
constexpr int CCAL_PF = 6;                   // prefetched 16-byte raw words per thread and chunk

struct CalCorrParams {
    const uint8_t *raw;          // in_layout 0: on-wire (n_pkt, n_ant, n_chan, 2t, 2p) bytes
    const _Float16 *re16;        // in_layout 1: fp16 planes (n_chan, 2t, 2p, n_pkt, n_ant)
    const _Float16 *im16;
    const float *cal_re;         // (n_chan, 2, n_ant) gains, NULL: operands are the voltages
    const float *cal_im;
    const float *tkeep;          // (n_acc, n_chan, 2) keep factors, NULL: none
    float *vis;                  // float (re, im) pairs
    float scale;
    int n_pkt, n_ant, n_chan, t_int, csf, fuse_pol, in_layout, out_layout, n_acc;
    int n_steps, n_cout, nbase, csf_c, n_chunks, kc, row_stride, n_tiles, n_pairs, n_items;
    int vec4;                    // raw input read 4 channels (16 bytes) at a time: csf_c, n_chan multiples of 4
};

// x = v * g in float32, products and sums rounded separately
__device__ __forceinline__ float2 cal_mul(float vr, float vi, float gr, float gi) {
#pragma clang fp contract(off)
    const float a = vr * gr;
    const float b = vi * gi;
    const float c = vr * gi;
    const float d = vi * gr;
    return make_float2(a - b, c + d);
}

__device__ __forceinline__ void cal_operand(const CalCorrParams &P, int ch, int p, int ant, int pkt,
                                            float vr, float vi, _Float16 &xr, _Float16 &xi) {
    if (P.cal_re != nullptr) {
        const int64_t gi_idx = ((int64_t)ch * 2 + p) * P.n_ant + ant;
        float gr = P.cal_re[gi_idx], gi = P.cal_im[gi_idx];
        if (P.tkeep != nullptr) {
            const int acc = pkt / (P.n_pkt / P.n_acc);
            const float tk = P.tkeep[((int64_t)acc * P.n_chan + ch) * 2 + p];
            gr = gr * tk;
            gi = gi * tk;
        }
        const float2 x = cal_mul(vr, vi, gr, gi);
        xr = (_Float16)x.x;
        xi = (_Float16)x.y;
    } else {
        xr = (_Float16)vr;
        xi = (_Float16)vi;
    }
}

__device__ __forceinline__ float nib_lo(uint32_t b) { return (float)((int)((b & 0xFu) ^ 8u) - 8); }
__device__ __forceinline__ float nib_hi(uint32_t b) { return (float)((int)(((b >> 4) & 0xFu) ^ 8u) - 8); }

template <int NQ, int MAXP>
__global__ __launch_bounds__(CCAL_THREADS) void corr_cal_kernel(CalCorrParams P) {
    extern __shared__ __align__(16) _Float16 lds16[];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, li = lane & 15;
    const int ppf = P.t_int / 2;
    const int kc = P.kc, RS = P.row_stride;
    // zero once: padded K slots and the rows past the last antenna stay zero
    const int n_vec = P.n_tiles * 16 * RS / 8;
    for (int i = tid; i < n_vec; i += CCAL_THREADS) reinterpret_cast<v4i *>(lds16)[i] = v4i{0, 0, 0, 0};
    __syncthreads();

    // vec4 input: slot sl = (packet, 4-channel group, antenna) of a chunk, antenna
    // fastest so that a wave's gain loads read consecutive antennas; one 16-byte
    // load of raw bytes each
    const int g4 = P.csf_c >> 2;
    const int n_slot = P.vec4 ? ppf * P.n_ant * g4 : 0;
    const bool gain_fast = P.cal_re != nullptr && P.tkeep == nullptr;
    auto load_slot = [&](int sl, int ch0, int step) -> uint4 {
        asm volatile("" : "+v"(sl));   // keep the slot's index math out of the item loop
        const int ant = sl % P.n_ant;
        const int rest = sl / P.n_ant;
        const int grp = rest % g4;
        const int pkt = step * ppf + rest / g4;
        return *reinterpret_cast<const uint4 *>(
            P.raw + ((((int64_t)pkt * P.n_ant + ant) * P.n_chan + ch0 + 4 * grp) << 2));
    };
    // one slot's operands into the antenna's LDS row; each channel's two steps
    // stored as one fp16 pair per part
    auto store_slot = [&](const uint4 w4, int sl, int ch0, int step) {
        asm volatile("" : "+v"(sl));
        const int ant = sl % P.n_ant;
        const int rest = sl / P.n_ant;
        const int grp = rest % g4;
        const int pkt_l = rest / g4;
        const int pkt = step * ppf + pkt_l;
        const int chb = ch0 + 4 * grp;
        const uint32_t wv[4] = {w4.x, w4.y, w4.z, w4.w};
        _Float16 *row = lds16 + ant * RS;
        if (gain_fast) {
            // gains without a time mask: the slot's 8 complex gains loaded at once,
            // then the arithmetic of cal_operand
            float gr[4][2], gim[4][2];
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int p = 0; p < 2; ++p) {
                    const int64_t gi_idx = ((int64_t)(chb + c) * 2 + p) * P.n_ant + ant;
                    gr[c][p] = P.cal_re[gi_idx];
                    gim[c][p] = P.cal_im[gi_idx];
                }
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const uint32_t word = wv[c];
                const int k = (pkt_l * P.csf_c + 4 * grp + c) * 2;
#pragma unroll
                for (int p = 0; p < 2; ++p) {
                    const uint32_t b0 = (word >> (8 * p)) & 0xFFu;
                    const uint32_t b1 = (word >> (8 * (2 + p))) & 0xFFu;
                    const float vr0 = (float)(_Float16)(nib_lo(b0) * P.scale);
                    const float vi0 = (float)(_Float16)(nib_hi(b0) * P.scale);
                    const float vr1 = (float)(_Float16)(nib_lo(b1) * P.scale);
                    const float vi1 = (float)(_Float16)(nib_hi(b1) * P.scale);
                    const float2 x0 = cal_mul(vr0, vi0, gr[c][p], gim[c][p]);
                    const float2 x1 = cal_mul(vr1, vi1, gr[c][p], gim[c][p]);
                    *reinterpret_cast<h2 *>(row + p * kc + k) = h2{(_Float16)x0.x, (_Float16)x1.x};
                    *reinterpret_cast<h2 *>(row + 2 * kc + p * kc + k) = h2{(_Float16)x0.y, (_Float16)x1.y};
                }
            }
        } else {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const uint32_t word = wv[c];                       // (t0 p0, t0 p1, t1 p0, t1 p1)
                const int ch = chb + c;
                const int k = (pkt_l * P.csf_c + 4 * grp + c) * 2;
#pragma unroll
                for (int p = 0; p < 2; ++p) {
                    const uint32_t b0 = (word >> (8 * p)) & 0xFFu;         // t0
                    const uint32_t b1 = (word >> (8 * (2 + p))) & 0xFFu;   // t1
                    const float vr0 = (float)(_Float16)(nib_lo(b0) * P.scale);
                    const float vi0 = (float)(_Float16)(nib_hi(b0) * P.scale);
                    const float vr1 = (float)(_Float16)(nib_lo(b1) * P.scale);
                    const float vi1 = (float)(_Float16)(nib_hi(b1) * P.scale);
                    _Float16 xr0, xi0, xr1, xi1;
                    cal_operand(P, ch, p, ant, pkt, vr0, vi0, xr0, xi0);
                    cal_operand(P, ch, p, ant, pkt, vr1, vi1, xr1, xi1);
                    *reinterpret_cast<h2 *>(row + p * kc + k) = h2{xr0, xr1};
                    *reinterpret_cast<h2 *>(row + 2 * kc + p * kc + k) = h2{xi0, xi1};
                }
            }
        }
    };
    // prefetch: when a chunk's slots fit CCAL_PF per thread, the raw words of the
    // next (item, chunk) are loaded into registers while this chunk's MFMAs run
    const bool pf = P.vec4 && n_slot <= CCAL_PF * CCAL_THREADS;
    uint4 wbuf[CCAL_PF];
    auto fetch = [&](int it_f, int chunk_f) {
        const int st_f = it_f / P.n_cout;
        const int ch0_f = (it_f - st_f * P.n_cout) * P.csf + chunk_f * P.csf_c;
#pragma unroll
        for (int k = 0; k < CCAL_PF; ++k) {
            const int sl = tid + k * CCAL_THREADS;
            if (sl < n_slot) wbuf[k] = load_slot(sl, ch0_f, st_f);
        }
    };
    if (pf && (int)blockIdx.x < P.n_items) fetch(blockIdx.x, 0);

    for (int item = blockIdx.x; item < P.n_items; item += gridDim.x) {
        const int step = item / P.n_cout;
        const int co = item - step * P.n_cout;
        v4f acc[MAXP][NQ][3];
#pragma unroll
        for (int j = 0; j < MAXP; ++j)
#pragma unroll
            for (int q = 0; q < NQ; ++q)
#pragma unroll
                for (int c = 0; c < 3; ++c) acc[j][q][c] = v4f{0.f, 0.f, 0.f, 0.f};

        for (int chunk = 0; chunk < P.n_chunks; ++chunk) {
            const int ch0 = co * P.csf + chunk * P.csf_c;
            if (pf) {
#pragma unroll
                for (int k = 0; k < CCAL_PF; ++k) {
                    const int sl = tid + k * CCAL_THREADS;
                    if (sl < n_slot) store_slot(wbuf[k], sl, ch0, step);
                }
            } else if (P.vec4) {
                for (int sl = tid; sl < n_slot; sl += CCAL_THREADS) store_slot(load_slot(sl, ch0, step), sl, ch0, step);
            } else if (P.in_layout == 0) {
                // one 4-byte word (t0 p0, t0 p1, t1 p0, t1 p1) per (packet, antenna, channel)
                const int n_words = ppf * P.n_ant * P.csf_c;
                for (int w = tid; w < n_words; w += CCAL_THREADS) {
                    const int ch_l = w % P.csf_c;
                    const int rest = w / P.csf_c;
                    const int ant = rest % P.n_ant;
                    const int pkt_l = rest / P.n_ant;
                    const int pkt = step * ppf + pkt_l;
                    const int ch = ch0 + ch_l;
                    const uint32_t word = *reinterpret_cast<const uint32_t *>(
                        P.raw + ((((int64_t)pkt * P.n_ant + ant) * P.n_chan + ch) << 2));
                    const int k = (pkt_l * P.csf_c + ch_l) * 2;
                    _Float16 *row = lds16 + ant * RS;
#pragma unroll
                    for (int p = 0; p < 2; ++p) {
                        const uint32_t b0 = (word >> (8 * p)) & 0xFFu;         // t0
                        const uint32_t b1 = (word >> (8 * (2 + p))) & 0xFFu;   // t1
                        const float vr0 = (float)(_Float16)(nib_lo(b0) * P.scale);
                        const float vi0 = (float)(_Float16)(nib_hi(b0) * P.scale);
                        const float vr1 = (float)(_Float16)(nib_lo(b1) * P.scale);
                        const float vi1 = (float)(_Float16)(nib_hi(b1) * P.scale);
                        _Float16 xr0, xi0, xr1, xi1;
                        cal_operand(P, ch, p, ant, pkt, vr0, vi0, xr0, xi0);
                        cal_operand(P, ch, p, ant, pkt, vr1, vi1, xr1, xi1);
                        *reinterpret_cast<h2 *>(row + p * kc + k) = h2{xr0, xr1};
                        *reinterpret_cast<h2 *>(row + 2 * kc + p * kc + k) = h2{xi0, xi1};
                    }
                }
            } else {
                // fp16 planes, antenna fastest
                const int n_el = P.csf_c * 4 * ppf * P.n_ant;
                for (int e = tid; e < n_el; e += CCAL_THREADS) {
                    const int ant = e % P.n_ant;
                    int rest = e / P.n_ant;
                    const int pkt_l = rest % ppf;
                    rest /= ppf;
                    const int tp = rest & 3;                 // t * 2 + p
                    const int ch_l = rest >> 2;
                    const int t = tp >> 1, p = tp & 1;
                    const int ch = ch0 + ch_l;
                    const int pkt = step * ppf + pkt_l;
                    const int64_t src = ((((int64_t)ch * 4 + tp) * P.n_pkt) + pkt) * P.n_ant + ant;
                    _Float16 xr, xi;
                    cal_operand(P, ch, p, ant, pkt, (float)P.re16[src], (float)P.im16[src], xr, xi);
                    const int k = (pkt_l * P.csf_c + ch_l) * 2 + t;
                    _Float16 *row = lds16 + ant * RS;
                    row[p * kc + k] = xr;
                    row[2 * kc + p * kc + k] = xi;
                }
            }
            __syncthreads();
            if (pf) {
                if (chunk + 1 < P.n_chunks) fetch(item, chunk + 1);
                else if (item + (int)gridDim.x < P.n_items) fetch(item + gridDim.x, 0);
            }

            // ---- MFMA over this wave's upper-triangle tile pairs (b tile <= a tile)
#pragma unroll
            for (int j = 0; j < MAXP; ++j) {
                const int pair = wave + CCAL_WAVES * j;
                if (pair < P.n_pairs) {
                    int at = 0;
                    while ((at + 1) * (at + 2) / 2 <= pair) ++at;
                    const int bt = pair - at * (at + 1) / 2;
                    const _Float16 *rowA = lds16 + (bt * 16 + li) * RS + g * 8;   // b antennas
                    const _Float16 *rowB = lds16 + (at * 16 + li) * RS + g * 8;   // a antennas
#pragma unroll
                    for (int q = 0; q < 2; ++q) {
                        const int qo = NQ == 1 ? 0 : q;
                        for (int kb = 0; kb < kc; kb += 32) {
                            const int off = q * kc + kb;
                            const v8h aR = *reinterpret_cast<const v8h *>(rowA + off);
                            const v8h aI = *reinterpret_cast<const v8h *>(rowA + 2 * kc + off);
                            const v8h bR = *reinterpret_cast<const v8h *>(rowB + off);
                            const v8h bI = *reinterpret_cast<const v8h *>(rowB + 2 * kc + off);
                            acc[j][qo][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aR, bR, acc[j][qo][0], 0, 0, 0);
                            acc[j][qo][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aI, bI, acc[j][qo][0], 0, 0, 0);
                            acc[j][qo][1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aR, bI, acc[j][qo][1], 0, 0, 0);
                            acc[j][qo][2] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aI, bR, acc[j][qo][2], 0, 0, 0);
                        }
                    }
                }
            }
            __syncthreads();   // the next chunk rewrites the rows
        }

        // ---- store V[b][a] for b <= a
#pragma unroll
        for (int j = 0; j < MAXP; ++j) {
            const int pair = wave + CCAL_WAVES * j;
            if (pair >= P.n_pairs) continue;
            int at = 0;
            while ((at + 1) * (at + 2) / 2 <= pair) ++at;
            const int bt = pair - at * (at + 1) / 2;
            const int a = at * 16 + li;
            if (a >= P.n_ant) continue;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int b = bt * 16 + 4 * g + v;
                if (b > a) continue;
                const int bls = a * (a + 1) / 2 + b;
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    int64_t idx;
                    if (P.out_layout == 1)
                        idx = (((int64_t)step * P.n_cout + co) * NQ + q) * P.nbase + bls;
                    else
                        idx = (((int64_t)step * P.nbase + bls) * P.n_cout + co) * NQ + q;
                    *reinterpret_cast<float2 *>(P.vis + 2 * idx) =
                        make_float2(acc[j][q][0][v], acc[j][q][1][v] - acc[j][q][2][v]);
                }
            }
        }
    }
}

// Calibrate fp16 planes: out = fp16(v * g), the operand arithmetic above.
__global__ __launch_bounds__(256) void apply_cal_kernel(const _Float16 *re, const _Float16 *im,
                                                        const float *cal_re, const float *cal_im,
                                                        _Float16 *out_re, _Float16 *out_im, int n_chan,
                                                        int n_pkt, int n_ant, int64_t cal_ch_stride) {
    const int64_t n = (int64_t)n_chan * 4 * n_pkt * n_ant;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int ant = (int)(i % n_ant);
        const int64_t r = i / n_ant / n_pkt;                 // ch * 4 + tp
        const int p = (int)(r & 1);
        const int ch = (int)(r >> 2);
        const int64_t gi_idx = ch * cal_ch_stride + (int64_t)p * n_ant + ant;
        const float2 x = cal_mul((float)re[i], (float)im[i], cal_re[gi_idx], cal_im[gi_idx]);
        out_re[i] = (_Float16)x.x;
        out_im[i] = (_Float16)x.y;
    }
}

template <int NQ, int MAXP>
static int launch_cal(const CalCorrParams &P, size_t lds, hipStream_t s) {
    const int grid = std::min(P.n_items, 4096);
    (void)hipFuncSetAttribute((const void *)corr_cal_kernel<NQ, MAXP>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    corr_cal_kernel<NQ, MAXP><<<grid, CCAL_THREADS, lds, s>>>(P);
    return check_launch("dsx_corr_cal");
}

template <int NQ>
static int launch_cal_q(const CalCorrParams &P, size_t lds, hipStream_t s) {
    const int need = (P.n_pairs + CCAL_WAVES - 1) / CCAL_WAVES;
    if (need <= 1) return launch_cal<NQ, 1>(P, lds, s);
    if (need <= 2) return launch_cal<NQ, 2>(P, lds, s);
    if (need <= 3) return launch_cal<NQ, 3>(P, lds, s);
    if (need <= 4) return launch_cal<NQ, 4>(P, lds, s);
    if (need <= 6) return launch_cal<NQ, 6>(P, lds, s);
    return launch_cal<NQ, 9>(P, lds, s);
}

}  // namespace dsx

using namespace dsx;

extern "C" int dsx_corr_cal(const void *volt_re, const void *volt_im, const float *cal_re,
                            const float *cal_im, const float *time_keep, int n_acc, float *vis,
                            int n_pkt, int n_ant, int n_chan, int t_int, int csf, int fuse_pol,
                            int input_layout, int out_layout, float scale, dsx_stream_t stream) {
    DSX_REQUIRE(volt_re && vis, "dsx_corr_cal: null pointer");
    DSX_REQUIRE(input_layout == 0 || input_layout == 1, "dsx_corr_cal: input_layout must be 0 or 1");
    DSX_REQUIRE(input_layout == 0 ? volt_im == nullptr : volt_im != nullptr,
                "dsx_corr_cal: volt_im must be NULL for raw input and set for fp16 planes");
    DSX_REQUIRE(out_layout == 0 || out_layout == 1, "dsx_corr_cal: out_layout must be 0 or 1");
    DSX_REQUIRE((cal_re == nullptr) == (cal_im == nullptr), "dsx_corr_cal: cal_re and cal_im go together");
    DSX_REQUIRE(time_keep == nullptr || cal_re != nullptr, "dsx_corr_cal: time_keep needs gains");
    DSX_REQUIRE(n_pkt > 0 && n_ant >= 2 && n_chan > 0, "dsx_corr_cal: bad shape");
    DSX_REQUIRE(t_int >= 2 && t_int % 2 == 0, "dsx_corr_cal: t_int must be even and >= 2");
    DSX_REQUIRE((2 * n_pkt) % t_int == 0, "dsx_corr_cal: 2*n_pkt must be a multiple of t_int");
    DSX_REQUIRE(csf >= 1 && n_chan % csf == 0, "dsx_corr_cal: csf must divide n_chan");
    DSX_REQUIRE(time_keep == nullptr || (n_acc >= 1 && n_pkt % n_acc == 0),
                "dsx_corr_cal: n_acc must divide n_pkt");
    if (n_ant > 128) return set_error(DSX_ERANGE, "dsx_corr_cal: n_ant=%d > 128", n_ant);
    CalCorrParams P{};
    P.raw = (const uint8_t *)volt_re;
    P.re16 = (const _Float16 *)volt_re;
    P.im16 = (const _Float16 *)volt_im;
    P.cal_re = cal_re;
    P.cal_im = cal_im;
    P.tkeep = time_keep;
    P.vis = vis;
    P.scale = scale;
    P.n_pkt = n_pkt;
    P.n_ant = n_ant;
    P.n_chan = n_chan;
    P.t_int = t_int;
    P.csf = csf;
    P.fuse_pol = fuse_pol ? 1 : 0;
    P.in_layout = input_layout;
    P.out_layout = out_layout;
    P.n_acc = time_keep ? n_acc : 1;
    P.n_steps = 2 * n_pkt / t_int;
    P.n_cout = n_chan / csf;
    P.nbase = n_ant * (n_ant + 1) / 2;
    P.n_tiles = (n_ant + 15) / 16;
    P.n_pairs = P.n_tiles * (P.n_tiles + 1) / 2;
    P.n_items = P.n_steps * P.n_cout;
    // channel chunk: the largest divisor of csf whose rows fit the LDS budget
    size_t lds = 0;
    for (int d = csf; d >= 1; --d) {
        if (csf % d) continue;
        const int kc = ((t_int * d + 31) / 32) * 32;
        const int rs = 4 * kc + 8;
        const size_t need = (size_t)P.n_tiles * 16 * rs * sizeof(_Float16);
        if (need <= CCAL_LDS_BUDGET || d == 1) {
            P.csf_c = d;
            P.kc = kc;
            P.row_stride = rs;
            lds = need;
            break;
        }
    }
    if (lds > 160 * 1024)
        return set_error(DSX_ERANGE, "dsx_corr_cal: %zu B of LDS needed (t_int too large)", lds);
    P.n_chunks = csf / P.csf_c;
    P.vec4 = input_layout == 0 && P.csf_c % 4 == 0 && n_chan % 4 == 0 && ((uintptr_t)volt_re & 15) == 0;
    hipStream_t s = as_stream(stream);
    return P.fuse_pol ? launch_cal_q<1>(P, lds, s) : launch_cal_q<2>(P, lds, s);
}

extern "C" int dsx_apply_cal(const void *re, const void *im, const float *cal_re, const float *cal_im,
                             int cal_per_chan, void *out_re, void *out_im, int n_chan, int n_pkt,
                             int n_ant, dsx_stream_t stream) {
    DSX_REQUIRE(re && im && cal_re && cal_im && out_re && out_im, "dsx_apply_cal: null pointer");
    DSX_REQUIRE(n_chan > 0 && n_pkt > 0 && n_ant > 0, "dsx_apply_cal: bad shape");
    const int64_t n = (int64_t)n_chan * 4 * n_pkt * n_ant;
    const int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
    apply_cal_kernel<<<blocks, 256, 0, as_stream(stream)>>>(
        (const _Float16 *)re, (const _Float16 *)im, cal_re, cal_im, (_Float16 *)out_re,
        (_Float16 *)out_im, n_chan, n_pkt, n_ant, cal_per_chan ? 2 * (int64_t)n_ant : 0);
    return check_launch("dsx_apply_cal");
}
