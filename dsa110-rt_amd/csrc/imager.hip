// Fused nearest-cell gridder + C2R FFT imager, and the CSR gather gridder.
//
// Per (subband s, step t) image:
//   1. scatter the subband's taps into an int32 Hermitian half plane
//      H[r][w] (rows r in [r0, r0 + rc) of [0, N/2]) with LDS atomics --
//      integer sums, so the grid is exact and order independent;
//   2. row transforms Y[r][j] = sum_w H[r][w] e^{+2 pi i w j / N}
//      (radix-2 DIT in LDS, float32);
//   3. column C2R  out[i][j] = sum_r c_r Re(Y[r][j] e^{+2 pi i r i / N})
//      (c_0 = c_{N/2} = 1, else 2) as a float32 MFMA product
//      [cos | -sin](i, r) x [Re Y | Im Y](r, j) with v_mfma_f32_16x16x4f32;
//   4. epilogue: scale / (2 N^2), zero the outer npad ring and the DC pixel.
// The tap signs (-1)^(row+col) and the Hermitian fold make this equal to
// edge_mask * fftshift(Re(ifft2(G))) of the reference grid G
// (src/dsart/image/imager.py:112-324).
#include "dsx_internal.h"

namespace dsx {

constexpr int IMG_THREADS = 256;
constexpr int IMG_MAX_KB = 20;          // <= 40 rows per chunk (2 rows = 4 K values per MFMA)
constexpr int IMG_TAP_BATCH = 24;
constexpr int IMG_JT = 4;               // output tiles per column-transform pass
constexpr int IMG_KB_GROUP = 4;         // column-transform B operands loaded per group of MFMAs
static_assert(IMG_MAX_KB % IMG_KB_GROUP == 0, "K-block groups must tile IMG_MAX_KB");
// folded column transform: K-blocks of 4 uv rows (IMG_MAX_KB / 2 of them), rounded up to whole groups
constexpr int IMG_FOLD_KB = ((IMG_MAX_KB / 2 + IMG_KB_GROUP - 1) / IMG_KB_GROUP) * IMG_KB_GROUP;
constexpr int TAP_CONJ = 1 << 20;
constexpr int TAP_NEG = 1 << 21;
constexpr int TAP_CELL = (1 << 20) - 1;

// In-register 16-point transform X[k] = sum_n x[n] e^{+2 pi i n k / 16} (natural
// order in and out), as 4 x 4: four 4-point transforms over n1 (n = 4 n1 + n2),
// twiddles e^{2 pi i n2 k1 / 16}, four 4-point transforms over n2 (k = k1 + 4 k2).
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ void dft4(float2 &a, float2 &b, float2 &c, float2 &d) {
    // X0 = a+b+c+d, X1 = a + i b - c - i d, X2 = a - b + c - d, X3 = a - i b - c + i d
    const float2 s0 = make_float2(a.x + c.x, a.y + c.y), d0 = make_float2(a.x - c.x, a.y - c.y);
    const float2 s1 = make_float2(b.x + d.x, b.y + d.y), d1 = make_float2(b.x - d.x, b.y - d.y);
    a = make_float2(s0.x + s1.x, s0.y + s1.y);
    c = make_float2(s0.x - s1.x, s0.y - s1.y);
    b = make_float2(d0.x - d1.y, d0.y + d1.x);      // d0 + i d1
    d = make_float2(d0.x + d1.y, d0.y - d1.x);      // d0 - i d1
}

__device__ __forceinline__ void dft16(float2 (&x)[16]) {
    constexpr float C1 = 0.92387953251128674f, S1 = 0.38268343236508978f, R2 = 0.70710678118654752f;
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) dft4(x[n2], x[4 + n2], x[8 + n2], x[12 + n2]);   // -> A[n2][k1] at x[4 k1 + n2]
    // twiddles e^{2 pi i n2 k1 / 16} for n2, k1 in 1..3
    const float2 w1 = make_float2(C1, S1), w2 = make_float2(R2, R2), w3 = make_float2(S1, C1);
    const float2 w6 = make_float2(-R2, R2), w9 = make_float2(-C1, -S1);
    x[4 + 1] = cmul(x[4 + 1], w1);
    x[4 + 2] = cmul(x[4 + 2], w2);
    x[4 + 3] = cmul(x[4 + 3], w3);
    x[8 + 1] = cmul(x[8 + 1], w2);
    x[8 + 2] = make_float2(-x[8 + 2].y, x[8 + 2].x);    // i
    x[8 + 3] = cmul(x[8 + 3], w6);
    x[12 + 1] = cmul(x[12 + 1], w3);
    x[12 + 2] = cmul(x[12 + 2], w6);
    x[12 + 3] = cmul(x[12 + 3], w9);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) dft4(x[4 * k1], x[4 * k1 + 1], x[4 * k1 + 2], x[4 * k1 + 3]);
    // X[k1 + 4 k2] sits at x[4 k1 + k2]: transpose the 4 x 4 index
    float2 y[16];
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1)
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) y[k1 + 4 * k2] = x[4 * k1 + k2];
#pragma unroll
    for (int k = 0; k < 16; ++k) x[k] = y[k];
}

struct ImgParams {
    const int32_t *vis;
    const int32_t *tap_src;
    const int32_t *tap_dst;
    const int32_t *tap_off;   // n_sub + 1 CSR offsets (device)
    const float2 *grid;       // dense mode: (n_items, N, N) complex64
    void *img;
    int n_steps, n_sub, per_sub, nbase, N, log2N, n_rows, rc_max, npad;
    float scale;
    int dense;                // 0: taps from int32 visibilities (+ mask); 1: dense grid, no mask;
                              // 3: taps from float32 visibilities (calibrated streams, + mask);
                              // 2: shift-and-add of cint8 uv streams (cube mode, below)
    // cube mode: item (f, t), t in [t_lo, t_lo + n_steps), f < n_sub:
    //   G[x][y] = sum_g cal_g(streams[g][t - shifts[f][g]][re, im][x][y]) for in-range rows,
    //   img[t][f] = mask * fftshift(Re(ifft2(scale * G))) into a [T][n_sub][N][N] cube
    const int8_t *streams;    // [n_chg][t_stream][2][N][N]
    const int32_t *shifts;    // [n_sub][n_chg] (device)
    const float *cal_scale;   // [n_chg] or [n_chg][t_stream] (cal_2d), NULL: unit scale
    const float *cal_off_re;  // same shapes, NULL: zero
    const float *cal_off_im;
    const float *mask;        // [N][N] (or [n_sub][N][N], mask_stride N^2) multiplied in the epilogue
    float *acc;               // float32 partial images when the output type is not float32
    int n_chg, t_stream, t_lo, cal_2d;
    int64_t mask_stride;      // 0: one mask for every item; N^2: one per f
};

// cint8 cube streams: the calibrated value of one (chgroup, row, cell)
__device__ __forceinline__ float cal_at(const float *a, int cal_2d, int g, int t_stream, int tau, float dflt) {
    if (a == nullptr) return dflt;
    return cal_2d ? a[(int64_t)g * t_stream + tau] : a[g];
}

// R16: N = 256 row transforms as two in-register 16-point passes (rows padded
// to RS = 272 entries so the pass-1 transposed writes are conflict-free);
// otherwise radix-2 passes over rows of RS = N entries.
constexpr int R16_RS = 272;

// MODE (= P.dense) is a compile-time constant so each mode carries only its own registers.
template <typename TOut, bool R16, int MODE>
__global__ __launch_bounds__(IMG_THREADS, 2) void image_kernel(ImgParams P) {
    constexpr int DENSE = MODE == 3 ? 0 : MODE;     // taps (0), dense grid (1), cube streams (2)
    constexpr bool VISF = MODE == 3;                // float32 visibilities, float LDS atomics
    extern __shared__ __align__(16) float2 smem[];
    const int N = P.N;
    const int RS = R16 ? R16_RS : N;         // row stride of H
    float2 *tw = smem;                       // N entries: e^{+2 pi i m / N}
    float2 *tw16 = smem + N;                 // R16: [j1][w2] = e^{+2 pi i w2 j1 / 256}
    float2 *H = smem + N + (R16 ? 256 : 0);  // rc_max x RS
    int2 *Hi = reinterpret_cast<int2 *>(H);
    const int tid = threadIdx.x;
    for (int m = tid; m < N; m += IMG_THREADS) {
        double sv, cv;
        sincospi(2.0 * (double)m / (double)N, &sv, &cv);
        tw[m] = make_float2((float)cv, (float)sv);
    }
    if (R16) {
        __syncthreads();
        for (int m = tid; m < 256; m += IMG_THREADS) tw16[m] = tw[((m >> 4) * (m & 15)) & 255];
    }
    const int wave = tid >> 6, lane = tid & 63;
    const int lr = lane & 15, lg = lane >> 4;
    const int n_tiles_1d = N / 16;
    const int n_tiles = n_tiles_1d * n_tiles_1d;
    const float out_scale = P.scale / (2.0f * (float)N * (float)N);
    const int n_cout = P.n_sub * P.per_sub;
    const int n_items = P.n_sub * P.n_steps;
    const int n_chunks = (P.n_rows + P.rc_max - 1) / P.rc_max;
    TOut *img = reinterpret_cast<TOut *>(P.img);

    // cube mode: per-item stream rows (-1: outside [0, t_stream)) after the H rows
    int *tau_s = reinterpret_cast<int *>(smem + N + (R16 ? 256 : 0) + P.rc_max * RS);
    for (int item = blockIdx.x; item < n_items; item += gridDim.x) {
        const int s = item / P.n_steps;
        const int t = item - s * P.n_steps;
        const int32_t *vslab = DENSE ? nullptr
                                       : P.vis + 2 * ((int64_t)t * n_cout + (int64_t)s * P.per_sub) * P.nbase;
        const int tap_lo = DENSE ? 0 : P.tap_off[s], tap_hi = DENSE ? 0 : P.tap_off[s + 1];
        const float2 *gsrc = DENSE == 1 ? P.grid + (int64_t)item * N * N : nullptr;
        const int64_t o_item = DENSE == 2 ? ((int64_t)(P.t_lo + t) * P.n_sub + s) * N * N
                                            : ((int64_t)s * P.n_steps + t) * N * N;
        TOut *out = img + o_item;
        float *acc_out = (DENSE == 2 && P.acc != nullptr) ? P.acc + o_item : nullptr;   // cube mode only
        if constexpr (DENSE == 2) {
            __syncthreads();                 // the previous item's rows are no longer read
            for (int g = tid; g < P.n_chg; g += IMG_THREADS) {
                const int tau = P.t_lo + t - P.shifts[(int64_t)s * P.n_chg + g];
                tau_s[g] = (tau >= 0 && tau < P.t_stream) ? tau : -1;
            }
        }
        for (int ck = 0; ck < n_chunks; ++ck) {
            const int r0 = ck * P.rc_max;
            const int rc = min(P.rc_max, P.n_rows - r0);
            __syncthreads();
            if constexpr (DENSE != 0) {
                // Hermitian fold of (-1)^(a+b) G[a][b]: H[r][w] = (-1)^(r+w) (G[r][w] + conj G[-r][-w]),
                // written straight into float32 rows in bit-reversed order
                for (int i = tid; i < rc * N; i += IMG_THREADS) {
                    const int r = i >> P.log2N, w = i & (N - 1);
                    const int rg = r0 + r;
                    const int64_t ca = (int64_t)rg * N + w;
                    const int64_t cb = (int64_t)((N - rg) & (N - 1)) * N + ((N - w) & (N - 1));
                    float2 a, b;
                    if constexpr (DENSE == 2) {
                        // shift-and-add of the chgroups' cint8 rows at this cell and its mirror
                        const int64_t plane = (int64_t)N * N;
                        if (P.cal_scale == nullptr && P.cal_off_re == nullptr && P.cal_off_im == nullptr) {
                            int ar = 0, ai = 0, br = 0, bi = 0;      // exact integer sums
                            for (int g = 0; g < P.n_chg; ++g) {
                                const int tau = tau_s[g];
                                if (tau < 0) continue;
                                const int8_t *row = P.streams + ((int64_t)g * P.t_stream + tau) * 2 * plane;
                                ar += row[ca];
                                ai += row[plane + ca];
                                br += row[cb];
                                bi += row[plane + cb];
                            }
                            a = make_float2((float)ar, (float)ai);
                            b = make_float2((float)br, (float)bi);
                        } else {
                            a = make_float2(0.f, 0.f);
                            b = make_float2(0.f, 0.f);
                            for (int g = 0; g < P.n_chg; ++g) {
                                const int tau = tau_s[g];
                                if (tau < 0) continue;
                                const int8_t *row = P.streams + ((int64_t)g * P.t_stream + tau) * 2 * plane;
                                const float sc = cal_at(P.cal_scale, P.cal_2d, g, P.t_stream, tau, 1.f);
                                const float ore = cal_at(P.cal_off_re, P.cal_2d, g, P.t_stream, tau, 0.f);
                                const float oim = cal_at(P.cal_off_im, P.cal_2d, g, P.t_stream, tau, 0.f);
                                a.x += fmaf(sc, (float)row[ca], ore);
                                a.y += fmaf(sc, (float)row[plane + ca], oim);
                                b.x += fmaf(sc, (float)row[cb], ore);
                                b.y += fmaf(sc, (float)row[plane + cb], oim);
                            }
                        }
                    } else {
                        a = gsrc[ca];
                        b = gsrc[cb];
                    }
                    const float sg = ((rg + w) & 1) ? -1.f : 1.f;
                    const int wr = R16 ? w : (int)(__brev((unsigned)w) >> (32 - P.log2N));
                    H[r * RS + wr] = make_float2(sg * (a.x + b.x), sg * (a.y - b.y));
                }
                __syncthreads();
            } else {
            for (int i = tid; i < rc * RS; i += IMG_THREADS) Hi[i] = make_int2(0, 0);
            __syncthreads();
            // ---- 1. scatter taps (int32 LDS atomics).  A thread takes its taps in
            //      batches of IMG_TAP_BATCH with every load of a batch in flight
            //      together (tap tables, then the visibilities they name); loads use
            //      a valid index for every slot, slots past the end add nothing.
            for (int i0 = tap_lo + tid; i0 < tap_hi; i0 += IMG_THREADS * IMG_TAP_BATCH) {
                int dst[IMG_TAP_BATCH], src[IMG_TAP_BATCH];
#pragma unroll
                for (int k = 0; k < IMG_TAP_BATCH; ++k) {
                    const int i = i0 + k * IMG_THREADS;
                    const int ic = i < tap_hi ? i : tap_lo;
                    dst[k] = P.tap_dst[ic];
                    src[k] = P.tap_src[ic];
                }
                if constexpr (VISF) {
                    // calibrated visibilities: float32 LDS atomics (sum order not fixed)
                    float2 v[IMG_TAP_BATCH];
#pragma unroll
                    for (int k = 0; k < IMG_TAP_BATCH; ++k)
                        v[k] = *reinterpret_cast<const float2 *>(vslab + 2 * (int64_t)src[k]);
#pragma unroll
                    for (int k = 0; k < IMG_TAP_BATCH; ++k) {
                        const int d = dst[k];
                        const int cell = d & TAP_CELL;
                        const int r = (cell >> P.log2N) - r0;
                        if (i0 + k * IMG_THREADS < tap_hi && r >= 0 && r < rc) {
                            const int w = cell & (N - 1);
                            float2 x = v[k];
                            if (d & TAP_CONJ) x.y = -x.y;
                            if (d & TAP_NEG) { x.x = -x.x; x.y = -x.y; }
                            float *c = reinterpret_cast<float *>(&H[r * RS + w]);
                            atomicAdd(c, x.x);
                            atomicAdd(c + 1, x.y);
                        }
                    }
                } else {
                int2 v[IMG_TAP_BATCH];
#pragma unroll
                for (int k = 0; k < IMG_TAP_BATCH; ++k)
                    v[k] = *reinterpret_cast<const int2 *>(vslab + 2 * (int64_t)src[k]);
#pragma unroll
                for (int k = 0; k < IMG_TAP_BATCH; ++k) {
                    const int d = dst[k];
                    const int cell = d & TAP_CELL;
                    const int r = (cell >> P.log2N) - r0;
                    if (i0 + k * IMG_THREADS < tap_hi && r >= 0 && r < rc) {
                        const int w = cell & (N - 1);
                        int2 x = v[k];
                        if (d & TAP_CONJ) x.y = -x.y;
                        if (d & TAP_NEG) { x.x = -x.x; x.y = -x.y; }
                        int *c = reinterpret_cast<int *>(&Hi[r * RS + w]);
                        atomicAdd(c, x.x);
                        atomicAdd(c + 1, x.y);
                    }
                }
                }
            }
            __syncthreads();
            // ---- int32 -> float32 in place, with the bit-reversal permutation
            //      (R16: the first 16-point pass converts as it loads)
            if (!R16)
            for (int i = tid; i < rc * N; i += IMG_THREADS) {
                const int r = i >> P.log2N, j = i & (N - 1);
                const int jr = (int)(__brev((unsigned)j) >> (32 - P.log2N));
                if constexpr (VISF) {
                    if (j < jr) {
                        const float2 a = H[r * N + j], b = H[r * N + jr];
                        H[r * N + j] = b;
                        H[r * N + jr] = a;
                    }
                } else if (j < jr) {
                    const int2 a = Hi[r * N + j], b = Hi[r * N + jr];
                    H[r * N + j] = make_float2((float)b.x, (float)b.y);
                    H[r * N + jr] = make_float2((float)a.x, (float)a.y);
                } else if (j == jr) {
                    const int2 a = Hi[i];
                    H[i] = make_float2((float)a.x, (float)a.y);
                }
            }
            __syncthreads();
            }
            // ---- 2. row transforms Y[r][j] = sum_w H[r][w] e^{+2 pi i w j / N}
            if constexpr (R16) {
                // pass 1 (task = row, w2): Z[w2][j1] = e^{2 pi i w2 j1 / 256} sum_w1 H[16 w1 + w2]
                // e^{2 pi i w1 j1 / 16}, stored at 17 w2 + j1.  The 16 tasks of a row are
                // 16 lanes of one wave, so a row's reads finish before its writes.
                const bool h_int = !DENSE && !VISF;
                for (int task = tid; task < rc * 16; task += IMG_THREADS) {
                    const int r = task >> 4, w2 = task & 15;
                    float2 x[16];
#pragma unroll
                    for (int w1 = 0; w1 < 16; ++w1) {
                        const int idx = r * RS + 16 * w1 + w2;
                        if (h_int) {
                            const int2 v = Hi[idx];
                            x[w1] = make_float2((float)v.x, (float)v.y);
                        } else {
                            x[w1] = H[idx];
                        }
                    }
                    dft16(x);
#pragma unroll
                    for (int j1 = 1; j1 < 16; ++j1) x[j1] = cmul(x[j1], tw16[j1 * 16 + w2]);
#pragma unroll
                    for (int j1 = 0; j1 < 16; ++j1) H[r * RS + 17 * w2 + j1] = x[j1];
                }
                __syncthreads();
                // pass 2 (task = row, j1): Y[j1 + 16 j2] = sum_w2 Z[w2][j1] e^{2 pi i w2 j2 / 16}
                for (int task = tid; task < rc * 16; task += IMG_THREADS) {
                    const int r = task >> 4, j1 = task & 15;
                    float2 x[16];
#pragma unroll
                    for (int w2 = 0; w2 < 16; ++w2) x[w2] = H[r * RS + 17 * w2 + j1];
                    dft16(x);
#pragma unroll
                    for (int j2 = 0; j2 < 16; ++j2) H[r * RS + j1 + 16 * j2] = x[j2];
                }
                __syncthreads();
            } else {
            const int halfN_ = N >> 1;
            for (int m = 1; m < N; m <<= 1) {
                const int stride = N / (2 * m);
                for (int b = tid; b < rc * halfN_; b += IMG_THREADS) {
                    const int r = b / halfN_;
                    const int k = b - r * halfN_;
                    const int pos = k & (m - 1);
                    const int i0 = ((k - pos) << 1) + pos;
                    const float2 w = tw[pos * stride];
                    const float2 x0 = H[r * N + i0];
                    const float2 x1 = H[r * N + i0 + m];
                    const float2 y1 = make_float2(x1.x * w.x - x1.y * w.y, x1.x * w.y + x1.y * w.x);
                    H[r * N + i0] = make_float2(x0.x + y1.x, x0.y + y1.y);
                    H[r * N + i0 + m] = make_float2(x0.x - y1.x, x0.y - y1.y);
                }
                __syncthreads();
            }
            }
            const int halfN = N >> 1;
            // ---- 3. column C2R on MFMA, folded.  With theta = 2 pi r i / N,
            //         C[i][j] = sum_r c_r cos(theta) Re Y[r][j] and S[i][j] = -sum_r c_r sin(theta) Im Y[r][j]
            //         give out[i] = C + S and out[N - i] = C - S (cos is even and sin odd in i), so
            //         only the output rows i in [0, N/2) are transformed: two MFMA chains per tile
            //         (A = c_r cos or -c_r sin, B = Re Y or Im Y, K = 4 uv rows per MFMA), and
            //         one LDS read of Y[r][j] feeds both
            constexpr int JT = IMG_JT;
            const int nkb = (rc + 3) / 4;
            const int n_fold = (halfN + 15) / 16;            // i-tiles covering rows [0, N/2)
            // ---- 4. epilogue of one pixel: scale, outer ring and DC pixel zeroed, cube
            //         mask, chunk partial sums
            auto emit = [&](int i, int j, float x) {
                const bool inside = DENSE || (i >= P.npad && i < N - P.npad && j >= P.npad &&
                                                j < N - P.npad && !(i == halfN && j == halfN));
                const int64_t o = (int64_t)i * N + j;
                float val = inside ? x * out_scale : 0.f;
                if constexpr (DENSE == 2) {
                    if (P.mask != nullptr) val *= P.mask[(int64_t)s * P.mask_stride + o];
                }
                if (DENSE == 2 && acc_out != nullptr) {
                    // partial images in float32; the last pass writes the output type
                    if (ck > 0) val += acc_out[o];
                    if (ck + 1 < n_chunks) acc_out[o] = val;
                    else st_from_float(out, o, val);
                } else {
                    if (ck > 0) val += ld_as_float(out, o);
                    st_from_float(out, o, val);
                }
            };
            // row N/2 on the VALU: theta = pi r there, so the sine terms vanish and
            // out[N/2][j] = sum_r c_r (-1)^r Re Y[r][j]
            for (int j = tid; j < N; j += IMG_THREADS) {
                float x = 0.f;
                for (int r = 0; r < rc; ++r) {
                    const int rg = r0 + r;
                    const float cr = (rg == 0 || rg == halfN) ? 1.f : 2.f;
                    x += ((rg & 1) ? -cr : cr) * H[r * RS + j].x;
                }
                emit(halfN, j, x);
            }
            // work units (i-tile, group of JT j-tiles) spread over the waves, so the
            // N/32 folded i-tiles leave no wave idle
            const int n_jg = (n_tiles_1d + JT - 1) / JT;
            for (int unit = wave; unit < n_fold * n_jg; unit += IMG_THREADS / 64) {
                const int it = unit / n_jg;
                const int jt = (unit - it * n_jg) * JT;
                const int i0 = it * 16;
                float a_c[IMG_FOLD_KB], a_s[IMG_FOLD_KB];
                {
                    const int i = i0 + lr;
#pragma unroll
                    for (int kb = 0; kb < IMG_FOLD_KB; ++kb) {
                        const int r = kb * 4 + lg;
                        float c = 0.f, sn = 0.f;
                        if (kb < nkb && r < rc) {
                            // c_r (1 or 2) rides on the A operand: scaling by 2 is exact, so
                            // the products equal those with c_r on Re Y / Im Y
                            const int rg = r0 + r;
                            const float cr = (rg == 0 || rg == halfN) ? 1.f : 2.f;
                            const float2 w = tw[(rg * i) & (N - 1)];
                            c = cr * w.x;
                            sn = -(cr * w.y);
                        }
                        a_c[kb] = c;
                        a_s[kb] = sn;
                    }
                }
                // JT output tiles per unit: independent accumulator chains interleave,
                // so one tile's MFMA latency hides behind the others'
                {
                    v4f acc_c[JT], acc_s[JT];
#pragma unroll
                    for (int h = 0; h < JT; ++h) {
                        acc_c[h] = v4f{0.f, 0.f, 0.f, 0.f};
                        acc_s[h] = v4f{0.f, 0.f, 0.f, 0.f};
                    }
                    // B operands in groups of IMG_KB_GROUP: every LDS load of a group is
                    // issued (clamped row and tile, in bounds) before the group's MFMAs
#pragma unroll
                    for (int kb0 = 0; kb0 < IMG_FOLD_KB; kb0 += IMG_KB_GROUP) {
                        if (kb0 < nkb) {
                            float bc[JT][IMG_KB_GROUP], bs[JT][IMG_KB_GROUP];
#pragma unroll
                            for (int h = 0; h < JT; ++h) {
                                const int jc = (jt + h < n_tiles_1d ? jt + h : jt) * 16;
#pragma unroll
                                for (int u = 0; u < IMG_KB_GROUP; ++u) {
                                    // rows past the chunk read row 0 (finite) against a zero A
                                    const int r = (kb0 + u) * 4 + lg;
                                    const int rr = (kb0 + u < nkb && r < rc) ? r : 0;
                                    const float2 y = H[rr * RS + jc + lr];
                                    bc[h][u] = y.x;
                                    bs[h][u] = y.y;
                                }
                            }
#pragma unroll
                            for (int u = 0; u < IMG_KB_GROUP; ++u)
                                if (kb0 + u < nkb) {
#pragma unroll
                                    for (int h = 0; h < JT; ++h) {
                                        acc_c[h] = __builtin_amdgcn_mfma_f32_16x16x4f32(a_c[kb0 + u], bc[h][u], acc_c[h], 0, 0, 0);
                                        acc_s[h] = __builtin_amdgcn_mfma_f32_16x16x4f32(a_s[kb0 + u], bs[h][u], acc_s[h], 0, 0, 0);
                                    }
                                }
                        }
                    }
#pragma unroll
                    for (int h = 0; h < JT; ++h) {
                    if (jt + h >= n_tiles_1d) break;
                    const int j = (jt + h) * 16 + lr;
                    // row i from C + S, row N - i from C - S
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int i = i0 + 4 * lg + v;
                        if (i < halfN) emit(i, j, acc_c[h][v] + acc_s[h][v]);
                        if (i >= 1 && i < halfN) emit(N - i, j, acc_c[h][v] - acc_s[h][v]);
                    }
                    }
                }
            }
        }
    }
}

// out[f][c] = sum_{i in [ptr[c], ptr[c+1])} vis[f][src[i]]   (one thread per (f, c))
template <typename T>
__global__ __launch_bounds__(256) void grid_gather_kernel(const T *vis, const int32_t *ptr,
                                                          const int32_t *src, T *out,
                                                          int n_steps, int n_src, int n_cells) {
    const int64_t n = (int64_t)n_steps * n_cells;
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int f = (int)(g / n_cells);
        const int c = (int)(g - (int64_t)f * n_cells);
        const T *v = vis + 2 * (int64_t)f * n_src;
        T re = 0, im = 0;
        for (int i = ptr[c]; i < ptr[c + 1]; ++i) {
            const int sidx = src[i];
            re += v[2 * sidx];
            im += v[2 * sidx + 1];
        }
        out[2 * g] = re;
        out[2 * g + 1] = im;
    }
}

// uv rows per LDS pass: the twiddle tables plus rc rows of RS float2 in 96 KB
static int img_rows_per_chunk(int n_grid, bool r16) {
    const size_t budget = 96 * 1024;
    const size_t rs = r16 ? R16_RS : n_grid;
    const size_t fixed = (size_t)(n_grid + (r16 ? 256 : 0)) * sizeof(float2);
    const int rc = (int)((budget - fixed) / (rs * sizeof(float2)));
    return std::min(rc, 2 * IMG_MAX_KB);
}

static size_t img_lds_bytes(int n_grid, int rc_max, bool r16, int n_chg = 0) {
    const size_t rs = r16 ? R16_RS : n_grid;
    return ((size_t)n_grid + (r16 ? 256 : 0) + (size_t)rc_max * rs) * sizeof(float2) + (size_t)n_chg * sizeof(int);
}

template <typename TOut, bool R16, int MODE>
static int launch_image_m(const ImgParams &P, int grid, size_t lds, hipStream_t s) {
    (void)hipFuncSetAttribute((const void *)image_kernel<TOut, R16, MODE>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    image_kernel<TOut, R16, MODE><<<grid, IMG_THREADS, lds, s>>>(P);
    return check_launch("dsx_image_subbands");
}

template <typename TOut, bool R16>
static int launch_image(const ImgParams &P, int grid, size_t lds, hipStream_t s) {
    switch (P.dense) {
        case 0: return launch_image_m<TOut, R16, 0>(P, grid, lds, s);
        case 1: return launch_image_m<TOut, R16, 1>(P, grid, lds, s);
        case 2: return launch_image_m<TOut, R16, 2>(P, grid, lds, s);
        default: return launch_image_m<TOut, R16, 3>(P, grid, lds, s);
    }
}

}  // namespace dsx

using namespace dsx;

static int image_subbands(const void *vis, int vis_f32, const int32_t *tap_src,
                          const int32_t *tap_dst, const int32_t *tap_off, void *img,
                          int n_steps, int n_sub, int per_sub, int nbase, int n_grid,
                          int n_rows, int npad, float scale, int img_dtype,
                          dsx_stream_t stream) {
    DSX_REQUIRE(vis && tap_src && tap_dst && tap_off && img, "dsx_image_subbands: null pointer");
    DSX_REQUIRE(n_steps > 0 && n_sub > 0 && per_sub > 0 && nbase > 0, "dsx_image_subbands: bad shape");
    DSX_REQUIRE(n_grid >= 16 && n_grid <= 1024 && (n_grid & (n_grid - 1)) == 0,
                "dsx_image_subbands: n_grid must be a power of two in [16, 1024]");
    DSX_REQUIRE(n_rows >= 1 && n_rows <= n_grid / 2 + 1, "dsx_image_subbands: bad n_rows %d", n_rows);
    DSX_REQUIRE(img_dtype == DSX_F16 || img_dtype == DSX_F32, "dsx_image_subbands: img_dtype");
    int log2N = 0;
    while ((1 << log2N) < n_grid) ++log2N;
    const bool r16 = n_grid == 256;
    int rc_max = img_rows_per_chunk(n_grid, r16);
    rc_max = std::min(rc_max, n_rows);
    if (rc_max < 1) return set_error(DSX_ERANGE, "dsx_image_subbands: n_grid too large");
    const int n_chunks = (n_rows + rc_max - 1) / rc_max;
    if (n_chunks > 1 && img_dtype != DSX_F32)
        return set_error(DSX_ENOTSUP,
                         "dsx_image_subbands: %d uv rows need %d passes; multi-pass imaging "
                         "accumulates in place and needs img_dtype DSX_F32", n_rows, n_chunks);
    hipStream_t s = as_stream(stream);
    ImgParams P{(const int32_t *)vis, tap_src, tap_dst, tap_off, nullptr, img, n_steps, n_sub, per_sub,
                nbase, n_grid, log2N, n_rows, rc_max, npad, scale, vis_f32 ? 3 : 0};
    const size_t lds = img_lds_bytes(n_grid, rc_max, r16);
    const int grid = std::min(n_sub * n_steps, 4096);
    if (img_dtype == DSX_F16)
        return r16 ? launch_image<_Float16, true>(P, grid, lds, s) : launch_image<_Float16, false>(P, grid, lds, s);
    return r16 ? launch_image<float, true>(P, grid, lds, s) : launch_image<float, false>(P, grid, lds, s);
}

extern "C" int dsx_image_subbands(const int32_t *vis, const int32_t *tap_src,
                                  const int32_t *tap_dst, const int32_t *tap_off, void *img,
                                  int n_steps, int n_sub, int per_sub, int nbase, int n_grid,
                                  int n_rows, int npad, float scale, int img_dtype,
                                  dsx_stream_t stream) {
    return image_subbands(vis, 0, tap_src, tap_dst, tap_off, img, n_steps, n_sub, per_sub, nbase, n_grid,
                          n_rows, npad, scale, img_dtype, stream);
}

extern "C" int dsx_image_subbands_f32(const float *vis, const int32_t *tap_src,
                                      const int32_t *tap_dst, const int32_t *tap_off, void *img,
                                      int n_steps, int n_sub, int per_sub, int nbase, int n_grid,
                                      int n_rows, int npad, float scale, int img_dtype,
                                      dsx_stream_t stream) {
    return image_subbands(vis, 1, tap_src, tap_dst, tap_off, img, n_steps, n_sub, per_sub, nbase, n_grid,
                          n_rows, npad, scale, img_dtype, stream);
}

extern "C" int dsx_grid_gather(const void *vis, const int32_t *cell_ptr, const int32_t *cell_src,
                               void *out, int n_steps, int n_src, int n_cells, int vis_dtype,
                               dsx_stream_t stream) {
    DSX_REQUIRE(vis && cell_ptr && cell_src && out, "dsx_grid_gather: null pointer");
    DSX_REQUIRE(n_steps > 0 && n_src > 0 && n_cells >= 0, "dsx_grid_gather: bad shape");
    if (n_cells == 0) return DSX_OK;
    const int64_t n = (int64_t)n_steps * n_cells;
    const int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
    hipStream_t s = as_stream(stream);
    if (vis_dtype == DSX_I32)
        grid_gather_kernel<int32_t><<<blocks, 256, 0, s>>>((const int32_t *)vis, cell_ptr, cell_src,
                                                           (int32_t *)out, n_steps, n_src, n_cells);
    else if (vis_dtype == DSX_F32)
        grid_gather_kernel<float><<<blocks, 256, 0, s>>>((const float *)vis, cell_ptr, cell_src,
                                                         (float *)out, n_steps, n_src, n_cells);
    else
        return set_error(DSX_EINVAL, "dsx_grid_gather: vis_dtype must be DSX_I32 or DSX_F32");
    return check_launch("dsx_grid_gather");
}

extern "C" int dsx_dirty_image(const void *grid, void *img, int n_images, int n_grid, float prescale,
                               dsx_stream_t stream) {
    DSX_REQUIRE(grid && img, "dsx_dirty_image: null pointer");
    DSX_REQUIRE(n_images > 0, "dsx_dirty_image: n_images must be > 0");
    DSX_REQUIRE(n_grid >= 16 && n_grid <= 1024 && (n_grid & (n_grid - 1)) == 0,
                "dsx_dirty_image: n_grid must be a power of two in [16, 1024]");
    int log2N = 0;
    while ((1 << log2N) < n_grid) ++log2N;
    const int n_rows = n_grid / 2 + 1;
    const bool r16 = n_grid == 256;
    int rc_max = std::min(img_rows_per_chunk(n_grid, r16), n_rows);
    if (rc_max < 1) return set_error(DSX_ERANGE, "dsx_dirty_image: n_grid too large");
    ImgParams P{nullptr, nullptr, nullptr, nullptr, (const float2 *)grid, img, n_images, 1, 1, 1,
                n_grid, log2N, n_rows, rc_max, 0, prescale, 1};
    const size_t lds = img_lds_bytes(n_grid, rc_max, r16);
    const int blocks = std::min(n_images, 4096);
    const int rc = r16 ? launch_image<float, true>(P, blocks, lds, as_stream(stream))
                       : launch_image<float, false>(P, blocks, lds, as_stream(stream));
    if (rc != DSX_OK) return rc;
    return DSX_OK;
}

extern "C" int dsx_image_cube(const int8_t *streams, const int32_t *shifts, int n_chg, int t_stream,
                              int n_fdm, int t_det, int t_lo, int n_grid, float prescale,
                              const float *cal_scale, const float *cal_off_re, const float *cal_off_im,
                              int cal_2d, const float *mask, int mask_per_fdm, void *cube, int cube_dtype,
                              float *acc, dsx_stream_t stream) {
    DSX_REQUIRE(streams && shifts && cube, "dsx_image_cube: null pointer");
    DSX_REQUIRE(n_chg > 0 && n_chg <= 4096 && t_stream > 0 && n_fdm > 0 && t_det > 0,
                "dsx_image_cube: bad shape");
    DSX_REQUIRE(t_lo >= 0 && t_lo < t_det, "dsx_image_cube: t_lo=%d not in [0, t_det=%d)", t_lo, t_det);
    DSX_REQUIRE(n_grid >= 16 && n_grid <= 1024 && (n_grid & (n_grid - 1)) == 0,
                "dsx_image_cube: n_grid must be a power of two in [16, 1024]");
    DSX_REQUIRE(cube_dtype == DSX_F16 || cube_dtype == DSX_F32, "dsx_image_cube: cube_dtype");
    DSX_REQUIRE((int64_t)n_fdm * (t_det - t_lo) <= INT32_MAX, "dsx_image_cube: too many images");
    int log2N = 0;
    while ((1 << log2N) < n_grid) ++log2N;
    const int n_rows = n_grid / 2 + 1;
    const bool r16 = n_grid == 256;
    int rc_max = std::min(img_rows_per_chunk(n_grid, r16), n_rows);
    if (rc_max < 1) return set_error(DSX_ERANGE, "dsx_image_cube: n_grid too large");
    const int n_chunks = (n_rows + rc_max - 1) / rc_max;
    if (cube_dtype != DSX_F32 && n_chunks > 1 && acc == nullptr)
        return set_error(DSX_EINVAL, "dsx_image_cube: a float16 cube needs a float32 acc buffer "
                                     "(%d uv row passes)", n_chunks);
    ImgParams P{};
    P.img = cube;
    P.n_steps = t_det - t_lo;
    P.n_sub = n_fdm;
    P.per_sub = 1;
    P.nbase = 1;
    P.N = n_grid;
    P.log2N = log2N;
    P.n_rows = n_rows;
    P.rc_max = rc_max;
    P.scale = prescale;
    P.dense = 2;
    P.streams = streams;
    P.shifts = shifts;
    P.cal_scale = cal_scale;
    P.cal_off_re = cal_off_re;
    P.cal_off_im = cal_off_im;
    P.mask = mask;
    P.mask_stride = mask_per_fdm ? (int64_t)n_grid * n_grid : 0;
    P.acc = cube_dtype == DSX_F32 ? nullptr : acc;
    P.n_chg = n_chg;
    P.t_stream = t_stream;
    P.t_lo = t_lo;
    P.cal_2d = cal_2d != 0;
    const size_t lds = img_lds_bytes(n_grid, rc_max, r16, n_chg);
    if (lds > 160 * 1024) return set_error(DSX_ERANGE, "dsx_image_cube: %zu B of LDS", lds);
    const int blocks = (int)std::min<int64_t>((int64_t)n_fdm * (t_det - t_lo), 4096);
    hipStream_t s = as_stream(stream);
    if (cube_dtype == DSX_F16)
        return r16 ? launch_image<_Float16, true>(P, blocks, lds, s) : launch_image<_Float16, false>(P, blocks, lds, s);
    return r16 ? launch_image<float, true>(P, blocks, lds, s) : launch_image<float, false>(P, blocks, lds, s);
}
