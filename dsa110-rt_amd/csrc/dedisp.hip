// Subband dedispersion fused with the boxcar bank, float64 running statistics
// and chunk-maximum candidate extraction.  The DM x time x pixel cube is never
// written: each work-group owns one pixel, each lane one DM trial.
//
// Work-group = 1 pixel x up to 256 DMs (lane -> DM, blockIdx.y -> DM group).
// Per time tile of TT centres the work-group stages, for every subband s,
// the window of steps any of its DMs reads into LDS (float32):
//     tau in [c0 - W/2 - shmax_s, c0 + TT + W/2 - shmin_s)
// Each lane then sweeps time in 64-centre chunks.  It keeps the exclusive
// prefix sum P of its dedispersed series in a register window
//     Pw[x] = P[c - W/2 + x],  x in [0, 64 + 2 (W/2)]
// computing every D(t) = sum_s win_s[t - shift[s][d]] exactly once, then
//     B_w(c + i) = Pw[i + W/2 - w/2 + w] - Pw[i + W/2 - w/2]      (all widths)
// The chunk maximum of each width's boxcar is turned into S/N with the
// float64 state read at kernel start; a chunk whose best S/N exceeds the
// threshold appends one candidate.  Block sums of complete centres are merged
// into the float64 state at the end (Chan), so detection in a block uses only
// earlier blocks' statistics.
#include "dsx_internal.h"

namespace dsx {

struct Cand {
    int32_t t;
    int32_t p;
    int16_t d;
    int16_t k;
    float snr;
};
static_assert(sizeof(Cand) == 16, "candidate record must be 16 bytes");

struct DedispParams {
    const void *ring;
    const int32_t *shift;       // [s][d]
    const int32_t *sub_info;    // [s][2] = (lds offset, shmax); shmin is 0 for DM 0
    const int32_t *valid_from;  // [d]
    const double *n_prev;       // [d]
    const double *blk_count;    // [d]
    double *mean;               // [p][d]
    double *m2;                 // [p][d]
    Cand *cand;
    int32_t *n_cand;
    int64_t T0, c_lo;
    int n_sub, n_pix, hist, row_len, n_dm, n_centres, tt, win_total, cand_cap;
    float threshold;
    double min_count;
};

constexpr int DD_CHUNK = 64;

constexpr int DD_THREADS = 256;   // DMs per work-group (grid.y covers the rest)

// D(t) of one DM from the staged windows; t_rel is relative to the tile's
// first centre and the subbands are summed in the register path's order.
__device__ __forceinline__ float window_d(const float *win, const int32_t *sub_info,
                                          const int32_t *shift, int n_sub, int n_dm, int d,
                                          int t_rel, int h2) {
    float acc = 0.f;
    for (int s = 0; s < n_sub; ++s)
        acc += win[sub_info[2 * s] + sub_info[2 * s + 1] - shift[s * n_dm + d] + t_rel + h2];
    return acc;
}

// Earliest centre i in [0, 64) of the chunk at tile offset cc whose width-2^k
// boxcar is largest (running window sum).  Runs only for chunks that produce
// a candidate; the reported S/N comes from the register maximum.
__device__ __forceinline__ int rescan_chunk(const float *win, const int32_t *sub_info,
                                            const int32_t *shift, int n_sub, int n_dm, int d, int k,
                                            int cc, int h2) {
    const int w = 1 << k;
    const int lo = cc - (w >> 1);                  // first sample of centre cc's box
    float run = 0.f;
    for (int j = 0; j < w; ++j) run += window_d(win, sub_info, shift, n_sub, n_dm, d, lo + j, h2);
    float best = run;
    int bi = 0;
    for (int i = 1; i < DD_CHUNK; ++i) {
        run += window_d(win, sub_info, shift, n_sub, n_dm, d, lo + i + w - 1, h2) -
               window_d(win, sub_info, shift, n_sub, n_dm, d, lo + i - 1, h2);
        if (run > best) {
            best = run;
            bi = i;
        }
    }
    return bi;
}

template <typename T, int NW>
__global__ __launch_bounds__(DD_THREADS) void dedisp_kernel(DedispParams P) {
    constexpr int WMAX = 1 << (NW - 1);
    constexpr int H2 = WMAX / 2;
    constexpr int WW = 2 * H2;
    constexpr int NP = DD_CHUNK + WW + 1;            // register prefix window
    extern __shared__ float win[];
    const T *ring = reinterpret_cast<const T *>(P.ring);
    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    const int d = blockIdx.y * DD_THREADS + tid;      // one lane per DM
    const bool active = d < P.n_dm;

    for (int p = blockIdx.x; p < P.n_pix; p += gridDim.x) {
        // ---- per-lane state from earlier blocks (float64)
        float wmu[NW], inv[NW];
        bool score = false;
        int vf = 0;
        double mean_a = 0.0, m2_a = 0.0, n_a = 0.0;
        if (active) {
            vf = P.valid_from[d];
            n_a = P.n_prev[d];
            mean_a = P.mean[(int64_t)p * P.n_dm + d];
            m2_a = P.m2[(int64_t)p * P.n_dm + d];
            const double sig = n_a > 0.0 ? sqrt(m2_a / n_a) : 0.0;
            score = (n_a >= P.min_count) && (n_a > 0.0) && (sig > 0.0);
#pragma unroll
            for (int k = 0; k < NW; ++k) {
                const double w = (double)(1 << k);
                wmu[k] = (float)(w * mean_a);
                inv[k] = score ? (float)(1.0 / (sig * sqrt(w))) : 0.f;
            }
        }
        double s1 = 0.0, s2 = 0.0;
        float Pw[NP];
#pragma unroll
        for (int x = 0; x < NP; ++x) Pw[x] = 0.f;

        for (int c0 = 0; c0 < P.n_centres; c0 += P.tt) {
            const int64_t cabs0 = P.c_lo + c0;
            const int tt = min(P.tt, P.n_centres - c0);
            // ---- stage the windows of every subband (float32)
            __syncthreads();
            for (int s = 0; s < P.n_sub; ++s) {
                const int off = P.sub_info[2 * s];
                const int shmax = P.sub_info[2 * s + 1];
                const int len = tt + WW + shmax;
                const int64_t tau_lo = cabs0 - H2 - shmax;
                const T *row = ring + ((int64_t)s * P.n_pix + p) * P.row_len;
                for (int e = tid; e < len; e += nthr) {
                    const int64_t pos = P.hist + (tau_lo + e) - P.T0;
                    float v = 0.f;
                    if (pos >= 0 && pos < P.row_len) v = ld_as_float(row, pos);
                    win[off + e] = v;
                }
            }
            __syncthreads();
            if (!active) continue;

            // index of step t for subband s in this tile: off_s + (t - tau_lo_s) - shift
            //   = off_s + t - cabs0 + H2 + shmax_s - shift[s][d]
            if (c0 == 0) {
                // prologue: Pw[0] = 0, Pw[x + 1] = Pw[x] + D(c_lo - H2 + x), x < WW
#pragma unroll
                for (int x = 0; x < NP; ++x) Pw[x] = 0.f;
                for (int s = 0; s < P.n_sub; ++s) {
                    const int base = P.sub_info[2 * s] + P.sub_info[2 * s + 1] - P.shift[s * P.n_dm + d];
                    // t = cabs0 - H2 + x  ->  index = base + x
#pragma unroll
                    for (int x = 0; x < WW; ++x) Pw[x + 1] += win[base + x];
                }
#pragma unroll
                for (int x = 0; x < WW; ++x) Pw[x + 1] += Pw[x];
            }

            for (int cc = 0; cc < tt; cc += DD_CHUNK) {
                const int64_t c = cabs0 + cc;
                // ---- D(c + H2 + i), i < 64, accumulated over subbands into the
                //      window slots, then turned into prefix sums in place
#pragma unroll
                for (int i = 0; i < DD_CHUNK; ++i) Pw[WW + 1 + i] = 0.f;
                for (int s = 0; s < P.n_sub; ++s) {
                    // t = c + H2 + i -> index = off + (c - cabs0) + 2 H2 + shmax - shift + i
                    const int base = P.sub_info[2 * s] + P.sub_info[2 * s + 1] - P.shift[s * P.n_dm + d] +
                                     cc + WW;
#pragma unroll
                    for (int i = 0; i < DD_CHUNK; ++i) Pw[WW + 1 + i] += win[base + i];
                }
#pragma unroll
                for (int i = 0; i < DD_CHUNK; ++i) Pw[WW + 1 + i] += Pw[WW + i];

                // ---- statistics of complete centres (D(c + i) = B_1)
                if (c >= vf) {
                    float a1 = 0.f, a2 = 0.f;
#pragma unroll
                    for (int i = 0; i < DD_CHUNK; ++i) {
                        const float v = Pw[i + H2 + 1] - Pw[i + H2];
                        a1 += v;
                        a2 = fmaf(v, v, a2);
                    }
                    s1 += (double)a1;
                    s2 += (double)a2;
                } else if (c + DD_CHUNK > vf) {
                    const int i_first = (int)(vf - c);    // first complete centre of the chunk
                    float a1 = 0.f, a2 = 0.f;
#pragma unroll
                    for (int i = 0; i < DD_CHUNK; ++i) {
                        const float v = Pw[i + H2 + 1] - Pw[i + H2];
                        const bool ok = i >= i_first;
                        a1 += ok ? v : 0.f;
                        a2 = fmaf(ok ? v : 0.f, v, a2);
                    }
                    s1 += (double)a1;
                    s2 += (double)a2;
                }

                // ---- boxcar chunk maxima and candidate
                if (score && c - H2 >= vf) {
                    float mb[NW];
#pragma unroll
                    for (int k = 0; k < NW; ++k) {
                        const int w = 1 << k;
                        const int lo = H2 - (w >> 1);
                        float m = -INFINITY;
#pragma unroll
                        for (int i = 0; i < DD_CHUNK; ++i) m = fmaxf(m, Pw[i + lo + w] - Pw[i + lo]);
                        mb[k] = m;
                    }
                    float best = -INFINITY;
                    int bk = 0;
#pragma unroll
                    for (int k = 0; k < NW; ++k) {
                        const float snr = (mb[k] - wmu[k]) * inv[k];
                        if (snr > best) {
                            best = snr;
                            bk = k;
                        }
                    }
                    if (best > P.threshold) {
                        // rare path: locate the earliest centre of the winning
                        // width by re-summing D from the LDS window
                        const int bi = rescan_chunk(win, P.sub_info, P.shift, P.n_sub, P.n_dm, d, bk,
                                                    cc, H2);
                        const int slot = atomicAdd(P.n_cand, 1);
                        if (slot < P.cand_cap) {
                            Cand cd;
                            cd.t = (int32_t)(c + bi);
                            cd.p = p;
                            cd.d = (int16_t)d;
                            cd.k = (int16_t)bk;
                            cd.snr = best;
                            P.cand[slot] = cd;
                        }
                    }
                }
                // ---- slide the prefix window by one chunk
#pragma unroll
                for (int x = 0; x <= WW; ++x) Pw[x] = Pw[x + DD_CHUNK];
            }
        }
        // ---- merge this block's statistics into the running state
        if (active) {
            const double nb = P.blk_count[d];
            if (nb > 0.0) {
                const double mean_b = s1 / nb;
                const double m2_b = s2 - s1 * s1 / nb;
                const double n = n_a + nb;
                const double delta = mean_b - mean_a;
                P.mean[(int64_t)p * P.n_dm + d] = mean_a + delta * (nb / n);
                P.m2[(int64_t)p * P.n_dm + d] = m2_a + m2_b + delta * delta * (n_a * nb / n);
            }
        }
    }
}

template <typename T, int NW>
static int launch_dedisp(const DedispParams &P, int threads, size_t lds, dim3 grid, hipStream_t s) {
    (void)hipFuncSetAttribute((const void *)dedisp_kernel<T, NW>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    dedisp_kernel<T, NW><<<grid, threads, lds, s>>>(P);
    return check_launch("dsx_dedisp_detect");
}

template <typename T>
static int dispatch_nw(int nw, const DedispParams &P, int threads, size_t lds, dim3 grid, hipStream_t s) {
    switch (nw) {
        case 1: return launch_dedisp<T, 1>(P, threads, lds, grid, s);
        case 2: return launch_dedisp<T, 2>(P, threads, lds, grid, s);
        case 3: return launch_dedisp<T, 3>(P, threads, lds, grid, s);
        case 4: return launch_dedisp<T, 4>(P, threads, lds, grid, s);
        case 5: return launch_dedisp<T, 5>(P, threads, lds, grid, s);
        case 6: return launch_dedisp<T, 6>(P, threads, lds, grid, s);
        case 7: return launch_dedisp<T, 7>(P, threads, lds, grid, s);
        default: return set_error(DSX_ENOTSUP, "dsx_dedisp_detect: n_widths=%d not in [1, 7]", nw);
    }
}

}  // namespace dsx

using namespace dsx;

extern "C" int dsx_dedisp_tile(int n_sub, const int32_t *host_shmax, int n_widths, int n_centres,
                               int lds_budget) {
    // Largest tile (multiple of 64, <= n_centres) whose windows fit the LDS budget.
    int64_t span = 0;
    for (int s = 0; s < n_sub; ++s) span += host_shmax[s];
    const int ww = n_widths > 1 ? (1 << (n_widths - 1)) : 0;
    int best = 0;
    for (int tt = 64; tt <= n_centres && tt <= 2048; tt += 64) {
        const int64_t floats = (int64_t)n_sub * (tt + ww) + span;
        if (floats * 4 <= lds_budget && n_centres % tt == 0) best = tt;
    }
    return best;
}

extern "C" int dsx_dedisp_detect(const void *ring, int ring_dtype, int n_sub, int n_pix, int hist,
                                 int ring_len, int64_t T0, const int32_t *shift, int n_dm,
                                 const int32_t *sub_info, const int32_t *valid_from,
                                 const double *n_prev, const double *blk_count, double *mean,
                                 double *m2, int64_t c_lo, int n_centres, int n_widths, int tile,
                                 int win_total, float threshold, double min_count, void *cand,
                                 int cand_cap, int32_t *n_cand, dsx_stream_t stream) {
    DSX_REQUIRE(ring && shift && sub_info && valid_from && n_prev && blk_count && mean && m2 &&
                    cand && n_cand,
                "dsx_dedisp_detect: null pointer");
    DSX_REQUIRE(ring_dtype == DSX_F16 || ring_dtype == DSX_F32, "dsx_dedisp_detect: ring_dtype");
    DSX_REQUIRE(n_sub > 0 && n_pix > 0 && n_dm > 0 && hist >= 0 && ring_len > hist,
                "dsx_dedisp_detect: bad shape");
    DSX_REQUIRE(n_dm <= 32767, "dsx_dedisp_detect: n_dm=%d > 32767", n_dm);
    DSX_REQUIRE(n_centres > 0 && n_centres % DD_CHUNK == 0,
                "dsx_dedisp_detect: n_centres must be a positive multiple of 64");
    DSX_REQUIRE(tile > 0 && tile % DD_CHUNK == 0 && n_centres % tile == 0,
                "dsx_dedisp_detect: tile must be a multiple of 64 dividing n_centres");
    DSX_REQUIRE(cand_cap >= 0, "dsx_dedisp_detect: cand_cap");
    const size_t lds = (size_t)win_total * sizeof(float);
    if (lds > 160 * 1024)
        return set_error(DSX_ERANGE, "dsx_dedisp_detect: %zu B of LDS for the time windows", lds);
    DedispParams P{};
    P.ring = ring;
    P.shift = shift;
    P.sub_info = sub_info;
    P.valid_from = valid_from;
    P.n_prev = n_prev;
    P.blk_count = blk_count;
    P.mean = mean;
    P.m2 = m2;
    P.cand = reinterpret_cast<Cand *>(cand);
    P.n_cand = n_cand;
    P.T0 = T0;
    P.c_lo = c_lo;
    P.n_sub = n_sub;
    P.n_pix = n_pix;
    P.hist = hist;
    P.row_len = ring_len;
    P.n_dm = n_dm;
    P.n_centres = n_centres;
    P.tt = tile;
    P.win_total = win_total;
    P.cand_cap = cand_cap;
    P.threshold = threshold;
    P.min_count = min_count;
    const int threads = std::min(DD_THREADS, ((n_dm + 63) / 64) * 64);
    const dim3 grid((unsigned)std::min(n_pix, 1 << 20), (unsigned)((n_dm + DD_THREADS - 1) / DD_THREADS));
    hipStream_t s = as_stream(stream);
    if (ring_dtype == DSX_F16) return dispatch_nw<_Float16>(n_widths, P, threads, lds, grid, s);
    return dispatch_nw<float>(n_widths, P, threads, lds, grid, s);
}
