// Subband dedispersion fused with the boxcar bank, float64 running statistics
// and chunk-maximum candidate extraction.  The DM x time x pixel cube is never
// written: each work-group owns one pixel, each lane one DM trial.
//
// Subband-pair partial sums.  For adjacent subbands a = 2g, b = 2g + 1 a DM
// reads a[t - sh_a] + b[t - sh_b] = Q_delta[t - sh_a] with
//     Q_delta[tau] = a[tau] + b[tau - delta],   delta = sh_b - sh_a,
// and a pair has far fewer distinct deltas than the plan has DMs (12 for the
// 256-trial 0-1500 plan over 16 subbands).  Per time tile the work-group
// stages every subband's window of steps in LDS (ring dtype), builds one
// float32 row per (pair, delta) trimmed to the span of steps its DMs read,
// and each lane then sums n_sub / 2 rows per dedispersed sample instead of
// n_sub subbands.  Plans whose pair rows do not fit the LDS budget (many
// subbands with few DMs each) run unpaired: the lanes sum the subband
// windows directly.
//
// Centring.  Each lane subtracts a bias b from every dedispersed sample,
// D' = D - b, where b is its float64 running mean rounded to 11 significant
// bits (before a DM's first block, dsx_dedisp_prime sets that mean to the
// block's own mean).  Prefix sums, boxcars
// and the block's statistics are accumulated on D', so their float32
// magnitudes stay at the noise level even for pixels holding a steady
// source; the float64 merge and the S/N add b back exactly.
//
// Each lane sweeps time in 64-centre chunks and keeps the exclusive prefix
// sum P of its centred series in a register window
//     Pw[x] = P[c - W/2 + x],  x in [0, 64 + 2 (W/2)]
// so every D'(t) is formed once; width-1 maxima and the block's statistics
// come from the fresh D' values, the other widths from prefix differences
//     B_w(c + i) = Pw[i + W/2 - w/2 + w] - Pw[i + W/2 - w/2].
// The chunk maximum of each width becomes an S/N with the float64 state read
// at kernel start; a chunk whose best S/N exceeds the threshold appends one
// candidate, located by recomputing the winning width's boxcars with the
// same arithmetic (bit-identical values).  Block sums of complete samples are
// merged into the float64 state at the end (Chan), so detection in a block
// uses only earlier blocks' statistics.
#include <algorithm>
#include <climits>
#include <type_traits>
#include <vector>

#include "dsx_internal.h"
#include "dedisp_plan.h"

namespace dsx {

struct Cand {
    int32_t t;      // centre step relative to c_lo (the caller adds the int64 c_lo)
    int32_t p;
    int16_t d;
    int16_t k;
    float snr;
};
static_assert(sizeof(Cand) == 16, "candidate record must be 16 bytes");

struct DedispParams {
    const void *ring;
    const int32_t *sub_info;    // [s][3] = (R offset, shmax_s, window length L_s)
    const int32_t *rows;        // [r][6] = (Q offset, length, R index of a at e = 0,
                                //           R index of b at e = 0, b's R bounds [lo, hi); empty
                                //           bounds for a pair without b)
    const int32_t *tbase;       // [g][d]: LDS element index of D(t)'s term g at t = tile start
    const int32_t *valid_from;  // [d]
    const double *n_prev;       // [d]
    const double *blk_count;    // [d]
    double *mean;               // [p][d] state of earlier blocks, merged in place when `merge`
    double *m2;                 // [p][d]
    Cand *cand;
    int32_t *n_cand;            // records appended (kept or not)
    int64_t T0, c_lo;
    int n_sub, n_pix, hist, row_len, n_dm, n_rows, n_terms, n_centres, tt, r_total, q_total,
        cand_cap;
    float threshold;
    int merge;                  // 0: detect only, the state is left unchanged
    double min_count;
};

// x rounded to 11 significant bits (fp16 precision, float32 range).  Centring
// on such a value adds no bits finer than the samples' own at the mean's
// magnitude, so sums of fp16-valued samples that were exact in float32 stay
// exact, and exact ties between boxes resolve as in exact arithmetic.
__device__ __forceinline__ float round_sig11(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7F800000u) == 0x7F800000u) return x;      // inf / nan
    u = (u + 0x1000u) & ~0x1FFFu;
    return __uint_as_float(u);
}


// BF16 storage of subband-pair rows (the reduced-precision partial sums): the
// float32 pair sum rounded to nearest even, read back by a 16-bit shift.
struct bf16s {
    uint16_t u;
    __device__ __forceinline__ operator float() const { return __uint_as_float((uint32_t)u << 16); }
};
__device__ __forceinline__ bf16s to_bf16(float f) {
    const uint32_t x = __float_as_uint(f);
    return bf16s{(uint16_t)((x + 0x7FFFu + ((x >> 16) & 1u)) >> 16)};
}

// Centred D'(t) = D(t) - bias of DM d from the term rows, in the main path's
// order; t_rel = t - (first centre of the tile).
template <typename E>
__device__ __forceinline__ float terms_d(const E *X, const int32_t *tbase, int n_terms, int n_dm, int d,
                                         int t_rel, float bias) {
    float acc = (float)X[tbase[d] + t_rel] - bias;
    for (int g = 1; g < n_terms; ++g) acc += (float)X[tbase[g * n_dm + d] + t_rel];
    return acc;
}

// Pw[OFF + j] (=|+=) (sum over terms g of X[base_g + rel + j]) - bias, j < N,
// summed as ((x_0 - bias) + x_1) + ...  With MAXT > 0 the term bases sit in
// this lane's LDS slots (tb[h * stride], two 16-bit bases per word) and the
// term loop is unrolled; otherwise they are read from global memory term by
// term.  EXACT: the plan has exactly MAXT terms (software-pipelined reads).
template <int MAXT, int N, int OFF, bool EXACT, typename E, int NP>
__device__ __forceinline__ void sum_terms(float (&Pw)[NP], const E *X, const uint32_t *tb, int stride,
                                          const int32_t *tbase, int n_terms, int n_dm, int d, int rel,
                                          float bias) {
    if constexpr (EXACT) {
        // exactly MAXT terms: term g + 1's N samples are read into one register
        // buffer while term g's are added from the other, so a wave keeps N / 2
        // LDS reads in flight behind its adds (the scheduling barriers keep the
        // two phases apart).  Same summation order as the guarded loop.
        uint32_t tw[(MAXT + 1) / 2];
#pragma unroll
        for (int h = 0; h < (MAXT + 1) / 2; ++h) tw[h] = tb[h * stride];
        float buf[2][N];
        {
            const E *x0 = X + (int)(tw[0] & 0xFFFF) + rel;
#pragma unroll
            for (int j = 0; j < N; ++j) buf[0][j] = (float)x0[j];
        }
#pragma unroll
        for (int g = 0; g < MAXT; ++g) {
            if (g + 1 < MAXT) {
                const E *xn = X + (int)((tw[(g + 1) >> 1] >> (16 * ((g + 1) & 1))) & 0xFFFF) + rel;
#pragma unroll
                for (int j = 0; j < N; ++j) buf[(g + 1) & 1][j] = (float)xn[j];
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int j = 0; j < N; ++j) {
                if (g == 0) Pw[OFF + j] = buf[0][j] - bias;
                else Pw[OFF + j] += buf[g & 1][j];
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    } else if constexpr (MAXT > 0) {
        {
            const E *x0 = X + (int)(tb[0] & 0xFFFF) + rel;
#pragma unroll
            for (int j = 0; j < N; ++j) Pw[OFF + j] = (float)x0[j] - bias;
        }
#pragma unroll
        for (int g = 1; g < MAXT; ++g) {
            if (g < n_terms) {
                const E *xg = X + (int)((tb[(g >> 1) * stride] >> (16 * (g & 1))) & 0xFFFF) + rel;
#pragma unroll
                for (int j = 0; j < N; ++j) Pw[OFF + j] += (float)xg[j];
            }
        }
    } else {
        {
            const E *x0 = X + tbase[d] + rel;
#pragma unroll
            for (int j = 0; j < N; ++j) Pw[OFF + j] = (float)x0[j] - bias;
        }
        for (int g = 1; g < n_terms; ++g) {
            const E *xg = X + tbase[g * n_dm + d] + rel;
#pragma unroll
            for (int j = 0; j < N; ++j) Pw[OFF + j] += (float)xg[j];
        }
    }
}

template <typename T, int NW, bool PAIRED, int MAXT, bool ASYNC = false, bool QB = false, bool EXACT = false>
__global__ __launch_bounds__(DD_THREADS, 2) void dedisp_kernel(DedispParams P) {
    static_assert(!QB || (PAIRED && !ASYNC), "BF16 rows are a layout of the pair-row plan");
    static_assert(!EXACT || MAXT > 0, "an exact term count is a register-base instance");
    using XE = std::conditional_t<QB, bf16s, float>;  // element of the rows the lanes sum
    constexpr int WMAX = 1 << (NW - 1);
    constexpr int H2 = WMAX / 2;
    constexpr int WW = 2 * H2;
    // centres per register pass: 16 for the pipelined exact-term sums (their
    // second read buffer takes the registers of the smaller prefix window)
    constexpr int SUB = EXACT ? DD_SUB / 2 : DD_SUB;
    constexpr int NP = SUB + WW + 1;                 // register prefix window
    extern __shared__ float lds[];
    float *Q = lds;                                   // [q_total] float32 pair rows (ASYNC: window buffer 1)
    float *R = lds + P.q_total;                       // [r_total] subband windows (ASYNC: window buffer 0)
    XE *Qx = reinterpret_cast<XE *>(lds);              // pair rows in their storage type
    const XE *X = PAIRED ? Qx : reinterpret_cast<const XE *>(R);   // rows the lanes sum (ASYNC: this tile's buffer)
    const float *Rw = R;                              // windows the pair rows are built from
    const T *ring = reinterpret_cast<const T *>(P.ring);
    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    const int d = blockIdx.y * DD_THREADS + tid;      // one lane per DM
    const bool active = d < P.n_dm;
    const int64_t c_hi = P.c_lo + P.n_centres;

    // term bases of this lane, two 16-bit LDS indices per word, in LDS slots
    // after the windows (kept out of registers: the chunk loop needs them all)
    constexpr int TBW = MAXT > 0 ? (MAXT + 1) / 2 : 1;
    constexpr int RBUFS = (PAIRED && ASYNC) ? 2 : 1;   // pair rows with LDS-direct loads: two window buffers
    uint32_t *tb = reinterpret_cast<uint32_t *>(lds + P.q_total + RBUFS * P.r_total) + tid;
    // subband window table (offset, shmax, length) copied to LDS, so the staging
    // loop's table reads never wait behind the window loads in flight
    int *si = reinterpret_cast<int *>(lds + P.q_total + RBUFS * P.r_total) + (MAXT > 0 ? TBW * nthr : 0);
    for (int i = tid; i < 3 * P.n_sub; i += nthr) si[i] = P.sub_info[i];
    if constexpr (MAXT > 0) {
#pragma unroll
        for (int h = 0; h < TBW; ++h) {
            uint32_t lo = 0, hi = 0;
            if (active && 2 * h < P.n_terms) lo = (uint32_t)P.tbase[(2 * h) * P.n_dm + d];
            if (active && 2 * h + 1 < P.n_terms) hi = (uint32_t)P.tbase[(2 * h + 1) * P.n_dm + d];
            tb[h * nthr] = lo | (hi << 16);
        }
    }
    const int lane = tid & 63;
    const int n_waves = nthr >> 6;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int64_t plane = (int64_t)P.n_pix * P.row_len;
    __syncthreads();                                      // si[] visible
    // fast staging: a wave owns at most DD_STAGE_SUBS subbands (s = wave +
    // u * n_waves) whose windows span at most DD_STAGE_ITERS x 64 steps
    bool fast = MAXT > 0 && P.n_sub <= n_waves * DD_STAGE_SUBS;
    for (int s = 0; s < P.n_sub; ++s) fast = fast && si[3 * s + 2] <= 64 * DD_STAGE_ITERS;

    // ASYNC (float32 history): the windows of the next tile -- same pixel, or
    // the first tile of the work-group's next pixel -- are copied from HBM by
    // LDS-direct loads (global_load_lds_dword: lane l fills slot base + l, no
    // registers hold the data) while the current tile computes, into the
    // other of two window buffers (with pair rows: the buffer the next tile's
    // rows are built from).  Steps outside the history are zero-filled by
    // ordinary stores.
    auto dma_stage = [&](int pp, int cc0, float *dst) {
        const int64_t pos0 = (int64_t)P.hist + (P.c_lo + cc0) - H2 - P.T0;
        const float *ring_p = reinterpret_cast<const float *>(P.ring) + (int64_t)pp * P.row_len;
        for (int s = wave; s < P.n_sub; s += n_waves) {
            const int off = si[3 * s];
            const int len = si[3 * s + 2];
            const int64_t q0 = pos0 - si[3 * s + 1];
            const float *row = ring_p + (int64_t)s * plane;
            for (int e0 = 0; e0 < len; e0 += 64) {
                const int e = e0 + lane;
                const int64_t pos = q0 + e;
                const bool in = e < len;
                if (in && pos >= 0 && pos < P.row_len)
                    __builtin_amdgcn_global_load_lds(
                        (const __attribute__((address_space(1))) void *)(row + pos),
                        (__attribute__((address_space(3))) void *)(dst + off + e0), 4, 0, 0);
                else if (in)
                    dst[off + e] = 0.f;
            }
        }
    };
    int tile_no = 0;
    if constexpr (ASYNC) {
        if ((int)blockIdx.x < P.n_pix) dma_stage(blockIdx.x, 0, R);
    }

    for (int p = blockIdx.x; p < P.n_pix; p += gridDim.x) {
        // ---- per-lane state from earlier blocks (float64)
        // Samples are centred by `bias` (float32).  A box of width w_k then
        // scores snr_k = (B'_k + w_k (bias - mean)) / (sigma sqrt(w_k)); w_k times
        // the residual and 1 / (sigma sqrt(w_k)) are power-of-two scalings of
        // the k = 0 / k = 1 values (exact in binary floating point), so three
        // floats carry them: rmu_k = 2^k rmu0, inv_k = 2^-(k/2) inv_{k mod 2}
        float bias = 0.f, rmu0 = 0.f, inv0 = 0.f, inv1 = 0.f;
        bool score = false;
        int vf = 0;
        if (active) {
            vf = P.valid_from[d];
            const double n_a = P.n_prev[d];
            const double mean_a = P.mean[(int64_t)p * P.n_dm + d];
            const double m2_a = P.m2[(int64_t)p * P.n_dm + d];
            const double sig = n_a > 0.0 ? sqrt(m2_a / n_a) : 0.0;
            score = (n_a >= P.min_count) && (n_a > 0.0) && (sig > 0.0);
            bias = round_sig11((float)mean_a);
            rmu0 = (float)((double)bias - mean_a);
            inv0 = score ? (float)(1.0 / sig) : 0.f;
            inv1 = score ? (float)(1.0 / (sig * sqrt(2.0))) : 0.f;
        }
        double s1 = 0.0, s2 = 0.0;
        float m1_tail = -INFINITY;    // max of D over [c, c + H2) for the next chunk
        float Pw[NP];
#pragma unroll
        for (int x = 0; x < NP; ++x) Pw[x] = 0.f;

        for (int c0 = 0; c0 < P.n_centres; c0 += P.tt) {
            const int64_t cabs0 = P.c_lo + c0;
            // ---- stage every subband's window (ring dtype): step tau of subband s
            //      lands at R[off_s + tau - (cabs0 - H2 - shmax_s)]
            if constexpr (ASYNC) {
                // this tile's loads were issued one tile ago: wait for this wave's,
                // then for every wave's (and for every wave to leave the other buffer)
                __builtin_amdgcn_s_waitcnt(0);
                __syncthreads();
                int np = p, nc = c0 + P.tt;
                if (nc >= P.n_centres) {
                    np = p + (int)gridDim.x;
                    nc = 0;
                }
                if constexpr (!PAIRED) {
                    X = (tile_no & 1) ? Q : R;
                    if (np < P.n_pix) dma_stage(np, nc, (tile_no & 1) ? R : Q);
                } else {
                    // pair rows: this tile's rows are built from one window buffer
                    // while the next tile's windows stream into the other
                    Rw = (tile_no & 1) ? R + P.r_total : R;
                    if (np < P.n_pix) dma_stage(np, nc, (tile_no & 1) ? R : R + P.r_total);
                }
                ++tile_no;
            } else {
            __syncthreads();
            const int64_t pos0 = (int64_t)P.hist + cabs0 - H2 - P.T0;   // ring position of e - shmax_s = 0
            const T *ring_p = ring + (int64_t)p * P.row_len;
            if (MAXT > 0 && fast) {
                // each wave issues the loads of all its subbands before its first
                // store (unconditional loads: clamped index, zero selected
                // afterwards), so they are in flight together
                float v[MAXT > 0 ? DD_STAGE_SUBS : 1][MAXT > 0 ? DD_STAGE_ITERS : 1];
#pragma unroll
                for (int u = 0; u < (MAXT > 0 ? DD_STAGE_SUBS : 1); ++u) {
                    const int s = wave + u * n_waves;
                    const bool live = s < P.n_sub;
                    const int len = live ? si[3 * s + 2] : 0;
                    const T *row = ring_p + (int64_t)(live ? s : 0) * plane;
                    const int64_t q0 = pos0 - (live ? si[3 * s + 1] : 0);
#pragma unroll
                    for (int k = 0; k < (MAXT > 0 ? DD_STAGE_ITERS : 1); ++k) {
                        const int e = lane + 64 * k;
                        const int64_t pos = q0 + e;
                        const bool ok = e < len && pos >= 0 && pos < P.row_len;
                        const T x = row[ok ? pos : 0];
                        v[u][k] = ok ? (float)x : 0.f;
                    }
                }
#pragma unroll
                for (int u = 0; u < (MAXT > 0 ? DD_STAGE_SUBS : 1); ++u) {
                    const int s = wave + u * n_waves;
                    if (s < P.n_sub) {
                        const int off = si[3 * s];
                        const int len = si[3 * s + 2];
#pragma unroll
                        for (int k = 0; k < (MAXT > 0 ? DD_STAGE_ITERS : 1); ++k)
                            if (lane + 64 * k < len) R[off + lane + 64 * k] = v[u][k];
                    }
                }
            } else {
                // generic staging: each wave stages whole subbands (wave-uniform s),
                // four loads in flight per lane
                for (int s = wave; s < P.n_sub; s += n_waves) {
                    const int off = si[3 * s];
                    const int len = si[3 * s + 2];
                    const T *row = ring_p + (int64_t)s * plane;
                    const int64_t q0 = pos0 - si[3 * s + 1];      // ring position of e = 0
                    for (int e0 = lane; e0 < len; e0 += 4 * 64) {
                        float v[4];
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            const int e = e0 + 64 * k;
                            const int64_t pos = q0 + e;
                            const bool ok = e < len && pos >= 0 && pos < P.row_len;
                            const T x = row[ok ? pos : 0];
                            v[k] = ok ? (float)x : 0.f;
                        }
#pragma unroll
                        for (int k = 0; k < 4; ++k)
                            if (e0 + 64 * k < len) R[off + e0 + 64 * k] = v[k];
                    }
                }
            }
            __syncthreads();
            }
            // ---- one float32 row per (pair, delta): Q[e] = R[a0 + e] + R[b0 + e];
            //      each wave builds every n_waves-th row
            if (PAIRED) {
                for (int r = wave; r < P.n_rows; r += n_waves) {
                    const int qoff = P.rows[6 * r], len = P.rows[6 * r + 1];
                    const int a0 = P.rows[6 * r + 2], b0 = P.rows[6 * r + 3];
                    const int blo = P.rows[6 * r + 4], bhi = P.rows[6 * r + 5];
#pragma unroll 4
                    for (int e = lane; e < len; e += 64) {
                        const int ib = b0 + e;
                        const bool okb = ib >= blo && ib < bhi;
                        const float va = Rw[a0 + e];
                        const float vb = Rw[okb ? ib : a0 + e];
                        const float v = okb ? va + vb : va;
                        if constexpr (QB) Qx[qoff + e] = to_bf16(v);
                        else Q[qoff + e] = v;
                    }
                }
            }
            __syncthreads();
            if (!active) continue;

            if (c0 == 0) {
                // prologue: D(c_lo - H2 + x), x < WW, into Pw[x + 1]
                if (WW > 0) sum_terms<MAXT, WW, 1, false>(Pw, X, tb, nthr, P.tbase, P.n_terms, P.n_dm, d, -H2, bias);
                // samples t = c_lo + (x - H2) for x >= H2 belong to this block
                float a1 = 0.f, a2 = 0.f, m = -INFINITY;
#pragma unroll
                for (int x = H2; x < WW; ++x) {
                    const float v = Pw[x + 1];
                    const int64_t t = P.c_lo + (x - H2);
                    const bool ok = t >= vf && t < c_hi;
                    a1 += ok ? v : 0.f;
                    a2 = fmaf(ok ? v : 0.f, v, a2);
                    m = fmaxf(m, v);
                }
                s1 += (double)a1;
                s2 += (double)a2;
                m1_tail = m;
                Pw[0] = 0.f;
#pragma unroll
                for (int x = 0; x < WW; ++x) Pw[x + 1] += Pw[x];
            }

            for (int cc = 0; cc < P.tt; cc += DD_CHUNK) {
                const int64_t c = cabs0 + cc;                  // first centre of the 64-centre chunk
                const float pbase = Pw[0];                       // P[c - H2], for the rescan
                const bool scored = score && c - H2 >= vf;
                float mb[NW];
#pragma unroll
                for (int k = 0; k < NW; ++k) mb[k] = -INFINITY;
                float m1n = -INFINITY;                           // width-1 maximum of the next chunk
                mb[0] = m1_tail;
#pragma unroll
                for (int j = 0; j < DD_CHUNK / SUB; ++j) {
                    const int cs = cc + j * SUB;                  // sub-chunk start (tile relative)
                    // ---- D(c + j S + H2 + i), i < S, summed over the term rows
                    sum_terms<MAXT, SUB, WW + 1, EXACT>(Pw, X, tb, nthr, P.tbase, P.n_terms, P.n_dm, d, cs + H2,
                                                    bias);
                    // ---- statistics of this block's complete samples
                    const int64_t t1 = c + j * SUB + H2;
                    float a1 = 0.f, a2 = 0.f;
                    const bool full = t1 >= vf && t1 + SUB <= c_hi;
                    if (full) {
#pragma unroll
                        for (int i = 0; i < SUB; ++i) a2 = fmaf(Pw[WW + 1 + i], Pw[WW + 1 + i], a2);
                    } else {
#pragma unroll
                        for (int i = 0; i < SUB; ++i) {
                            const float v = Pw[WW + 1 + i];
                            const bool ok = t1 + i >= vf && t1 + i < c_hi;
                            a1 += ok ? v : 0.f;
                            a2 = fmaf(ok ? v : 0.f, v, a2);
                        }
                    }
                    // ---- width-1 maxima from the fresh samples: t < c + 64 belong to this chunk
                    const int split = DD_CHUNK - H2 - j * SUB;
#pragma unroll
                    for (int i = 0; i < SUB; ++i) {
                        if (i < split) mb[0] = fmaxf(mb[0], Pw[WW + 1 + i]);
                        else m1n = fmaxf(m1n, Pw[WW + 1 + i]);
                    }
                    // ---- prefix sums in place
#pragma unroll
                    for (int i = 0; i < SUB; ++i) Pw[WW + 1 + i] += Pw[WW + i];
                    if (full) a1 = Pw[WW + SUB] - Pw[WW];
                    s1 += (double)a1;
                    s2 += (double)a2;
                    // ---- boxcar maxima of the wider widths over this sub-chunk's centres
                    if (scored) {
#pragma unroll
                        for (int k = 1; k < NW; ++k) {
                            const int w = 1 << k;
                            const int lo = H2 - (w >> 1);
                            float m = mb[k];
#pragma unroll
                            for (int i = 0; i < SUB; ++i) m = fmaxf(m, Pw[i + lo + w] - Pw[i + lo]);
                            mb[k] = m;
                        }
                    }
                    // ---- slide the prefix window by one sub-chunk
#pragma unroll
                    for (int x = 0; x <= WW; ++x) Pw[x] = Pw[x + SUB];
                }
                m1_tail = m1n;

                // ---- chunk maximum over widths and candidate
                if (scored) {
                    float best = -INFINITY;
                    int bk = 0;
#pragma unroll
                    for (int k = 0; k < NW; ++k) {
                        const float snr = (mb[k] + ldexpf(rmu0, k)) * ldexpf((k & 1) ? inv1 : inv0, -(k >> 1));
                        if (snr > best) {
                            best = snr;
                            bk = k;
                        }
                    }
                    if (best > P.threshold) {
                        // rare path: earliest centre of the winning width.  D and the
                        // prefix chain are recomputed from the LDS rows in the main
                        // path's order, so the boxcars are bit-identical to the maxima.
                        const float target = mb[bk];
                        int bi = 0;
                        if (bk == 0) {
                            for (int i = 0; i < DD_CHUNK; ++i) {
                                if (terms_d(X, P.tbase, P.n_terms, P.n_dm, d, cc + i, bias) == target) {
                                    bi = i;
                                    break;
                                }
                            }
                        } else {
                            const int w = 1 << bk;
                            const int lo = H2 - (w >> 1);
                            // chains at prefix positions x = i + lo and x = i + lo + w, where
                            // position x is P[c - H2 + x] = pbase + sum_{x' < x} D(c - H2 + x')
                            float pa = pbase, pb = pbase;
                            for (int x = 0; x < lo + w; ++x) {
                                const float dv = terms_d(X, P.tbase, P.n_terms, P.n_dm, d, cc - H2 + x, bias);
                                if (x < lo) pa = dv + pa;
                                pb = dv + pb;
                            }
                            for (int i = 0; i < DD_CHUNK; ++i) {
                                if (pb - pa == target) {
                                    bi = i;
                                    break;
                                }
                                pa = terms_d(X, P.tbase, P.n_terms, P.n_dm, d, cc - H2 + lo + i, bias) + pa;
                                pb = terms_d(X, P.tbase, P.n_terms, P.n_dm, d, cc - H2 + lo + w + i, bias) + pb;
                            }
                        }
                        const int slot = atomicAdd(P.n_cand, 1);
                        if (slot < P.cand_cap) {
                            Cand cd;
                            cd.t = (int32_t)(c + bi - P.c_lo);
                            cd.p = p;
                            cd.d = (int16_t)d;
                            cd.k = (int16_t)bk;
                            cd.snr = best;
                            P.cand[slot] = cd;
                        }
                    }
                }
            }
        }
        // ---- merge this block's statistics into the running state
        // (s1, s2 are sums of the centred samples D - bias)
        if (active && P.merge) {
            const double nb = P.blk_count[d];
            if (nb > 0.0) {
                // earlier blocks' state, read again rather than kept in registers
                const int64_t i = (int64_t)p * P.n_dm + d;
                const double n_a = P.n_prev[d];
                const double mean_a = P.mean[i];
                const double m2_a = P.m2[i];
                const double m2_b = s2 - s1 * s1 / nb;
                const double n = n_a + nb;
                const double delta = ((double)bias - mean_a) + s1 / nb;     // mean_b - mean_a
                P.mean[i] = mean_a + delta * (nb / n);
                P.m2[i] = m2_a + m2_b + delta * delta * (n_a * nb / n);
            }
        }
    }
}

// Centre of a DM's first block: for each (pixel, DM) with no state yet
// (n_prev[d] == 0), the mean of its complete samples in the block,
//     D(t) = sum_s row_s[hist + t - shift[s][d] - T0],  max(valid_from[d], c_lo) <= t < c_hi,
// written as the running mean.  A work-group owns one pixel and 256 DMs and
// walks each subband's row in chunks: the chunk's exclusive prefix sums
// (float64, LDS) give every lane the two ends of its range, so a DM costs two
// lookups per subband.  With no samples behind it the mean only centres the
// block's sums; the merge weighs it by 0.
constexpr int DD_PRIME_CHUNK = 2048;          // row positions per LDS prefix
constexpr int DD_PRIME_PER_THREAD = DD_PRIME_CHUNK / DD_THREADS;

template <typename T>
__global__ __launch_bounds__(DD_THREADS) void dedisp_prime_kernel(const T *ring, int n_sub, int n_pix, int hist,
                                                                  int row_len, int64_t T0, const int32_t *shift,
                                                                  int n_dm, const int32_t *valid_from,
                                                                  const double *n_prev, int64_t c_lo,
                                                                  int n_centres, double *mean) {
    __shared__ double pre[DD_PRIME_CHUNK + 1];     // pre[i] = sum of the chunk's positions < i
    __shared__ double part[DD_THREADS];
    const int tid = threadIdx.x;
    const int d = blockIdx.y * DD_THREADS + tid;
    const int p = blockIdx.x;
    const int64_t c_hi = c_lo + n_centres;
    bool need = d < n_dm && n_prev[d] == 0.0;
    const int64_t t_b = need ? (valid_from[d] > c_lo ? (int64_t)valid_from[d] : c_lo) : c_hi;
    need = need && t_b < c_hi;
    double acc = 0.0;
    for (int s = 0; s < n_sub; ++s) {
        const T *row = ring + ((int64_t)s * n_pix + p) * row_len;
        const int sh = need ? shift[(int64_t)s * n_dm + d] : 0;
        // this lane's range of row positions [a, b)
        const int64_t a = hist + t_b - sh - T0, b = hist + c_hi - sh - T0;
        double carry = 0.0;
        for (int c0 = 0; c0 < row_len; c0 += DD_PRIME_CHUNK) {
            // exclusive prefix of positions [c0, c0 + DD_PRIME_CHUNK)
            double v[DD_PRIME_PER_THREAD];
            double run = 0.0;
#pragma unroll
            for (int k = 0; k < DD_PRIME_PER_THREAD; ++k) {
                const int pos = c0 + tid * DD_PRIME_PER_THREAD + k;
                v[k] = pos < row_len ? (double)row[pos] : 0.0;
                run += v[k];
            }
            part[tid] = run;
            __syncthreads();
            for (int off = 1; off < DD_THREADS; off <<= 1) {       // inclusive scan of thread totals
                const double x = tid >= off ? part[tid - off] : 0.0;
                __syncthreads();
                part[tid] += x;
                __syncthreads();
            }
            double q = part[tid] - run;                              // exclusive
#pragma unroll
            for (int k = 0; k < DD_PRIME_PER_THREAD; ++k) {
                pre[tid * DD_PRIME_PER_THREAD + k] = q;
                q += v[k];
            }
            if (tid == DD_THREADS - 1) pre[DD_PRIME_CHUNK] = q;
            __syncthreads();
            if (need) {
                if (a >= c0 && a < c0 + DD_PRIME_CHUNK) acc -= carry + pre[a - c0];
                if (b >= c0 && b < c0 + DD_PRIME_CHUNK) acc += carry + pre[b - c0];
            }
            carry += pre[DD_PRIME_CHUNK];
            __syncthreads();                                         // pre and part are reused
        }
        if (need && b == row_len) acc += carry;
    }
    if (need) mean[(int64_t)p * n_dm + d] = acc / (double)(c_hi - t_b);
}

// Plans with at most 8 (or 16) terms keep the lanes' term bases in registers.
template <typename T, int NW, bool PAIRED, int MAXT, bool ASYNC = false, bool QB = false, bool EXACT = false>
static int launch_maxt(const DedispParams &P, int threads, size_t lds, dim3 grid, hipStream_t s) {
    (void)hipFuncSetAttribute((const void *)dedisp_kernel<T, NW, PAIRED, MAXT, ASYNC, QB, EXACT>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    dedisp_kernel<T, NW, PAIRED, MAXT, ASYNC, QB, EXACT><<<grid, threads, lds, s>>>(P);
    return check_launch("dsx_dedisp_detect");
}

// float32 history with LDS-direct window loads: unpaired plan with two window
// buffers, or pair rows with one window buffer
template <int NW, bool PAIRED>
static int launch_async(const DedispParams &P, int threads, size_t lds, dim3 grid, hipStream_t s) {
    // exact term counts: the pipelined sums
    if constexpr (!PAIRED) {
        if (P.n_terms == 16) return launch_maxt<float, NW, PAIRED, 16, true, false, true>(P, threads, lds, grid, s);
    } else {
        if (P.n_terms == 8) return launch_maxt<float, NW, PAIRED, 8, true, false, true>(P, threads, lds, grid, s);
    }
    if (P.n_terms <= 8) return launch_maxt<float, NW, PAIRED, 8, true>(P, threads, lds, grid, s);
    if (P.n_terms <= 16) return launch_maxt<float, NW, PAIRED, 16, true>(P, threads, lds, grid, s);
    return launch_maxt<float, NW, PAIRED, 0, true>(P, threads, lds, grid, s);
}

template <bool PAIRED>
static int dispatch_async(int nw, const DedispParams &P, int threads, size_t lds, dim3 grid, hipStream_t s) {
    switch (nw) {
        case 1: return launch_async<1, PAIRED>(P, threads, lds, grid, s);
        case 2: return launch_async<2, PAIRED>(P, threads, lds, grid, s);
        case 3: return launch_async<3, PAIRED>(P, threads, lds, grid, s);
        case 4: return launch_async<4, PAIRED>(P, threads, lds, grid, s);
        case 5: return launch_async<5, PAIRED>(P, threads, lds, grid, s);
        case 6: return launch_async<6, PAIRED>(P, threads, lds, grid, s);
        case 7: return launch_async<7, PAIRED>(P, threads, lds, grid, s);
        default: return set_error(DSX_ENOTSUP, "dsx_dedisp_detect: n_widths=%d not in [1, 7]", nw);
    }
}

template <typename T, int NW, bool PAIRED, bool QB>
static int launch_dedisp(const DedispParams &P, int threads, size_t lds, dim3 grid, hipStream_t s) {
    if constexpr (!QB && !PAIRED) {
        if (P.n_terms == 16) return launch_maxt<T, NW, PAIRED, 16, false, false, true>(P, threads, lds, grid, s);
    } else if constexpr (!QB) {
        if (P.n_terms == 8) return launch_maxt<T, NW, PAIRED, 8, false, false, true>(P, threads, lds, grid, s);
    }
    if (P.n_terms <= 8) return launch_maxt<T, NW, PAIRED, 8, false, QB>(P, threads, lds, grid, s);
    if (P.n_terms <= 16) return launch_maxt<T, NW, PAIRED, 16, false, QB>(P, threads, lds, grid, s);
    return launch_maxt<T, NW, PAIRED, 0, false, QB>(P, threads, lds, grid, s);
}

template <typename T, bool PAIRED, bool QB = false>
static int dispatch_nw(int nw, const DedispParams &P, int threads, size_t lds, dim3 grid, hipStream_t s) {
    switch (nw) {
        case 1: return launch_dedisp<T, 1, PAIRED, QB>(P, threads, lds, grid, s);
        case 2: return launch_dedisp<T, 2, PAIRED, QB>(P, threads, lds, grid, s);
        case 3: return launch_dedisp<T, 3, PAIRED, QB>(P, threads, lds, grid, s);
        case 4: return launch_dedisp<T, 4, PAIRED, QB>(P, threads, lds, grid, s);
        case 5: return launch_dedisp<T, 5, PAIRED, QB>(P, threads, lds, grid, s);
        case 6: return launch_dedisp<T, 6, PAIRED, QB>(P, threads, lds, grid, s);
        case 7: return launch_dedisp<T, 7, PAIRED, QB>(P, threads, lds, grid, s);
        default: return set_error(DSX_ENOTSUP, "dsx_dedisp_detect: n_widths=%d not in [1, 7]", nw);
    }
}

}  // namespace dsx

using namespace dsx;

extern "C" int dsx_dedisp_detect(const void *ring, int ring_dtype, int n_sub, int n_pix, int hist,
                                 int ring_len, int64_t T0, const int32_t *sub_info, const int32_t *rows,
                                 int n_rows, const int32_t *tbase, int n_terms, int n_dm,
                                 const int32_t *valid_from, const double *n_prev, const double *blk_count,
                                 double *mean, double *m2, int merge, int64_t c_lo, int n_centres,
                                 int n_widths, int tile, int r_total, int q_total, int flags, float threshold,
                                 double min_count, void *cand, int cand_cap, int32_t *n_cand,
                                 dsx_stream_t stream) {
    DSX_REQUIRE(ring && sub_info && tbase && valid_from && n_prev && blk_count && mean && m2 && cand && n_cand,
                "dsx_dedisp_detect: null pointer");
    DSX_REQUIRE(ring_dtype == DSX_F16 || ring_dtype == DSX_F32, "dsx_dedisp_detect: ring_dtype");
    DSX_REQUIRE(n_sub > 0 && n_pix > 0 && n_dm > 0 && hist >= 0 && ring_len > hist,
                "dsx_dedisp_detect: bad shape");
    DSX_REQUIRE(n_dm <= 32767, "dsx_dedisp_detect: n_dm=%d > 32767", n_dm);
    const bool paired = n_rows > 0;
    const bool async = (flags & DSX_DD_DMA) != 0;
    const bool row_bf16 = (flags & DSX_DD_ROWS_BF16) != 0;
    DSX_REQUIRE((flags & ~(DSX_DD_DMA | DSX_DD_ROWS_BF16)) == 0, "dsx_dedisp_detect: flags=%d", flags);
    DSX_REQUIRE(paired ? (rows && n_terms == (n_sub + 1) / 2 && q_total > 0 && n_rows <= DD_MAX_ROWS)
                       : (n_terms == n_sub && !row_bf16 && (async ? q_total == r_total : q_total == 0)),
                "dsx_dedisp_detect: n_rows=%d / n_terms=%d / q_total=%d / flags=%d do not describe a plan for n_sub=%d",
                n_rows, n_terms, q_total, flags, n_sub);
    DSX_REQUIRE(!(row_bf16 && async), "dsx_dedisp_detect: BF16 pair rows are not combined with LDS-direct loads");
    DSX_REQUIRE(!async || ring_dtype == DSX_F32, "dsx_dedisp_detect: LDS-direct window loads need a float32 history");
    DSX_REQUIRE(n_centres > 0 && n_centres % DD_CHUNK == 0,
                "dsx_dedisp_detect: n_centres must be a positive multiple of 64");
    DSX_REQUIRE(tile > 0 && tile % DD_CHUNK == 0 && n_centres % tile == 0,
                "dsx_dedisp_detect: tile must be a multiple of 64 dividing n_centres");
    DSX_REQUIRE(n_widths >= 1 && n_widths <= 7, "dsx_dedisp_detect: n_widths=%d", n_widths);
    DSX_REQUIRE(cand_cap >= 0 && r_total > 0, "dsx_dedisp_detect: sizes");
    const int threads = std::min(DD_THREADS, ((n_dm + 63) / 64) * 64);
    // pair rows with LDS-direct loads keep two window buffers after the rows
    const size_t r_bufs = (paired && async) ? 2 : 1;
    const size_t lds =
        ((size_t)q_total + r_bufs * (size_t)r_total + (size_t)tb_words(n_terms) * threads + 3 * (size_t)n_sub) * 4;
    if (lds > 160 * 1024)
        return set_error(DSX_ERANGE, "dsx_dedisp_detect: %zu B of LDS for the time windows", lds);
    DedispParams P{};
    P.ring = ring;
    P.sub_info = sub_info;
    P.rows = rows;
    P.tbase = tbase;
    P.valid_from = valid_from;
    P.n_prev = n_prev;
    P.blk_count = blk_count;
    P.mean = mean;
    P.m2 = m2;
    P.merge = merge != 0;
    P.cand = reinterpret_cast<Cand *>(cand);
    P.n_cand = n_cand;
    P.T0 = T0;
    P.c_lo = c_lo;
    P.n_sub = n_sub;
    P.n_pix = n_pix;
    P.hist = hist;
    P.row_len = ring_len;
    P.n_dm = n_dm;
    P.n_rows = n_rows;
    P.n_terms = n_terms;
    P.n_centres = n_centres;
    P.tt = tile;
    P.r_total = r_total;
    P.q_total = q_total;
    P.cand_cap = cand_cap;
    P.threshold = threshold;
    P.min_count = min_count;
    const dim3 grid((unsigned)std::min(n_pix, 1 << 20), (unsigned)((n_dm + DD_THREADS - 1) / DD_THREADS));
    hipStream_t s = as_stream(stream);
    if (async) return paired ? dispatch_async<true>(n_widths, P, threads, lds, grid, s)
                             : dispatch_async<false>(n_widths, P, threads, lds, grid, s);
    if (row_bf16)
        return ring_dtype == DSX_F16 ? dispatch_nw<_Float16, true, true>(n_widths, P, threads, lds, grid, s)
                                     : dispatch_nw<float, true, true>(n_widths, P, threads, lds, grid, s);
    if (ring_dtype == DSX_F16)
        return paired ? dispatch_nw<_Float16, true>(n_widths, P, threads, lds, grid, s)
                      : dispatch_nw<_Float16, false>(n_widths, P, threads, lds, grid, s);
    return paired ? dispatch_nw<float, true>(n_widths, P, threads, lds, grid, s)
                  : dispatch_nw<float, false>(n_widths, P, threads, lds, grid, s);
}

extern "C" int dsx_dedisp_prime(const void *ring, int ring_dtype, int n_sub, int n_pix, int hist, int ring_len,
                                int64_t T0, const int32_t *shift, int n_dm, const int32_t *valid_from,
                                const double *n_prev, int64_t c_lo, int n_centres, double *mean,
                                dsx_stream_t stream) {
    DSX_REQUIRE(ring && shift && valid_from && n_prev && mean, "dsx_dedisp_prime: null pointer");
    DSX_REQUIRE(ring_dtype == DSX_F16 || ring_dtype == DSX_F32, "dsx_dedisp_prime: ring_dtype");
    DSX_REQUIRE(n_sub > 0 && n_pix > 0 && n_dm > 0 && hist >= 0 && ring_len > hist && n_centres > 0,
                "dsx_dedisp_prime: bad shape");
    // samples t in [c_lo, c_lo + n_centres) sit at ring positions hist + t - shift - T0
    DSX_REQUIRE(hist + (c_lo - T0) >= 0 && hist + (c_lo + n_centres - T0) <= (int64_t)ring_len,
                "dsx_dedisp_prime: centres [c_lo, c_lo + n_centres) are not held by the ring");
    DSX_REQUIRE(n_dm <= 32767, "dsx_dedisp_prime: n_dm=%d > 32767", n_dm);
    const dim3 grid((unsigned)n_pix, (unsigned)((n_dm + DD_THREADS - 1) / DD_THREADS));
    hipStream_t s = as_stream(stream);
    if (ring_dtype == DSX_F16)
        dedisp_prime_kernel<_Float16><<<grid, DD_THREADS, 0, s>>>(
            reinterpret_cast<const _Float16 *>(ring), n_sub, n_pix, hist, ring_len, T0, shift, n_dm, valid_from,
            n_prev, c_lo, n_centres, mean);
    else
        dedisp_prime_kernel<float><<<grid, DD_THREADS, 0, s>>>(
            reinterpret_cast<const float *>(ring), n_sub, n_pix, hist, ring_len, T0, shift, n_dm, valid_from,
            n_prev, c_lo, n_centres, mean);
    return check_launch("dsx_dedisp_prime");
}

// Exponential forgetting of the running statistics: after a block's merge,
// the state's weight n_prev[d] and M2 m2[p][d] are multiplied by `decay`
// (the mean is a ratio and stays).  Each earlier block then weighs
// decay^age in later merges and the sigma sqrt(m2 / n) is unchanged by the
// scaling itself.
__global__ void __launch_bounds__(256) stats_decay_kernel(double *__restrict__ m2, int64_t n_m2,
                                                          double *__restrict__ n_prev, int n_dm, double decay) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_m2; i += stride) m2[i] *= decay;
    if (blockIdx.x == 0)
        for (int d = threadIdx.x; d < n_dm; d += blockDim.x) n_prev[d] *= decay;
}

extern "C" int dsx_stats_decay(double *m2, int64_t n_m2, double *n_prev, int n_dm, double decay,
                               dsx_stream_t stream) {
    DSX_REQUIRE(m2 && n_prev, "dsx_stats_decay: null pointer");
    DSX_REQUIRE(n_m2 > 0 && n_dm > 0, "dsx_stats_decay: sizes");
    DSX_REQUIRE(decay > 0.0 && decay <= 1.0, "dsx_stats_decay: decay=%g outside (0, 1]", decay);
    const int64_t blocks = std::min<int64_t>((n_m2 + 255) / 256, 4096);
    stats_decay_kernel<<<(unsigned)blocks, 256, 0, as_stream(stream)>>>(m2, n_m2, n_prev, n_dm, decay);
    return check_launch("dsx_stats_decay");
}
