// Detector over a materialised search cube [T][F][H][W] (the reference's
// Detector.forward input): centred time boxcars of several widths, the
// per-cell maximum S/N over widths, and the n_top strongest cells.
//
//   1. cube_cumsum_kernel: per cell column, the exclusive time prefix
//        cs[j][c] = sum_{t < j - H} x[t][c],  j in [0, T + 2H] (zeros before
//      the cube, the total after it), float32 -- H = W_max / 2;
//   2. cube_queue_kernel: for centres t in [t_base, t_base + n_out) and every
//      cell, B_k = cs[t - w_k/2 + w_k + H] - cs[t - w_k/2 + H] (the box
//      [t - w_k/2, t - w_k/2 + w_k), zeros outside the cube),
//      snr_k = B_k * sigma_inv_k, the maximum over k (strict >, ties -> lowest
//      k) appended to a queue when it beats the threshold;
//   3. top_records_kernel (one work-group): the n_top queue records with the
//      largest (snr, -flat index) by an 8-pass radix select on that 64-bit
//      key, then a bitonic sort in LDS, written in descending order;
//   4. cube_samples_kernel: raw box sums of every width at given cells (the
//      Layer-2 noise estimator's sample).
// The queue and its counter are scratch: dsx_cube_candidates takes them from
// a workspace argument.
#include <algorithm>

#include "dsx_internal.h"

namespace dsx {

struct CubeRec {
    int32_t t;        // centre relative to t_base
    int32_t cell;     // flat (f, row, col) cell index
    int32_t k;        // winning width index
    float snr;
};
static_assert(sizeof(CubeRec) == 16, "cube record must be 16 bytes");

constexpr int CD_MAX_K = 16;
constexpr int CD_TOP_THREADS = 1024;
constexpr int CD_MAX_TOP = 4096;

template <typename T>
__global__ __launch_bounds__(256) void cube_cumsum_kernel(const T *x, float *cs, int n_t, int64_t n_cell, int h) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_cell) return;
    for (int j = 0; j <= h; ++j) cs[(int64_t)j * n_cell + c] = 0.f;
    float run = 0.f;
    for (int t = 0; t < n_t; ++t) {
        run += (float)x[(int64_t)t * n_cell + c];
        cs[(int64_t)(t + h + 1) * n_cell + c] = run;
    }
    for (int j = n_t + h + 1; j <= n_t + 2 * h; ++j) cs[(int64_t)j * n_cell + c] = run;
}

__global__ void zero_kernel(int32_t *p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

struct QueueParams {
    const float *cs;
    int64_t n_cell;
    int h, t_base, n_out, n_k;
    int w[CD_MAX_K];
    float inv[CD_MAX_K];
    float threshold;
    CubeRec *queue;
    int32_t *count;
    int cap;
};

__global__ __launch_bounds__(256) void cube_queue_kernel(QueueParams P) {
    const int64_t n = (int64_t)P.n_out * P.n_cell;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const int t = (int)(e / P.n_cell);
        const int64_t c = e - (int64_t)t * P.n_cell;
        const int tc = P.t_base + t;
        float best = -INFINITY;
        int bk = 0;
        for (int k = 0; k < P.n_k; ++k) {
            const int lo = tc - (P.w[k] >> 1) + P.h;
            const float b = P.cs[(int64_t)(lo + P.w[k]) * P.n_cell + c] - P.cs[(int64_t)lo * P.n_cell + c];
            const float snr = b * P.inv[k];
            if (snr > best) {
                best = snr;
                bk = k;
            }
        }
        if (best > P.threshold) {
            const int slot = atomicAdd(P.count, 1);
            if (slot < P.cap) {
                CubeRec r;
                r.t = t;
                r.cell = (int32_t)c;
                r.k = bk;
                r.snr = best;
                P.queue[slot] = r;
            }
        }
    }
}

// 64-bit order key: larger snr first, then the smaller flat index (t * n_cell + cell)
__device__ __forceinline__ uint64_t rec_key(const CubeRec &r, int64_t n_cell) {
    uint32_t u = __float_as_uint(r.snr);
    u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);               // monotonic in snr
    const uint32_t idx = (uint32_t)((int64_t)r.t * n_cell + r.cell);
    return ((uint64_t)u << 32) | (uint64_t)(~idx);
}

__global__ __launch_bounds__(CD_TOP_THREADS) void top_records_kernel(const CubeRec *queue, const int32_t *count,
                                                                     int cap, int64_t n_cell, int n_top,
                                                                     CubeRec *out, int32_t *n_out,
                                                                     int32_t *n_above) {
    __shared__ uint32_t hist[256];
    __shared__ uint64_t s_prefix;
    __shared__ int s_need;
    __shared__ uint64_t keys[CD_MAX_TOP];
    __shared__ int src[CD_MAX_TOP];
    __shared__ int s_fill;
    const int tid = threadIdx.x;
    const int n = min(*count, cap);
    const int m = min(n, n_top);
    if (tid == 0) {
        *n_above = *count;
        s_prefix = 0;
        s_need = m;                      // records still to take at or below the current prefix
        s_fill = 0;
    }
    __syncthreads();
    if (m == 0) {
        if (tid == 0) *n_out = 0;
        return;
    }
    // radix select: the m-th largest key K*, eight 8-bit digits from the top
    uint64_t mask = 0;
    for (int pass = 0; pass < 8; ++pass) {
        const int shift = 56 - 8 * pass;
        for (int i = tid; i < 256; i += CD_TOP_THREADS) hist[i] = 0;
        __syncthreads();
        const uint64_t prefix = s_prefix;
        for (int i = tid; i < n; i += CD_TOP_THREADS) {
            const uint64_t key = rec_key(queue[i], n_cell);
            if ((key & mask) == prefix) atomicAdd(&hist[(key >> shift) & 0xFF], 1u);
        }
        __syncthreads();
        if (tid == 0) {
            int need = s_need;
            int d = 255;
            for (; d > 0; --d) {
                if ((int)hist[d] >= need) break;
                need -= (int)hist[d];
            }
            s_need = need;
            s_prefix = prefix | ((uint64_t)d << shift);
        }
        __syncthreads();
        mask |= (uint64_t)0xFF << shift;
    }
    // keys are distinct (flat indices are), so exactly m records have key >= K*
    const uint64_t kstar = s_prefix;
    for (int i = tid; i < n; i += CD_TOP_THREADS) {
        const uint64_t key = rec_key(queue[i], n_cell);
        if (key >= kstar) {
            const int slot = atomicAdd(&s_fill, 1);
            if (slot < CD_MAX_TOP) {
                keys[slot] = key;
                src[slot] = i;
            }
        }
    }
    __syncthreads();
    // bitonic sort of the m keys, descending (padded to a power of two with zeros)
    int np = 1;
    while (np < m) np <<= 1;
    for (int i = m + tid; i < np; i += CD_TOP_THREADS) {
        keys[i] = 0;
        src[i] = -1;
    }
    __syncthreads();
    for (int size = 2; size <= np; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = tid; i < np; i += CD_TOP_THREADS) {
                const int j = i ^ stride;
                if (j > i) {
                    const bool desc = (i & size) == 0;
                    const uint64_t a = keys[i], b = keys[j];
                    if (desc ? (a < b) : (a > b)) {
                        keys[i] = b;
                        keys[j] = a;
                        const int t = src[i];
                        src[i] = src[j];
                        src[j] = t;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int i = tid; i < m; i += CD_TOP_THREADS) out[i] = queue[src[i]];
    if (tid == 0) *n_out = m;
}

__global__ __launch_bounds__(256) void cube_samples_kernel(const float *cs, int64_t n_cell, int h, int t_base,
                                                           int n_out, const int64_t *cells, int n_s, QueueParams P,
                                                           float *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_s) return;
    const int64_t e = cells[i];
    if (e < 0 || e >= (int64_t)n_out * n_cell) {          // not a cell of the centres: no sample
        for (int k = 0; k < P.n_k; ++k) out[(int64_t)k * n_s + i] = NAN;
        return;
    }
    const int t = (int)(e / n_cell);
    const int64_t c = e - (int64_t)t * n_cell;
    const int tc = t_base + t;
    for (int k = 0; k < P.n_k; ++k) {
        const int lo = tc - (P.w[k] >> 1) + h;
        out[(int64_t)k * n_s + i] = cs[(int64_t)(lo + P.w[k]) * n_cell + c] - cs[(int64_t)lo * n_cell + c];
    }
}

static int fill_widths(QueueParams &P, const int32_t *widths, const float *sigma_inv, int n_k, int h) {
    if (n_k < 1 || n_k > CD_MAX_K) return set_error(DSX_EINVAL, "n_k=%d not in [1, %d]", n_k, CD_MAX_K);
    P.n_k = n_k;
    for (int k = 0; k < n_k; ++k) {
        if (widths[k] < 1 || widths[k] / 2 > h || widths[k] - widths[k] / 2 > h + 1)
            return set_error(DSX_EINVAL, "width %d outside the cumsum padding (w_max/2 = %d)", widths[k], h);
        P.w[k] = widths[k];
        P.inv[k] = sigma_inv ? sigma_inv[k] : 1.f;
    }
    return DSX_OK;
}

}  // namespace dsx

using namespace dsx;

extern "C" int dsx_cube_cumsum(const void *cube, int cube_dtype, float *cs, int n_t, int64_t n_cell, int h,
                               dsx_stream_t stream) {
    DSX_REQUIRE(cube && cs, "dsx_cube_cumsum: null pointer");
    DSX_REQUIRE(n_t > 0 && n_cell > 0 && h >= 0, "dsx_cube_cumsum: bad shape");
    DSX_REQUIRE(cube_dtype == DSX_F16 || cube_dtype == DSX_F32, "dsx_cube_cumsum: cube_dtype");
    const unsigned blocks = (unsigned)((n_cell + 255) / 256);
    hipStream_t s = as_stream(stream);
    if (cube_dtype == DSX_F16)
        cube_cumsum_kernel<_Float16><<<blocks, 256, 0, s>>>((const _Float16 *)cube, cs, n_t, n_cell, h);
    else
        cube_cumsum_kernel<float><<<blocks, 256, 0, s>>>((const float *)cube, cs, n_t, n_cell, h);
    return check_launch("dsx_cube_cumsum");
}

namespace dsx {

// Scratch layout (workspace argument or the static array): the record counter
// in the first 256 bytes, then the queue of 16-byte records.
constexpr int64_t CD_SCRATCH_HEAD = 256;

// queue + top selection with the queue and counter in device scratch
static int cube_candidates_impl(const float *cs, int n_t, int64_t n_cell, int h, const int32_t *host_widths,
                                const float *host_sigma_inv, int n_k, float threshold, int t_base, int n_out,
                                int n_top, void *out, int32_t *n_out_dev, int32_t *n_above, char *scratch,
                                int64_t scratch_bytes, hipStream_t s, const char *who) {
    if (n_out <= 0 || t_base < 0 || t_base + n_out > n_t)
        return set_error(DSX_EINVAL, "%s: centres [t_base, t_base + n_out) must lie in [0, n_t)", who);
    if ((int64_t)n_out * n_cell >= (int64_t)UINT32_MAX)
        return set_error(DSX_ERANGE, "%s: n_out * n_cell exceeds 2^32 - 1 flat indices", who);
    if (n_top < 1 || n_top > CD_MAX_TOP) return set_error(DSX_EINVAL, "%s: n_top=%d not in [1, %d]", who, n_top, CD_MAX_TOP);
    const int64_t cap = (scratch_bytes - CD_SCRATCH_HEAD) / (int64_t)sizeof(CubeRec);
    if (cap < 1 || cap > INT32_MAX) return set_error(DSX_EINVAL, "%s: the scratch holds no record", who);
    int32_t *count = reinterpret_cast<int32_t *>(scratch);
    CubeRec *queue = reinterpret_cast<CubeRec *>(scratch + CD_SCRATCH_HEAD);
    QueueParams P{};
    P.cs = cs;
    P.n_cell = n_cell;
    P.h = h;
    P.t_base = t_base;
    P.n_out = n_out;
    P.threshold = threshold;
    P.queue = queue;
    P.count = count;
    P.cap = (int)cap;
    const int rc = fill_widths(P, host_widths, host_sigma_inv, n_k, h);
    if (rc != DSX_OK) return rc;
    // the counter is cleared by a kernel (no memset node if the caller captures the stream)
    zero_kernel<<<1, 64, 0, s>>>(count, 1);
    const int64_t n = (int64_t)n_out * n_cell;
    const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 1 << 16);
    cube_queue_kernel<<<blocks, 256, 0, s>>>(P);
    int e = check_launch(who);
    if (e != DSX_OK) return e;
    top_records_kernel<<<1, CD_TOP_THREADS, 0, s>>>(queue, count, (int)cap, n_cell, n_top,
                                                    reinterpret_cast<CubeRec *>(out), n_out_dev, n_above);
    return check_launch(who);
}

}  // namespace dsx

extern "C" int dsx_cube_candidates(const float *cs, int n_t, int64_t n_cell, int h, const int32_t *host_widths,
                                   const float *host_sigma_inv, int n_k, float threshold, int t_base, int n_out,
                                   int n_top, void *out, int32_t *n_out_dev, int32_t *n_above,
                                   void *workspace, int64_t workspace_bytes, dsx_stream_t stream) {
    DSX_REQUIRE(cs && host_widths && out && n_out_dev && n_above && workspace,
                "dsx_cube_candidates: null pointer");
    DSX_REQUIRE(n_t > 0 && n_cell > 0 && h >= 0, "dsx_cube_candidates: bad shape");
    return cube_candidates_impl(cs, n_t, n_cell, h, host_widths, host_sigma_inv, n_k, threshold, t_base, n_out,
                                n_top, out, n_out_dev, n_above, reinterpret_cast<char *>(workspace),
                                workspace_bytes, as_stream(stream), "dsx_cube_candidates");
}
extern "C" int dsx_cube_box_samples(const float *cs, int n_t, int64_t n_cell, int h, const int32_t *host_widths,
                                    int n_k, int t_base, int n_out, const int64_t *cells, int n_s, float *out,
                                    dsx_stream_t stream) {
    DSX_REQUIRE(cs && host_widths && cells && out, "dsx_cube_box_samples: null pointer");
    DSX_REQUIRE(n_t > 0 && n_cell > 0 && h >= 0 && n_s >= 0 && t_base >= 0 && n_out > 0 && t_base + n_out <= n_t,
                "dsx_cube_box_samples: bad shape");
    if (n_s == 0) return DSX_OK;
    QueueParams P{};
    const int rc = fill_widths(P, host_widths, nullptr, n_k, h);
    if (rc != DSX_OK) return rc;
    cube_samples_kernel<<<(n_s + 255) / 256, 256, 0, as_stream(stream)>>>(cs, n_cell, h, t_base, n_out, cells,
                                                                               n_s, P, out);
    return check_launch("dsx_cube_box_samples");
}

// ---------------------------------------------------------------------------
// No-allocation variant: the counter and the queue live in a file-scope device
// array (each GPU holds its own copy until the process ends).  Host code uses
// the per-device address from hipGetSymbolAddress, never the array's name.
// ---------------------------------------------------------------------------
#include <mutex>

namespace dsx {

constexpr int64_t CD_STATIC_RECORDS = 1 << 20;                    // 16 MiB of queue per GPU
constexpr int64_t CD_STATIC_BYTES = CD_SCRATCH_HEAD + CD_STATIC_RECORDS * (int64_t)sizeof(CubeRec);
constexpr int CD_MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[CD_STATIC_BYTES];

static std::mutex g_scratch_mu;
static void *g_scratch_ptr[CD_MAX_DEVICES];

// Address of this device's scratch, looked up once per device with the device current.
static int device_scratch(char **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return set_error(DSX_ELAUNCH, "dsx_cube_candidates_noalloc: hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= CD_MAX_DEVICES)
        return set_error(DSX_ERANGE, "dsx_cube_candidates_noalloc: device %d >= %d", dev, CD_MAX_DEVICES);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            return set_error(DSX_ELAUNCH, "dsx_cube_candidates_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = static_cast<char *>(g_scratch_ptr[dev]);
    return DSX_OK;
}

}  // namespace dsx

extern "C" int dsx_cube_candidates_noalloc(const float *cs, int n_t, int64_t n_cell, int h,
                                           const int32_t *host_widths, const float *host_sigma_inv, int n_k,
                                           float threshold, int t_base, int n_out, int n_top, void *out,
                                           int32_t *n_out_dev, int32_t *n_above, dsx_stream_t stream) {
    DSX_REQUIRE(cs && host_widths && out && n_out_dev && n_above, "dsx_cube_candidates_noalloc: null pointer");
    DSX_REQUIRE(n_t > 0 && n_cell > 0 && h >= 0, "dsx_cube_candidates_noalloc: bad shape");
    DSX_REQUIRE(n_top >= 1 && n_top <= CD_MAX_TOP, "dsx_cube_candidates_noalloc: n_top=%d beyond the maximum %d",
                n_top, CD_MAX_TOP);
    char *scr = nullptr;
    const int rc = device_scratch(&scr);
    if (rc != DSX_OK) return rc;
    return cube_candidates_impl(cs, n_t, n_cell, h, host_widths, host_sigma_inv, n_k, threshold, t_base, n_out,
                                n_top, out, n_out_dev, n_above, scr, CD_STATIC_BYTES, as_stream(stream),
                                "dsx_cube_candidates_noalloc");
}

extern "C" int64_t dsx_cube_candidates_noalloc_capacity(void) { return CD_STATIC_RECORDS; }
