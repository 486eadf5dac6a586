// Subband image history ("ring"): time-major rows per (subband, pixel).
//
// ring_new[s][p][0:hist)            <- ring_old[s][p][prev_len : prev_len + hist)
// ring_new[s][p][hist : hist + n)   <- img[s][0:n][p]   (64 x 64 LDS transpose)
// Old and new buffers alternate, so the history move never overlaps itself.
#include "dsx_internal.h"

namespace dsx {

template <typename T>
__global__ __launch_bounds__(256) void ring_hist_kernel(const T *old_ring, T *new_ring, int64_t n_rows,
                                                        int hist, int row_len, int prev_len) {
    const int64_t n = n_rows * hist;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = i / hist;
        const int k = (int)(i - row * hist);
        new_ring[row * row_len + k] = old_ring[row * row_len + prev_len + k];
    }
}

template <typename T>
__global__ __launch_bounds__(256) void ring_transpose_kernel(const T *img, T *ring, int n_sub,
                                                             int n_pix, int n_new, int hist,
                                                             int row_len) {
    __shared__ T tile[64][65];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;   // 64 x 4
    const int tiles_p = (n_pix + 63) / 64, tiles_t = (n_new + 63) / 64;
    const int64_t n_tiles = (int64_t)n_sub * tiles_p * tiles_t;
    for (int64_t b = blockIdx.x; b < n_tiles; b += gridDim.x) {
        const int s = (int)(b / ((int64_t)tiles_p * tiles_t));
        const int rem = (int)(b - (int64_t)s * tiles_p * tiles_t);
        const int tt = rem / tiles_p, tp = rem - tt * tiles_p;
        const int t0 = tt * 64, p0 = tp * 64;
        __syncthreads();
        for (int r = ty; r < 64; r += 4) {
            const int t = t0 + r, p = p0 + tx;
            if (t < n_new && p < n_pix) tile[r][tx] = img[((int64_t)s * n_new + t) * n_pix + p];
        }
        __syncthreads();
        for (int r = ty; r < 64; r += 4) {
            const int p = p0 + r, t = t0 + tx;
            if (t < n_new && p < n_pix)
                ring[((int64_t)s * n_pix + p) * row_len + hist + t] = tile[tx][r];
        }
    }
}

// fp16 history with 16-byte accesses: needs n_pix and n_new multiples of 64 and
// hist, row_len multiples of 8 (then every 8-pixel group of an image row and
// every 8-step group of a history row is 16-byte aligned).  A 64 x 64 tile is
// read as 8-pixel vectors, and each thread writes 8 consecutive steps of one
// pixel, gathered from 8 LDS rows that the wave reads contiguously.
__global__ __launch_bounds__(256) void ring_transpose16_kernel(const _Float16 *img, _Float16 *ring, int n_sub,
                                                               int n_pix, int n_new, int hist, int row_len) {
    __shared__ __align__(16) uint16_t tile[64][72];    // [step][pixel], 144-byte rows
    const int tid = threadIdx.x;
    const int tiles_p = n_pix / 64, tiles_t = n_new / 64;
    const int64_t n_tiles = (int64_t)n_sub * tiles_p * tiles_t;
    for (int64_t b = blockIdx.x; b < n_tiles; b += gridDim.x) {
        const int s = (int)(b / ((int64_t)tiles_p * tiles_t));
        const int rem = (int)(b - (int64_t)s * tiles_p * tiles_t);
        const int tt = rem / tiles_p, tp = rem - tt * tiles_p;
        const int t0 = tt * 64, p0 = tp * 64;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int i = tid + 256 * k;
            const int r = i >> 3, j = i & 7;
            const uint4 v = *reinterpret_cast<const uint4 *>(img + ((int64_t)s * n_new + t0 + r) * n_pix + p0 + 8 * j);
            *reinterpret_cast<uint4 *>(&tile[r][8 * j]) = v;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int i = tid + 256 * k;
            const int q = i & 63, c = i >> 6;
            uint32_t w[4];
#pragma unroll
            for (int m = 0; m < 4; ++m)
                w[m] = (uint32_t)tile[8 * c + 2 * m][q] | ((uint32_t)tile[8 * c + 2 * m + 1][q] << 16);
            // plain stores: each lane writes another pixel's row, and only the
            // cache merges those 16-byte pieces (non-temporal: 3.5 -> 15 ms per block)
            *reinterpret_cast<uint4 *>(ring + ((int64_t)s * n_pix + p0 + q) * row_len + hist + t0 + 8 * c) =
                make_uint4(w[0], w[1], w[2], w[3]);
        }
    }
}

// float32 history with 16-byte accesses (n_pix, n_new multiples of 64; hist, row_len
// multiples of 4).  A TT-step x 64-pixel tile (TT = 256, 128 or 64, the largest
// dividing the block) is read as 4-pixel vectors (16 lanes cover a 256-byte
// image-row segment) and written as 4-step vectors, TT / 4 lanes per pixel
// (TT x 4 contiguous bytes of its history row).  The tile's 4-float groups
// are XOR-swizzled by row group, so the column reads of the write phase spread
// over the LDS banks.  History writes here and in the carry kernels are
// non-temporal stores (each wave writes whole 256-byte row pieces): the history
// is streamed far past the caches, and the hint took the carry + transpose of a
// 96-antenna block from 4.70 to 4.18 ms.
template <int TT>
__global__ __launch_bounds__(256) void ring_transpose16f_kernel(const float *img, float *ring, int n_sub,
                                                                int n_pix, int n_new, int hist, int row_len) {
    __shared__ __align__(16) float tile[TT][64];        // [step][pixel group ^ (step / 4)]
    constexpr int PER = TT * 16 / 256;                   // 16-byte pieces per thread and phase
    const int tid = threadIdx.x;
    const int tiles_p = n_pix / 64, tiles_t = n_new / TT;
    const int64_t n_tiles = (int64_t)n_sub * tiles_p * tiles_t;
    for (int64_t b = blockIdx.x; b < n_tiles; b += gridDim.x) {
        const int s = (int)(b / ((int64_t)tiles_p * tiles_t));
        const int rem = (int)(b - (int64_t)s * tiles_p * tiles_t);
        const int tt = rem / tiles_p, tp = rem - tt * tiles_p;
        const int t0 = tt * TT, p0 = tp * 64;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int i = tid + 256 * k;
            const int r = i >> 4, j = i & 15;                     // step r, pixels 4j .. 4j + 3
            *reinterpret_cast<float4 *>(&tile[r][4 * (j ^ ((r >> 2) & 15))]) =
                *reinterpret_cast<const float4 *>(img + ((int64_t)s * n_new + t0 + r) * n_pix + p0 + 4 * j);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int i = tid + 256 * k;
            const int q = i / (TT / 4), c = i % (TT / 4);         // pixel q, steps 4c .. 4c + 3
            const int col = 4 * ((q >> 2) ^ (c & 15)) + (q & 3);
            __builtin_nontemporal_store(v4f{tile[4 * c][col], tile[4 * c + 1][col], tile[4 * c + 2][col], tile[4 * c + 3][col]},
                                        reinterpret_cast<v4f *>(ring + ((int64_t)s * n_pix + p0 + q) * row_len + hist + t0 + 4 * c));
        }
    }
}

// History carry in 16-byte pieces (hist, row_len, prev_len multiples of 16 bytes' worth of elements).
template <typename T>
__global__ __launch_bounds__(256) void ring_hist_vec_kernel(const T *old_ring, T *new_ring, int64_t n_rows,
                                                            int hist, int row_len, int prev_len) {
    constexpr int V = 16 / sizeof(T);
    const int hv = hist / V;
    const int64_t n = n_rows * hv;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = i / hv;
        const int k = (int)(i - row * hv);
        __builtin_nontemporal_store(*reinterpret_cast<const v4i *>(old_ring + row * row_len + prev_len + V * k),
                                    reinterpret_cast<v4i *>(new_ring + row * row_len + V * k));
    }
}

// History carry in 16-byte pieces (hist, row_len, prev_len multiples of 8).
__global__ __launch_bounds__(256) void ring_hist16_kernel(const _Float16 *old_ring, _Float16 *new_ring,
                                                          int64_t n_rows, int hist, int row_len, int prev_len) {
    const int hv = hist >> 3;
    const int64_t n = n_rows * hv;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = i / hv;
        const int k = (int)(i - row * hv);
        __builtin_nontemporal_store(*reinterpret_cast<const v4i *>(old_ring + row * row_len + prev_len + 8 * k),
                                    reinterpret_cast<v4i *>(new_ring + row * row_len + 8 * k));
    }
}

// out[s][t][k] = sum_{j < f} in[s][f t + j][k]: f-fold time downsampling of
// (n_planes, n_in, n_inner) planes, float32 accumulation, one thread per output
// element (coalesced along k).
template <typename T>
__global__ __launch_bounds__(256) void downsample_time_kernel(const T *in, T *out, int64_t n_out_total,
                                                              int n_in, int64_t n_inner, int f) {
    const int n_out = n_in / f;
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n_out_total;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = g % n_inner;
        const int64_t rest = g / n_inner;
        const int t = (int)(rest % n_out);
        const int64_t s = rest / n_out;
        const T *src = in + (s * n_in + (int64_t)f * t) * n_inner + k;
        float acc = 0.f;
        for (int j = 0; j < f; ++j) acc += ld_as_float(src, (int64_t)j * n_inner);
        st_from_float(out, g, acc);
    }
}

// The same for float32 with n_inner a multiple of 4: one 16-byte access per
// 4 consecutive elements, summed in the same order.
__global__ __launch_bounds__(256) void downsample_time_f4_kernel(const float *in, float *out, int64_t n_out4,
                                                                 int n_in, int64_t n_inner4, int f) {
    const int n_out = n_in / f;
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n_out4;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = g % n_inner4;
        const int64_t rest = g / n_inner4;
        const int t = (int)(rest % n_out);
        const int64_t s = rest / n_out;
        const float4 *src = reinterpret_cast<const float4 *>(in) + (s * n_in + (int64_t)f * t) * n_inner4 + k;
        float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
        for (int j = 0; j < f; ++j) {
            const float4 x = src[(int64_t)j * n_inner4];
            acc.x += x.x;
            acc.y += x.y;
            acc.z += x.z;
            acc.w += x.w;
        }
        reinterpret_cast<float4 *>(out)[g] = acc;
    }
}

// Static-sky subtraction on a block of images img[s][t][p] (the image-domain
// form of StaticSkyMean, src/dsart/services/corr_fast_integration.py:1049-1246:
// imaging is linear, so subtracting a per-cell mean grid from every step
// subtracts its image).  Per (subband, pixel): the block mean of the images
// (float64 sum); when `subtract`, every step loses the mean of the previous
// blocks' means, wsum / count; then this block's mean enters the window
// (ring[head], evicting the oldest when the window is full).
template <typename T>
__global__ __launch_bounds__(256) void static_sky_kernel(T *img, int n_sub, int n_steps, int64_t n_pix,
                                                         float *ring, double *wsum, int window, int head,
                                                         int count, int subtract) {
    const int64_t n = (int64_t)n_sub * n_pix;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = i / n_pix, p = i - s * n_pix;
        T *col = img + s * (int64_t)n_steps * n_pix + p;
        // one pass: the block sum of the images as they arrive, and the earlier
        // blocks' mean subtracted from each
        double acc = 0.0;
        if (subtract && count > 0) {
            const float m = (float)(wsum[i] / (double)count);
            for (int t = 0; t < n_steps; ++t) {
                const int64_t o = (int64_t)t * n_pix;
                const float x = ld_as_float(col, o);
                acc += (double)x;
                st_from_float(col, o, x - m);
            }
        } else {
            for (int t = 0; t < n_steps; ++t) acc += (double)ld_as_float(col, (int64_t)t * n_pix);
        }
        const float mean_now = (float)(acc / (double)n_steps);
        float *slot = ring + (int64_t)head * n + i;
        double w = wsum[i];
        if (count == window) w -= (double)*slot;
        *slot = mean_now;
        wsum[i] = w + (double)mean_now;
    }
}

}  // namespace dsx

using namespace dsx;

extern "C" int dsx_ring_advance(const void *ring_old, void *ring_new, const void *img, int n_sub,
                                int n_pix, int hist, int cap, int prev_len, int n_new, int dtype,
                                dsx_stream_t stream) {
    DSX_REQUIRE(ring_old && ring_new && img, "dsx_ring_advance: null pointer");
    DSX_REQUIRE(ring_old != ring_new, "dsx_ring_advance: old and new ring must differ");
    DSX_REQUIRE(n_sub > 0 && n_pix > 0 && hist >= 0 && cap > 0, "dsx_ring_advance: bad shape");
    DSX_REQUIRE(prev_len >= 0 && prev_len <= cap, "dsx_ring_advance: prev_len out of range");
    DSX_REQUIRE(n_new > 0 && n_new <= cap, "dsx_ring_advance: n_new out of range");
    DSX_REQUIRE(dtype == DSX_F16 || dtype == DSX_F32, "dsx_ring_advance: dtype");
    const int row_len = hist + cap;
    const int64_t n_rows = (int64_t)n_sub * n_pix;
    hipStream_t s = as_stream(stream);
    const int64_t nh = n_rows * hist;
    const int hb = (int)std::min<int64_t>((nh + 255) / 256, 65536);
    const int64_t n_tiles = (int64_t)n_sub * ((n_pix + 63) / 64) * ((n_new + 63) / 64);
    const int tb = (int)std::min<int64_t>(n_tiles, 65536);
    const bool vec16 = dtype == DSX_F16 && n_pix % 64 == 0 && n_new % 64 == 0 && hist % 8 == 0 &&
                       row_len % 8 == 0 && prev_len % 8 == 0 && ((uintptr_t)img & 15) == 0 &&
                       ((uintptr_t)ring_old & 15) == 0 && ((uintptr_t)ring_new & 15) == 0;
    const bool vec16f = dtype == DSX_F32 && n_pix % 64 == 0 && n_new % 64 == 0 && hist % 4 == 0 &&
                        row_len % 4 == 0 && prev_len % 4 == 0 && ((uintptr_t)img & 15) == 0 &&
                        ((uintptr_t)ring_old & 15) == 0 && ((uintptr_t)ring_new & 15) == 0;
    if (vec16f) {
        const int64_t nv = n_rows * (hist / 4);
        if (nv > 0)
            ring_hist_vec_kernel<float><<<(int)std::min<int64_t>((nv + 255) / 256, 65536), 256, 0, s>>>(
                (const float *)ring_old, (float *)ring_new, n_rows, hist, row_len, prev_len);
        if (n_new % 256 == 0) {
            const int tb3 = (int)std::min<int64_t>((int64_t)n_sub * (n_pix / 64) * (n_new / 256), 65536);
            ring_transpose16f_kernel<256><<<tb3, 256, 0, s>>>((const float *)img, (float *)ring_new, n_sub, n_pix,
                                                              n_new, hist, row_len);
        } else if (n_new % 128 == 0) {
            const int tb2 = (int)std::min<int64_t>((int64_t)n_sub * (n_pix / 64) * (n_new / 128), 65536);
            ring_transpose16f_kernel<128><<<tb2, 256, 0, s>>>((const float *)img, (float *)ring_new, n_sub, n_pix,
                                                              n_new, hist, row_len);
        } else {
            ring_transpose16f_kernel<64><<<tb, 256, 0, s>>>((const float *)img, (float *)ring_new, n_sub, n_pix,
                                                            n_new, hist, row_len);
        }
    } else if (vec16) {
        const int64_t nv = n_rows * (hist / 8);
        if (nv > 0)
            ring_hist16_kernel<<<(int)std::min<int64_t>((nv + 255) / 256, 65536), 256, 0, s>>>(
                (const _Float16 *)ring_old, (_Float16 *)ring_new, n_rows, hist, row_len, prev_len);
        ring_transpose16_kernel<<<tb, 256, 0, s>>>((const _Float16 *)img, (_Float16 *)ring_new, n_sub, n_pix,
                                                   n_new, hist, row_len);
    } else if (dtype == DSX_F16) {
        if (nh > 0)
            ring_hist_kernel<_Float16><<<hb, 256, 0, s>>>((const _Float16 *)ring_old, (_Float16 *)ring_new,
                                                           n_rows, hist, row_len, prev_len);
        ring_transpose_kernel<_Float16><<<tb, 256, 0, s>>>((const _Float16 *)img, (_Float16 *)ring_new,
                                                            n_sub, n_pix, n_new, hist, row_len);
    } else {
        if (nh > 0)
            ring_hist_kernel<float><<<hb, 256, 0, s>>>((const float *)ring_old, (float *)ring_new,
                                                        n_rows, hist, row_len, prev_len);
        ring_transpose_kernel<float><<<tb, 256, 0, s>>>((const float *)img, (float *)ring_new, n_sub,
                                                         n_pix, n_new, hist, row_len);
    }
    return check_launch("dsx_ring_advance");
}

extern "C" int dsx_downsample_time(const void *in, void *out, int n_planes, int n_in, int64_t n_inner,
                                   int factor, int dtype, dsx_stream_t stream) {
    DSX_REQUIRE(in && out, "dsx_downsample_time: null pointer");
    DSX_REQUIRE(n_planes > 0 && n_in > 0 && n_inner > 0, "dsx_downsample_time: bad shape");
    DSX_REQUIRE(factor >= 1 && n_in % factor == 0, "dsx_downsample_time: factor %d must divide n_in %d",
                factor, n_in);
    DSX_REQUIRE(in != out, "dsx_downsample_time: in and out must differ");
    DSX_REQUIRE(dtype == DSX_F16 || dtype == DSX_F32, "dsx_downsample_time: dtype");
    const int64_t total = (int64_t)n_planes * (n_in / factor) * n_inner;
    const int blocks = (int)std::min<int64_t>((total + 255) / 256, 1 << 20);
    hipStream_t s = as_stream(stream);
    if (dtype == DSX_F32 && n_inner % 4 == 0 && ((uintptr_t)in & 15) == 0 && ((uintptr_t)out & 15) == 0) {
        const int64_t total4 = total / 4;
        const int blocks4 = (int)std::min<int64_t>((total4 + 255) / 256, 1 << 20);
        downsample_time_f4_kernel<<<blocks4, 256, 0, s>>>((const float *)in, (float *)out, total4, n_in,
                                                         n_inner / 4, factor);
    } else if (dtype == DSX_F16)
        downsample_time_kernel<_Float16><<<blocks, 256, 0, s>>>((const _Float16 *)in, (_Float16 *)out, total,
                                                                 n_in, n_inner, factor);
    else
        downsample_time_kernel<float><<<blocks, 256, 0, s>>>((const float *)in, (float *)out, total, n_in,
                                                              n_inner, factor);
    return check_launch("dsx_downsample_time");
}

extern "C" int dsx_static_sky(void *img, int img_dtype, int n_sub, int n_steps, int64_t n_pix, float *ring,
                              double *wsum, int window, int head, int count, int subtract,
                              dsx_stream_t stream) {
    DSX_REQUIRE(img && ring && wsum, "dsx_static_sky: null pointer");
    DSX_REQUIRE(n_sub > 0 && n_steps > 0 && n_pix > 0, "dsx_static_sky: bad shape");
    DSX_REQUIRE(window >= 1 && head >= 0 && head < window && count >= 0 && count <= window,
                "dsx_static_sky: window %d, head %d, count %d", window, head, count);
    DSX_REQUIRE(img_dtype == DSX_F16 || img_dtype == DSX_F32, "dsx_static_sky: img_dtype");
    const int64_t n = (int64_t)n_sub * n_pix;
    const int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
    hipStream_t st = as_stream(stream);
    if (img_dtype == DSX_F16)
        static_sky_kernel<_Float16><<<blocks, 256, 0, st>>>((_Float16 *)img, n_sub, n_steps, n_pix, ring, wsum,
                                                             window, head, count, subtract);
    else
        static_sky_kernel<float><<<blocks, 256, 0, st>>>((float *)img, n_sub, n_steps, n_pix, ring, wsum,
                                                          window, head, count, subtract);
    return check_launch("dsx_static_sky");
}
