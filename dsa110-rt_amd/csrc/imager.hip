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
constexpr int IMG_MAX_KB = 32;          // <= 64 rows per chunk (2 rows = 4 K values per MFMA)
constexpr int TAP_CONJ = 1 << 20;
constexpr int TAP_NEG = 1 << 21;
constexpr int TAP_CELL = (1 << 20) - 1;

struct ImgParams {
    const int32_t *vis;
    const int32_t *tap_src;
    const int32_t *tap_dst;
    const int32_t *tap_off;   // n_sub + 1 CSR offsets (device)
    const float2 *grid;       // dense mode: (n_items, N, N) complex64
    void *img;
    int n_steps, n_sub, per_sub, nbase, N, log2N, n_rows, rc_max, npad;
    float scale;
    int dense;                // 0: taps from int32 visibilities (+ mask); 1: dense grid, no mask
};

template <typename TOut>
__global__ __launch_bounds__(IMG_THREADS) void image_kernel(ImgParams P) {
    extern __shared__ __align__(16) float2 smem[];
    const int N = P.N;
    float2 *tw = smem;                       // N entries: e^{+2 pi i m / N}
    float2 *H = smem + N;                    // rc_max x N
    int2 *Hi = reinterpret_cast<int2 *>(H);
    const int tid = threadIdx.x;
    for (int m = tid; m < N; m += IMG_THREADS) {
        double sv, cv;
        sincospi(2.0 * (double)m / (double)N, &sv, &cv);
        tw[m] = make_float2((float)cv, (float)sv);
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

    for (int item = blockIdx.x; item < n_items; item += gridDim.x) {
        const int s = item / P.n_steps;
        const int t = item - s * P.n_steps;
        const int32_t *vslab = P.dense ? nullptr
                                       : P.vis + 2 * ((int64_t)t * n_cout + (int64_t)s * P.per_sub) * P.nbase;
        const int tap_lo = P.dense ? 0 : P.tap_off[s], tap_hi = P.dense ? 0 : P.tap_off[s + 1];
        const float2 *gsrc = P.dense ? P.grid + (int64_t)item * N * N : nullptr;
        TOut *out = img + ((int64_t)s * P.n_steps + t) * N * N;
        for (int ck = 0; ck < n_chunks; ++ck) {
            const int r0 = ck * P.rc_max;
            const int rc = min(P.rc_max, P.n_rows - r0);
            __syncthreads();
            if (P.dense) {
                // Hermitian fold of (-1)^(a+b) G[a][b]: H[r][w] = (-1)^(r+w) (G[r][w] + conj G[-r][-w]),
                // written straight into float32 rows in bit-reversed order
                for (int i = tid; i < rc * N; i += IMG_THREADS) {
                    const int r = i >> P.log2N, w = i & (N - 1);
                    const int rg = r0 + r;
                    const float2 a = gsrc[(int64_t)rg * N + w];
                    const float2 b = gsrc[(int64_t)((N - rg) & (N - 1)) * N + ((N - w) & (N - 1))];
                    const float sg = ((rg + w) & 1) ? -1.f : 1.f;
                    const int wr = (int)(__brev((unsigned)w) >> (32 - P.log2N));
                    H[r * N + wr] = make_float2(sg * (a.x + b.x), sg * (a.y - b.y));
                }
                __syncthreads();
            } else {
            for (int i = tid; i < rc * N; i += IMG_THREADS) Hi[i] = make_int2(0, 0);
            __syncthreads();
            // ---- 1. scatter taps (int32 LDS atomics)
            for (int i = tap_lo + tid; i < tap_hi; i += IMG_THREADS) {
                const int d = P.tap_dst[i];
                const int cell = d & TAP_CELL;
                const int r = (cell >> P.log2N) - r0;
                if (r < 0 || r >= rc) continue;
                const int w = cell & (N - 1);
                int2 v = *reinterpret_cast<const int2 *>(vslab + 2 * (int64_t)P.tap_src[i]);
                if (d & TAP_CONJ) v.y = -v.y;
                if (d & TAP_NEG) { v.x = -v.x; v.y = -v.y; }
                int *c = reinterpret_cast<int *>(&Hi[r * N + w]);
                atomicAdd(c, v.x);
                atomicAdd(c + 1, v.y);
            }
            __syncthreads();
            // ---- int32 -> float32 in place, with the bit-reversal permutation
            for (int i = tid; i < rc * N; i += IMG_THREADS) {
                const int r = i >> P.log2N, j = i & (N - 1);
                const int jr = (int)(__brev((unsigned)j) >> (32 - P.log2N));
                if (j < jr) {
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
            // ---- 2. radix-2 DIT row transforms
            const int halfN = N >> 1;
            for (int m = 1; m < N; m <<= 1) {
                const int stride = N / (2 * m);
                for (int b = tid; b < rc * halfN; b += IMG_THREADS) {
                    const int r = b / halfN;
                    const int k = b - r * halfN;
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
            // ---- 3. column C2R on MFMA: D[i][j] = sum_k A[i][k] B[k][j],
            //         k = 2 r + part: A = (cos, -sin)(2 pi r i / N), B = c_r (Re, Im) Y[r][j]
            const int nkb = (2 * rc + 3) / 4;
            for (int it = wave; it < n_tiles_1d; it += IMG_THREADS / 64) {
                const int i0 = it * 16;
                float a_op[IMG_MAX_KB];
                {
                    const int i = i0 + lr;
#pragma unroll
                    for (int kb = 0; kb < IMG_MAX_KB; ++kb) {
                        const int k = kb * 4 + lg;
                        const int r = k >> 1;
                        float a = 0.f;
                        if (kb < nkb && r < rc) {
                            const float2 w = tw[((r0 + r) * i) & (N - 1)];
                            a = (k & 1) ? -w.y : w.x;
                        }
                        a_op[kb] = a;
                    }
                }
                for (int jt = 0; jt < n_tiles_1d; ++jt) {
                    const int j0 = jt * 16;
                    v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int kb = 0; kb < IMG_MAX_KB; ++kb) {
                        if (kb < nkb) {
                            const int k = kb * 4 + lg;
                            const int r = k >> 1;
                            float bv = 0.f;
                            if (r < rc) {
                                const int rg = r0 + r;
                                const float cr = (rg == 0 || rg == halfN) ? 1.f : 2.f;
                                const float2 y = H[r * N + j0 + lr];
                                bv = cr * ((k & 1) ? y.y : y.x);
                            }
                            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a_op[kb], bv, acc, 0, 0, 0);
                        }
                    }
                    // ---- 4. epilogue
                    const int j = j0 + lr;
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int i = i0 + 4 * lg + v;
                        const bool inside = P.dense || (i >= P.npad && i < N - P.npad && j >= P.npad &&
                                                        j < N - P.npad && !(i == halfN && j == halfN));
                        float val = inside ? acc[v] * out_scale : 0.f;
                        const int64_t o = (int64_t)i * N + j;
                        if (ck > 0) val += ld_as_float(out, o);
                        st_from_float(out, o, val);
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

}  // namespace dsx

using namespace dsx;

extern "C" int dsx_image_subbands(const int32_t *vis, const int32_t *tap_src,
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
    const size_t budget = 96 * 1024;
    int rc_max = (int)((budget - (size_t)n_grid * sizeof(float2)) / ((size_t)n_grid * sizeof(float2)));
    rc_max = std::min(rc_max, 2 * IMG_MAX_KB);
    rc_max = std::min(rc_max, n_rows);
    if (rc_max < 1) return set_error(DSX_ERANGE, "dsx_image_subbands: n_grid too large");
    const int n_chunks = (n_rows + rc_max - 1) / rc_max;
    if (n_chunks > 1 && img_dtype != DSX_F32)
        return set_error(DSX_ENOTSUP,
                         "dsx_image_subbands: %d uv rows need %d passes; multi-pass imaging "
                         "accumulates in place and needs img_dtype DSX_F32", n_rows, n_chunks);
    hipStream_t s = as_stream(stream);
    ImgParams P{vis, tap_src, tap_dst, tap_off, nullptr, img, n_steps, n_sub, per_sub, nbase, n_grid,
                log2N, n_rows, rc_max, npad, scale, 0};
    const size_t lds = (size_t)n_grid * sizeof(float2) * (1 + rc_max);
    const int grid = std::min(n_sub * n_steps, 4096);
    if (img_dtype == DSX_F16) {
        (void)hipFuncSetAttribute((const void *)image_kernel<_Float16>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        image_kernel<_Float16><<<grid, IMG_THREADS, lds, s>>>(P);
    } else {
        (void)hipFuncSetAttribute((const void *)image_kernel<float>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        image_kernel<float><<<grid, IMG_THREADS, lds, s>>>(P);
    }
    return check_launch("dsx_image_subbands");
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
    const size_t budget = 96 * 1024;
    int rc_max = (int)((budget - (size_t)n_grid * sizeof(float2)) / ((size_t)n_grid * sizeof(float2)));
    rc_max = std::min(std::min(rc_max, 2 * IMG_MAX_KB), n_rows);
    if (rc_max < 1) return set_error(DSX_ERANGE, "dsx_dirty_image: n_grid too large");
    ImgParams P{nullptr, nullptr, nullptr, nullptr, (const float2 *)grid, img, n_images, 1, 1, 1,
                n_grid, log2N, n_rows, rc_max, 0, prescale, 1};
    const size_t lds = (size_t)n_grid * sizeof(float2) * (1 + rc_max);
    (void)hipFuncSetAttribute((const void *)image_kernel<float>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    image_kernel<float><<<std::min(n_images, 4096), IMG_THREADS, lds, as_stream(stream)>>>(P);
    return check_launch("dsx_dirty_image");
}
