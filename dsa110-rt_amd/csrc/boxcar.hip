// Stand-alone detector primitives behind the reference detector API:
// centred boxcar along time, and the fused multi-width boxcar argmax over a
// padded cumulative sum.
#include "dsx_internal.h"

namespace dsx {

// out[i][j] = sum_{k<w} x[i - w/2 + k][j], zeros outside [0, n_t).
// One thread per column j; each output is summed directly in window order.
__global__ __launch_bounds__(256) void boxcar_kernel(const float *x, float *out, int n_t,
                                                     int n_inner, int w) {
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= n_inner) return;
    const int lo_off = w / 2;
    for (int i = 0; i < n_t; ++i) {
        const int lo = i - lo_off;
        float acc = 0.f;
        for (int k = 0; k < w; ++k) {
            const int t = lo + k;
            if (t >= 0 && t < n_t) acc += x[(int64_t)t * n_inner + j];
        }
        out[(int64_t)i * n_inner + j] = acc;
    }
}

template <typename TI, typename TO>
__global__ __launch_bounds__(256) void multi_boxcar_kernel(const TI *cs, const int32_t *widths,
                                                           const int32_t *offsets, const float *sinv,
                                                           int n_k, TO *out_max, int16_t *out_win,
                                                           int n_out, int64_t n_inner, int t_base,
                                                           int n_tpad) {
    const int64_t n = (int64_t)n_out * n_inner;
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int t = (int)(g / n_inner);
        const int64_t j = g - (int64_t)t * n_inner;
        const int ta = t + t_base;
        float best = 0.f;
        int win = 0;
        for (int k = 0; k < n_k; ++k) {
            const int o = offsets[k], w = widths[k];
            float snr = -INFINITY;
            if (ta + o >= 0 && ta + o + w < n_tpad) {
                const float lo = (float)cs[(int64_t)(ta + o) * n_inner + j];
                const float hi = (float)cs[(int64_t)(ta + o + w) * n_inner + j];
                snr = (hi - lo) * sinv[k];
            }
            if (k == 0 || snr > best) {
                best = snr;
                win = k;
            }
        }
        out_max[g] = (TO)best;
        out_win[g] = (int16_t)win;
    }
}

template <typename TI>
static int launch_mb(const TI *cs, const int32_t *widths, const int32_t *offsets, const float *sinv,
                     int n_k, void *out_max, int out_dtype, int16_t *out_win, int n_out,
                     int n_inner, int t_base, int n_tpad, hipStream_t s) {
    const int64_t n = (int64_t)n_out * n_inner;
    const int blocks = (int)std::min<int64_t>((n + 255) / 256, 1 << 20);
    if (out_dtype == DSX_F32)
        multi_boxcar_kernel<TI, float><<<blocks, 256, 0, s>>>(cs, widths, offsets, sinv, n_k,
                                                              (float *)out_max, out_win, n_out,
                                                              n_inner, t_base, n_tpad);
    else if (out_dtype == DSX_F16)
        multi_boxcar_kernel<TI, _Float16><<<blocks, 256, 0, s>>>(cs, widths, offsets, sinv, n_k,
                                                                 (_Float16 *)out_max, out_win,
                                                                 n_out, n_inner, t_base, n_tpad);
    else
        return set_error(DSX_EINVAL, "dsx_multi_boxcar_argmax: out_dtype");
    return check_launch("dsx_multi_boxcar_argmax");
}

}  // namespace dsx

using namespace dsx;

extern "C" int dsx_boxcar(const float *x, float *out, int n_t, int n_inner, int width,
                          dsx_stream_t stream) {
    DSX_REQUIRE(x && out, "dsx_boxcar: null pointer");
    DSX_REQUIRE(n_t > 0 && n_inner > 0 && width >= 1 && width <= n_t, "dsx_boxcar: bad shape");
    const int blocks = (n_inner + 255) / 256;
    boxcar_kernel<<<blocks, 256, 0, as_stream(stream)>>>(x, out, n_t, n_inner, width);
    return check_launch("dsx_boxcar");
}

extern "C" int dsx_multi_boxcar_argmax(const void *cs, int cs_dtype, const int32_t *widths,
                                       const int32_t *offsets, const float *sigma_inv, int n_k,
                                       void *out_max, int out_dtype, int16_t *out_win, int n_tpad,
                                       int n_out, int n_inner, int t_base, dsx_stream_t stream) {
    DSX_REQUIRE(cs && widths && offsets && sigma_inv && out_max && out_win,
                "dsx_multi_boxcar_argmax: null pointer");
    DSX_REQUIRE(n_k >= 1 && n_k <= 32, "dsx_multi_boxcar_argmax: n_k must be in [1, 32]");
    DSX_REQUIRE(n_out > 0 && n_inner > 0 && t_base >= 0 && n_tpad > 0,
                "dsx_multi_boxcar_argmax: bad shape");
    hipStream_t s = as_stream(stream);
    if (cs_dtype == DSX_F32)
        return launch_mb<float>((const float *)cs, widths, offsets, sigma_inv, n_k, out_max, out_dtype,
                                out_win, n_out, n_inner, t_base, n_tpad, s);
    if (cs_dtype == DSX_F16)
        return launch_mb<_Float16>((const _Float16 *)cs, widths, offsets, sigma_inv, n_k, out_max,
                                   out_dtype, out_win, n_out, n_inner, t_base, n_tpad, s);
    return set_error(DSX_EINVAL, "dsx_multi_boxcar_argmax: cs_dtype");
}
