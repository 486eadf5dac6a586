// Stand-alone shift-and-add combine (the reference's per-chgroup uv-stream
// combine): out[t][k] = sum_g in[g][t0 + t - shift[g]][k], zero outside
// [0, T_stream).  Elements are float32 (complex streams pass 2 per cell).
#include "dsx_internal.h"

namespace dsx {

__global__ __launch_bounds__(256) void combine_kernel(const float *in, const int32_t *shift, float *out,
                                                      int n_g, int t_stream, int64_t n_inner, int t0,
                                                      int n_t) {
    const int64_t n = (int64_t)n_t * n_inner;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int t = (int)(i / n_inner);
        const int64_t k = i - (int64_t)t * n_inner;
        float acc = 0.f;
        for (int g = 0; g < n_g; ++g) {
            const int ts = t0 + t - shift[g];
            if (ts >= 0 && ts < t_stream) acc += in[((int64_t)g * t_stream + ts) * n_inner + k];
        }
        out[i] = acc;
    }
}

}  // namespace dsx

using namespace dsx;

extern "C" int dsx_combine(const float *streams, const int32_t *shift, float *out, int n_streams,
                           int t_stream, int64_t n_inner, int t0, int n_t, dsx_stream_t stream) {
    DSX_REQUIRE(streams && shift && out, "dsx_combine: null pointer");
    DSX_REQUIRE(n_streams > 0 && t_stream > 0 && n_inner > 0 && n_t > 0, "dsx_combine: bad shape");
    const int64_t n = (int64_t)n_t * n_inner;
    const int blocks = (int)std::min<int64_t>((n + 255) / 256, 1 << 20);
    combine_kernel<<<blocks, 256, 0, as_stream(stream)>>>(streams, shift, out, n_streams, t_stream, n_inner,
                                                          t0, n_t);
    return check_launch("dsx_combine");
}
