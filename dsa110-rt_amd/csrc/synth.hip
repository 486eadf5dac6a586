// Counter-based synthetic 4-bit voltages (bit-identical to dsx.synth.synth_block_np).
#include "dsx_internal.h"

namespace dsx {

__device__ __forceinline__ int top16(uint64_t seed, uint64_t idx) {
    uint64_t z = seed + idx * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (int)(z >> 48);
}

// round half to even of v / 1024, clipped to [-7, 7]
__device__ __forceinline__ int quant(int64_t v) {
    int64_t q = v >> 10;
    const int64_t rem = v - (q << 10);
    q += (rem > 512) + ((rem == 512) && (q & 1));
    return (int)(q < -7 ? -7 : (q > 7 ? 7 : q));
}

__global__ __launch_bounds__(256) void synth_kernel(uint8_t *raw, uint64_t seed, int64_t sample0,
                                                    int n_pkt, int n_ant, int n_chan,
                                                    const int32_t *lut, int n_burst,
                                                    const int64_t *bstart, const int32_t *blen,
                                                    const int32_t *bph) {
    // one thread per (pkt, ant, chan) word of 4 bytes (2 t x 2 pol)
    const int64_t n = (int64_t)n_pkt * n_ant * n_chan;
    for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w < n;
         w += (int64_t)gridDim.x * blockDim.x) {
        const int ch = (int)(w % n_chan);
        const int64_t r = w / n_chan;
        const int ant = (int)(r % n_ant);
        const int64_t pkt = r / n_ant;
        uint32_t word = 0;
        for (int t = 0; t < 2; ++t) {
            const int64_t samp = sample0 + 2 * pkt + t;
            int64_t add_re = 0, add_im = 0;
            for (int b = 0; b < n_burst; ++b) {
                const int64_t st = bstart[(int64_t)b * n_chan + ch];
                if (samp >= st && samp < st + blen[b]) {
                    const int64_t o = (((int64_t)b * n_ant + ant) * n_chan + ch) * 2;
                    add_re += bph[o];
                    add_im += bph[o + 1];
                }
            }
            for (int p = 0; p < 2; ++p) {
                const uint64_t base = (uint64_t)((((samp * n_ant + ant) * n_chan + ch) * 2 + p) * 2);
                const int qr = quant((int64_t)lut[top16(seed, base)] + add_re);
                const int qi = quant((int64_t)lut[top16(seed, base + 1)] + add_im);
                const uint32_t byte = (uint32_t)((qr & 0xF) | ((qi & 0xF) << 4));
                word |= byte << (8 * (2 * t + p));
            }
        }
        *reinterpret_cast<uint32_t *>(raw + 4 * w) = word;
    }
}

}  // namespace dsx

using namespace dsx;

extern "C" int dsx_synth_block(uint8_t *raw, uint64_t seed, int64_t sample0, int n_pkt, int n_ant,
                               int n_chan, const int32_t *lut, int n_burst,
                               const int64_t *burst_start, const int32_t *burst_len,
                               const int32_t *burst_phasor, dsx_stream_t stream) {
    DSX_REQUIRE(raw && lut, "dsx_synth_block: null pointer");
    DSX_REQUIRE(n_pkt > 0 && n_ant > 0 && n_chan > 0 && sample0 >= 0 && sample0 % 2 == 0,
                "dsx_synth_block: bad shape");
    DSX_REQUIRE(n_burst == 0 || (burst_start && burst_len && burst_phasor),
                "dsx_synth_block: burst tables missing");
    const int64_t n = (int64_t)n_pkt * n_ant * n_chan;
    const int blocks = (int)std::min<int64_t>((n + 255) / 256, 1 << 20);
    synth_kernel<<<blocks, 256, 0, as_stream(stream)>>>(raw, seed, sample0, n_pkt, n_ant, n_chan, lut,
                                                        n_burst, burst_start, burst_len, burst_phasor);
    return check_launch("dsx_synth_block");
}
