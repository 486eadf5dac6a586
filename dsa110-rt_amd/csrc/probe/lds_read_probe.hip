// LDS read throughput of one MI355X (gfx950) for the access patterns of the
// dedispersion kernel's window reads.  Stand-alone measurement program, not
// part of libdsx:
//   hipcc -O3 --offload-arch=gfx950 -Xclang -target-feature -Xclang -packed-fp32-ops \
//       dsa110-rt_amd/csrc/probe/lds_read_probe.hip -o /tmp/lds_read_probe && /tmp/lds_read_probe
// Each work-group (256 threads, 64 KB of LDS) fills its array, then every lane
// reads 4 floats per iteration for ITERS iterations and adds them (one add per
// float, well under the VALU issue rate).  Every index is masked into the
// array.  Lane bytes per second = work-groups x 256 x ITERS x 16 / time; the
// shader clock under load comes from clock64() against the 100 MHz
// wall_clock64() inside the kernel.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

constexpr int NF = 16384;                 // floats per work-group
constexpr int MASK = NF - 1 - 1024;       // base indices stay below NF - 1024
constexpr int ITERS = 4096;

enum Pattern { CONSEC_B32, BROADCAST_B32, PERLANE_B32, CLUSTER_B32, CONSEC_B128, PERLANE_B64, PERLANE_B128,
               CLUSTER_B128, CLUSTER_B64 };

template <int P>
__global__ __launch_bounds__(256) void lds_read(float *out, long long *clk, int salt) {
    __shared__ __align__(16) float s[NF];
    const int tid = threadIdx.x, lane = tid & 63;
    for (int i = tid; i < NF; i += 256) s[i] = (float)(i ^ salt);
    __syncthreads();
    const long long c0 = clock64(), w0 = wall_clock64();
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
    const int spread = tid * 37;                                   // per-lane bases, distinct banks
    const int cluster = (lane * 3) / 8 + 300 * (tid >> 6);        // neighbouring lanes within a few words
    for (int it = 0; it < ITERS; ++it) {
        if constexpr (P == CONSEC_B32) {
            const int b = (it * 256 + lane) & MASK;
            a0 += s[b]; a1 += s[b + 64]; a2 += s[b + 128]; a3 += s[b + 192];
        } else if constexpr (P == BROADCAST_B32) {
            const int b = (it * 256) & MASK;
            a0 += s[b]; a1 += s[b + 64]; a2 += s[b + 128]; a3 += s[b + 192];
        } else if constexpr (P == PERLANE_B32 || P == CLUSTER_B32) {
            const int b = ((P == PERLANE_B32 ? spread : cluster) + it * 4) & MASK;
            a0 += s[b]; a1 += s[b + 1]; a2 += s[b + 2]; a3 += s[b + 3];
        } else if constexpr (P == CONSEC_B128) {
            const int b = (it * 256 + lane * 4) & MASK & ~3;
            const float4 v = *reinterpret_cast<const float4 *>(&s[b]);
            a0 += v.x; a1 += v.y; a2 += v.z; a3 += v.w;
        } else if constexpr (P == PERLANE_B64 || P == CLUSTER_B64) {
            const int b = (((P == PERLANE_B64 ? spread : cluster) & ~1) + it * 4) & MASK & ~1;
            const float2 u = *reinterpret_cast<const float2 *>(&s[b]);
            const float2 v = *reinterpret_cast<const float2 *>(&s[b + 2]);
            a0 += u.x; a1 += u.y; a2 += v.x; a3 += v.y;
        } else {
            const int b = ((P == PERLANE_B128 ? spread : cluster) + it * 4) & MASK & ~3;
            const float4 v = *reinterpret_cast<const float4 *>(&s[b]);
            a0 += v.x; a1 += v.y; a2 += v.z; a3 += v.w;
        }
    }
    const long long c1 = clock64(), w1 = wall_clock64();
    out[blockIdx.x * 256 + tid] = a0 + a1 + a2 + a3;
    if (tid == 0) {
        clk[2 * blockIdx.x] = c1 - c0;
        clk[2 * blockIdx.x + 1] = w1 - w0;
    }
}

#define CHECK(x)                                                                        \
    do {                                                                                \
        hipError_t e_ = (x);                                                            \
        if (e_ != hipSuccess) {                                                         \
            std::printf("%s failed: %s\n", #x, hipGetErrorString(e_));                  \
            std::exit(1);                                                               \
        }                                                                               \
    } while (0)

template <int P>
void run(const char *name, int n_cu, int nblk, float *out, long long *clk) {
    hipEvent_t a, b;
    CHECK(hipEventCreate(&a));
    CHECK(hipEventCreate(&b));
    lds_read<P><<<nblk, 256>>>(out, clk, 1);
    CHECK(hipDeviceSynchronize());
    CHECK(hipEventRecord(a));
    lds_read<P><<<nblk, 256>>>(out, clk, 2);
    CHECK(hipEventRecord(b));
    CHECK(hipEventSynchronize(b));
    float ms = 0.f;
    CHECK(hipEventElapsedTime(&ms, a, b));
    std::vector<long long> h(2 * nblk);
    CHECK(hipMemcpy(h.data(), clk, h.size() * sizeof(long long), hipMemcpyDeviceToHost));
    double cyc = 0, wall = 0;
    for (int i = 0; i < nblk; ++i) {
        cyc += (double)h[2 * i];
        wall += (double)h[2 * i + 1];
    }
    const double ghz = cyc / (wall / 100e6) / 1e9;
    const double bytes = (double)nblk * 256 * ITERS * 16.0;
    const double tbps = bytes / (ms * 1e-3) / 1e12;
    std::printf("%-34s %8.3f ms  clock %.2f GHz  %6.1f TB/s  %6.1f B/clk/CU\n", name, ms, ghz, tbps,
                bytes / (ms * 1e-3) / (ghz * 1e9) / n_cu);
    CHECK(hipEventDestroy(a));
    CHECK(hipEventDestroy(b));
}

int main() {
    int dev = 0;
    CHECK(hipGetDevice(&dev));
    hipDeviceProp_t p;
    CHECK(hipGetDeviceProperties(&p, dev));
    const int n_cu = p.multiProcessorCount;
    std::printf("%s: %d CUs, LDS per work-group %zu B\n", p.gcnArchName, n_cu, p.sharedMemPerBlock);
    const int nblk = n_cu * 16;
    float *out = nullptr;
    long long *clk = nullptr;
    CHECK(hipMalloc(&out, (size_t)nblk * 256 * sizeof(float)));
    CHECK(hipMalloc(&clk, (size_t)2 * nblk * sizeof(long long)));
    run<CONSEC_B32>("b32, consecutive lanes", n_cu, nblk, out, clk);
    run<BROADCAST_B32>("b32, one address per wave", n_cu, nblk, out, clk);
    run<PERLANE_B32>("b32 pairs, per-lane bases", n_cu, nblk, out, clk);
    run<CLUSTER_B32>("b32 pairs, clustered bases", n_cu, nblk, out, clk);
    run<CONSEC_B128>("b128, consecutive lanes", n_cu, nblk, out, clk);
    run<PERLANE_B64>("b64, per-lane bases", n_cu, nblk, out, clk);
    run<PERLANE_B128>("b128, per-lane bases", n_cu, nblk, out, clk);
    run<CLUSTER_B128>("b128, clustered bases", n_cu, nblk, out, clk);
    run<CLUSTER_B64>("b64, clustered bases", n_cu, nblk, out, clk);
    CHECK(hipFree(out));
    CHECK(hipFree(clk));
    return 0;
}
