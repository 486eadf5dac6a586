// Host-side driver for dsx_dedisp_plan under AddressSanitizer and
// UndefinedBehaviorSanitizer: every mode over a set of shift tables
// (the 16-subband 256-DM node, an odd subband count, tiny and ragged
// plans, shifts that are all equal), plus the error paths.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/dsx.h"

static std::vector<int32_t> table(int n_sub, int n_dm, double span, unsigned seed) {
    std::vector<int32_t> sh((size_t)n_sub * n_dm);
    srand(seed);
    for (int s = 0; s < n_sub; ++s)
        for (int d = 0; d < n_dm; ++d) {
            const double x = span * (double)d / (double)(n_dm > 1 ? n_dm - 1 : 1) * (double)(n_sub - s) / n_sub;
            sh[(size_t)s * n_dm + d] = (int32_t)(x + 0.5) + (rand() % 3 == 0 ? (rand() % 2) : 0);
        }
    return sh;
}

static int run(const std::vector<int32_t> &sh, int n_sub, int n_dm, int n_widths, int n_centres, int dtype,
               int budget, int mode) {
    std::vector<int32_t> sub_info(3 * (size_t)n_sub), tbase((size_t)n_sub * n_dm), info(8);
    const int rows_cap = n_sub * n_dm + 1;
    std::vector<int32_t> rows(6 * (size_t)rows_cap);
    const int rc = dsx_dedisp_plan(sh.data(), n_sub, n_dm, n_widths, n_centres, dtype, budget, mode,
                                   sub_info.data(), rows.data(), rows_cap, tbase.data(), info.data());
    if (rc == DSX_OK) {
        // every term base lies inside the windows / rows the plan reserves (pair rows
        // are indexed in row elements: two per 4-byte word when stored as BF16)
        const int64_t total = info[5] ? (int64_t)info[4] * ((info[7] & DSX_DD_ROWS_BF16) ? 2 : 1)
                                      : (int64_t)info[3] + info[4];
        for (int64_t i = 0; i < (int64_t)info[2] * n_dm; ++i)
            if (tbase[i] < 0 || tbase[i] >= total) {
                fprintf(stderr, "term base %d outside [0, %lld)\n", tbase[i], (long long)total);
                return 2;
            }
    }
    return rc == DSX_OK ? 0 : 1;
}

int main() {
    int fails = 0, ok = 0, refused = 0;
    const int shapes[][4] = {{16, 256, 6, 2048}, {7, 300, 5, 256}, {1, 1, 1, 64}, {64, 64, 4, 256},
                             {3, 17, 7, 128}, {16, 96, 7, 1024}};
    for (const auto &sp : shapes) {
        for (double span : {0.0, 40.0, 160.0}) {
            const auto sh = table(sp[0], sp[1], span, 7u + (unsigned)sp[0]);
            for (int mode = 0; mode <= 5; ++mode)
                for (int dtype : {DSX_F32, DSX_F16})
                    for (int budget : {16 * 1024, 48 * 1024, 160 * 1024}) {
                        const int r = run(sh, sp[0], sp[1], sp[2], sp[3], dtype, budget, mode);
                        if (r == 2) ++fails;
                        else if (r == 0) ++ok;
                        else ++refused;
                    }
        }
    }
    // argument errors are refused, never read out of bounds
    std::vector<int32_t> sh(16 * 256, 0), si(48), rows(6), tb(16 * 256), info(8);
    if (dsx_dedisp_plan(sh.data(), 16, 256, 6, 100, DSX_F32, 65536, 0, si.data(), rows.data(), 1, tb.data(),
                        info.data()) == DSX_OK) ++fails;
    if (dsx_dedisp_plan(sh.data(), 16, 256, 9, 2048, DSX_F32, 65536, 0, si.data(), rows.data(), 1, tb.data(),
                        info.data()) == DSX_OK) ++fails;
    if (dsx_dedisp_plan(nullptr, 16, 256, 6, 2048, DSX_F32, 65536, 0, si.data(), rows.data(), 1, tb.data(),
                        info.data()) == DSX_OK) ++fails;
    printf("plans built %d, refused %d, bad %d\n", ok, refused, fails);
    return fails == 0 && ok > 0 ? 0 : 1;
}
