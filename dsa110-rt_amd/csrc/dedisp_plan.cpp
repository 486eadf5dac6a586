// Host-side LDS planner of the fused dedispersion kernel (dsx_dedisp_plan):
// per-subband window offsets, subband-pair rows and each lane's term bases for
// one DM plan, block length and LDS budget.  Plain C++ (no device code), so it
// also builds with the host compiler alone for sanitizer runs.
#include <algorithm>
#include <climits>
#include <vector>

#include "dedisp_plan.h"
#include "dsx_internal.h"

namespace dsx {

// Pair rows of a plan: per pair g the sorted distinct deltas and, per delta,
// the range [lo, hi] of sh_a over the DMs that use it.
struct PairRows {
    std::vector<int> a, b, delta, lo, hi;   // one entry per row
    std::vector<int> dmin;                   // first DM (lane order) using the row
    std::vector<int> first;                  // first row of pair g, size n_pair + 1
};

static PairRows pair_rows(const int32_t *sh, int n_sub, int n_dm) {
    PairRows pr;
    const int n_pair = (n_sub + 1) / 2;
    pr.first.push_back(0);
    for (int g = 0; g < n_pair; ++g) {
        const int a = 2 * g, b = 2 * g + 1 < n_sub ? 2 * g + 1 : -1;
        std::vector<int> dl;
        for (int d = 0; d < n_dm; ++d)
            dl.push_back(b >= 0 ? sh[(int64_t)b * n_dm + d] - sh[(int64_t)a * n_dm + d] : 0);
        std::vector<int> u = dl;
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        for (int v : u) {
            int lo = INT32_MAX, hi = INT32_MIN, dm0 = n_dm;
            for (int d = 0; d < n_dm; ++d)
                if (dl[d] == v) {
                    lo = std::min(lo, (int)sh[(int64_t)a * n_dm + d]);
                    hi = std::max(hi, (int)sh[(int64_t)a * n_dm + d]);
                    dm0 = std::min(dm0, d);
                }
            pr.dmin.push_back(dm0);
            pr.a.push_back(a);
            pr.b.push_back(b);
            pr.delta.push_back(v);
            pr.lo.push_back(lo);
            pr.hi.push_back(hi);
        }
        pr.first.push_back((int)pr.a.size());
    }
    return pr;
}
}  // namespace dsx

using namespace dsx;

extern "C" int dsx_dedisp_plan(const int32_t *host_shift, int n_sub, int n_dm, int n_widths,
                               int n_centres, int ring_dtype, int lds_budget, int mode,
                               int32_t *sub_info, int32_t *rows, int rows_cap, int32_t *tbase,
                               int32_t *info) {
    DSX_REQUIRE(host_shift && sub_info && rows && tbase && info, "dsx_dedisp_plan: null pointer");
    DSX_REQUIRE(n_sub > 0 && n_sub <= 1024, "dsx_dedisp_plan: n_sub=%d not in [1, 1024]", n_sub);
    DSX_REQUIRE(n_dm > 0 && n_dm <= 32767, "dsx_dedisp_plan: n_dm=%d", n_dm);
    DSX_REQUIRE(n_widths >= 1 && n_widths <= 7, "dsx_dedisp_plan: n_widths=%d", n_widths);
    DSX_REQUIRE(n_centres > 0 && n_centres % DD_CHUNK == 0,
                "dsx_dedisp_plan: n_centres must be a positive multiple of 64");
    DSX_REQUIRE(ring_dtype == DSX_F16 || ring_dtype == DSX_F32, "dsx_dedisp_plan: ring_dtype");
    DSX_REQUIRE(mode >= 0 && mode <= 5, "dsx_dedisp_plan: mode=%d not in [0, 5]", mode);
    DSX_REQUIRE((mode != 3 && mode != 4) || ring_dtype == DSX_F32,
                "dsx_dedisp_plan: mode %d (LDS-direct window loads) needs a float32 history", mode);
    const int ww = n_widths > 1 ? (1 << (n_widths - 1)) : 0;
    const int h2 = ww / 2;
    std::vector<int> shmax(n_sub), shmin(n_sub);
    for (int s = 0; s < n_sub; ++s) {
        int hi = host_shift[(int64_t)s * n_dm], lo = hi;
        for (int d = 1; d < n_dm; ++d) {
            hi = std::max(hi, (int)host_shift[(int64_t)s * n_dm + d]);
            lo = std::min(lo, (int)host_shift[(int64_t)s * n_dm + d]);
        }
        shmax[s] = hi;
        shmin[s] = lo;
    }
    const PairRows pr = pair_rows(host_shift, n_sub, n_dm);
    const int n_prow = (int)pr.a.size();
    auto r_elems = [&](int tt) {
        int64_t r = 0;
        for (int s = 0; s < n_sub; ++s) r += tt + ww + shmax[s] - shmin[s];
        return r;
    };
    // pair-row offsets.  Lane d reads row r at qoff[r] + hi_r - sh_a[d] (+ the
    // step), a run that descends with d inside a row.  With `banks`, each
    // pair's rows are laid out in the order their DMs appear and every row is
    // padded so that its first lane's address sits one bank below the last
    // lane of the previous row: the lanes of a wave then read one descending
    // run of distinct banks (64 of them, 4 bytes each) across row boundaries.
    const bool banks = mode == 4;
    auto q_layout = [&](int tt, std::vector<int> *qoff) {
        int64_t q = 0;
        std::vector<int> order(n_prow);
        for (int r = 0; r < n_prow; ++r) order[r] = r;
        if (banks)
            for (size_t g = 0; g + 1 < pr.first.size(); ++g)
                std::sort(order.begin() + pr.first[g], order.begin() + pr.first[g + 1],
                          [&](int x, int y) { return pr.dmin[x] < pr.dmin[y]; });
        if (qoff) qoff->assign(n_prow, 0);
        int64_t prev_base = 0;     // address of the previous row's last lane
        for (size_t g = 0; g + 1 < pr.first.size(); ++g)
            for (int i = pr.first[g]; i < pr.first[g + 1]; ++i) {
                const int r = order[i];
                const int span = pr.hi[r] - pr.lo[r];
                if (banks && i > pr.first[g]) {
                    // want (q + span) = prev_base - 1  (mod 64)
                    const int64_t want = ((prev_base - 1 - span) % 64 + 64) % 64;
                    q += ((want - q) % 64 + 64) % 64;
                }
                if (qoff) (*qoff)[r] = (int)q;
                prev_base = q;
                q += tt + ww + span;
            }
        return q;
    };
    // bytes of one pair-row element: mode 5 stores the rows as BF16
    const int q_esz = mode == 5 ? 2 : 4;
    auto q_elems = [&](int tt) { return (q_layout(tt, nullptr) * q_esz + 3) / 4; };   // in 4-byte words
    // largest tile within the LDS budget, preferring tiles whose windows fit
    // the batched staging of the fast kernel
    const int threads = std::min(DD_THREADS, ((n_dm + 63) / 64) * 64);
    auto best_tile = [&](bool paired, bool dbl) {
        int best = 0, best_fast = 0;
        const int n_terms = paired ? (n_sub + 1) / 2 : n_sub;
        for (int t = DD_CHUNK; t <= n_centres && t <= 2048; t += DD_CHUNK) {
            const int64_t bytes = (paired ? q_elems(t) * 4 : 0) + r_elems(t) * 4 * (dbl ? 2 : 1) +
                                  ((int64_t)tb_words(n_terms) * threads + 3 * n_sub) * 4;
            if (n_centres % t != 0 || bytes > lds_budget) continue;
            best = t;
            bool fits = true;
            for (int s2 = 0; s2 < n_sub; ++s2) fits = fits && t + ww + shmax[s2] - shmin[s2] <= 64 * DD_STAGE_ITERS;
            if (fits) best_fast = t;
        }
        return best_fast > 0 ? best_fast : best;
    };
    bool paired = false;
    int tt = 0;
    if (mode != 2 && mode != 3 && n_prow <= DD_MAX_ROWS) {
        tt = best_tile(true, mode == 4);
        paired = tt > 0;
    }
    if (!paired) {
        if (mode == 1 || mode == 4 || mode == 5)
            return set_error(DSX_ERANGE, "dsx_dedisp_plan: pair rows need %lld B of LDS > %d",
                             (long long)(q_elems(DD_CHUNK) * 4 + r_elems(DD_CHUNK) * 4), lds_budget);
        tt = best_tile(false, mode == 3);
    }
    if (tt == 0)
        return set_error(DSX_ERANGE, "dsx_dedisp_plan: the 64-centre tile needs %lld B of LDS > %d",
                         (long long)(r_elems(DD_CHUNK) * 4), lds_budget);
    const int n_rows = paired ? n_prow : 0;
    if (n_rows > rows_cap)
        return set_error(DSX_ERANGE, "dsx_dedisp_plan: %d pair rows > rows_cap %d", n_rows, rows_cap);
    // subband windows: R index of step tau for subband s = off_s + tau - (tile start - W/2 - shmax_s)
    int r_total = 0;
    std::vector<int> roff(n_sub);
    for (int s = 0; s < n_sub; ++s) {
        const int len = tt + ww + shmax[s] - shmin[s];
        roff[s] = r_total;
        sub_info[3 * s] = r_total;
        sub_info[3 * s + 1] = shmax[s];
        sub_info[3 * s + 2] = len;
        r_total += len;
    }
    int q_total = mode == 3 ? r_total : 0;     // mode 3: the second window buffer
    if (paired) {
        // row r covers tau in [tile start - W/2 - hi_r, + len_r) of its pair's DMs
        std::vector<int> qoff;                         // in row elements
        q_total = (int)((q_layout(tt, &qoff) * q_esz + 3) / 4);
        for (int r = 0; r < n_prow; ++r) {
            const int a = pr.a[r], b = pr.b[r];
            const int len = tt + ww + pr.hi[r] - pr.lo[r];
            rows[6 * r] = qoff[r];
            rows[6 * r + 1] = len;
            rows[6 * r + 2] = roff[a] + shmax[a] - pr.hi[r];
            if (b >= 0) {
                rows[6 * r + 3] = roff[b] + shmax[b] - pr.hi[r] - pr.delta[r];
                rows[6 * r + 4] = roff[b];
                rows[6 * r + 5] = roff[b] + tt + ww + shmax[b] - shmin[b];
            } else {
                rows[6 * r + 3] = 0;
                rows[6 * r + 4] = 0;
                rows[6 * r + 5] = 0;
            }
        }
        const int n_pair = (n_sub + 1) / 2;
        for (int g = 0; g < n_pair; ++g) {
            const int a = 2 * g, b = 2 * g + 1 < n_sub ? 2 * g + 1 : -1;
            for (int d = 0; d < n_dm; ++d) {
                const int sa = host_shift[(int64_t)a * n_dm + d];
                const int dl = b >= 0 ? host_shift[(int64_t)b * n_dm + d] - sa : 0;
                int r = pr.first[g];
                while (pr.delta[r] != dl) ++r;
                tbase[(int64_t)g * n_dm + d] = qoff[r] + h2 + pr.hi[r] - sa;
            }
        }
    } else {
        for (int s = 0; s < n_sub; ++s)
            for (int d = 0; d < n_dm; ++d)
                tbase[(int64_t)s * n_dm + d] = roff[s] + h2 + shmax[s] - host_shift[(int64_t)s * n_dm + d];
    }
    const int n_terms = paired ? (n_sub + 1) / 2 : n_sub;
    info[0] = tt;
    info[1] = n_rows;
    info[2] = n_terms;                              // terms per dedispersed sample
    info[3] = r_total;
    info[4] = q_total;
    info[5] = paired ? 1 : 0;
    const int64_t r_bufs = mode == 4 ? 2 : 1;     // mode 4: two window buffers after the pair rows
    info[6] = (int)(((int64_t)q_total + r_bufs * r_total + (int64_t)tb_words(n_terms) * threads + 3 * n_sub) * 4);
    info[7] = ((mode == 3 || mode == 4) ? DSX_DD_DMA : 0) | (mode == 5 ? DSX_DD_ROWS_BF16 : 0);
    return DSX_OK;
}

