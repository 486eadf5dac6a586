/*
 * Float64 restatement of subband dedispersion + centred boxcars + chunk
 * maxima + block statistics -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * Follows:
 *   shift-and-add combine, zero outside the data:
 *       src/dsart/fine_dm/combiner.py:625-653  (out[t] += stream[t - shift])
 *   centred unweighted boxcar [c - w/2, c - w/2 + w - 1]:
 *       src/dsart/detector/forward.py:163-247,456-504
 *   per-(centre chunk, cell) maxima, ties -> lowest filter then earliest centre:
 *       src/dsart/x2/ref/search2x_ref.py:264-282 (ref_chunk_max)
 * with the S/N of width k taken against the running per-(DM, pixel) mean and
 * variance of the dedispersed series: snr = (B_k - w_k mu) / (sigma sqrt(w_k)).
 *
 * Validity gate (the reference rejects cubes whose stream history has not
 * filled, src/dsart/fine_dm/combiner.py:580-585): a dedispersed sample at t
 * is complete once every subband contributes, t >= max_s shift[d][s].  Only
 * complete samples enter the block statistics, and a centre chunk is scored
 * only when all of its boxes are complete (first centre - wmax/2 >= that
 * bound) and the DM already has at least ``min_count`` complete samples of
 * statistics from earlier blocks.
 *
 * Image layout: img[s][p][tau - tau0] (time contiguous per pixel).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

int oracle_search_block(
    const double *img, int n_sub, int n_pix, int64_t tau0, int n_tau,
    const int32_t *shifts, int n_dm,
    int64_t c_lo, int n_centres, int chunk,
    const int32_t *widths, int n_w,
    const double *st_n, const double *st_mean, const double *st_m2, double min_count,
    double *best_snr, int64_t *best_t, int32_t *best_k,
    double *blk_sum, double *blk_sumsq, double *blk_count)
{
    if (n_centres % chunk) return -1;
    int wmax = 1;
    for (int k = 0; k < n_w; ++k) if (widths[k] > wmax) wmax = widths[k];
    const int half = wmax / 2;
    /* D over [c_lo - half, c_lo + n_centres + half) */
    const int nd = n_centres + 2 * half + 1;
    const int n_chunks = n_centres / chunk;
    int status = 0;

#pragma omp parallel for schedule(dynamic)
    for (int d = 0; d < n_dm; ++d) {
        double *D = (double *)malloc(sizeof(double) * nd);
        double *P = (double *)malloc(sizeof(double) * (nd + 1));
        if (!D || !P) {
            free(D); free(P);
#pragma omp atomic write
            status = -2;
            continue;
        }
        int64_t valid_from = 0;
        for (int s = 0; s < n_sub; ++s)
            if (shifts[(size_t)d * n_sub + s] > valid_from) valid_from = shifts[(size_t)d * n_sub + s];
        int64_t cnt = 0;
        for (int64_t c = c_lo; c < c_lo + n_centres; ++c) if (c >= valid_from) ++cnt;
        blk_count[d] = (double)cnt;
        for (int p = 0; p < n_pix; ++p) {
            const int64_t t_first = c_lo - half;
            for (int i = 0; i < nd; ++i) {
                const int64_t t = t_first + i;
                double acc = 0.0;
                for (int s = 0; s < n_sub; ++s) {
                    const int64_t tau = t - shifts[(size_t)d * n_sub + s];
                    if (tau < 0) continue;                 /* before stream start */
                    const int64_t j = tau - tau0;
                    if (j < 0 || j >= n_tau) {
#pragma omp atomic write
                        status = -3;
                        continue;
                    }
                    acc += img[((size_t)s * n_pix + p) * n_tau + j];
                }
                D[i] = (t < 0) ? 0.0 : acc;
            }
            P[0] = 0.0;
            for (int i = 0; i < nd; ++i) P[i + 1] = P[i] + D[i];
            /* block statistics over the complete centres */
            double s1 = 0.0, s2 = 0.0;
            for (int i = 0; i < n_centres; ++i) {
                if (c_lo + i < valid_from) continue;
                const double v = D[half + i];
                s1 += v;
                s2 += v * v;
            }
            blk_sum[(size_t)d * n_pix + p] = s1;
            blk_sumsq[(size_t)d * n_pix + p] = s2;
            /* detection against the state before this block */
            const double n0 = st_n[d];
            const double mu = n0 > 0 ? st_mean[(size_t)d * n_pix + p] : 0.0;
            const double var = n0 > 0 ? st_m2[(size_t)d * n_pix + p] / n0 : 0.0;
            const double sig = sqrt(var);
            for (int ch = 0; ch < n_chunks; ++ch) {
                double bsnr = -INFINITY;
                int bt = -1, bk = -1;
                const int64_t first_box = c_lo + (int64_t)ch * chunk - half;
                if (n0 >= min_count && n0 > 0 && sig > 0.0 && first_box >= valid_from) {
                    for (int k = 0; k < n_w; ++k) {
                        const int w = widths[k];
                        double mb = -INFINITY;
                        int mt = -1;
                        for (int i = ch * chunk; i < (ch + 1) * chunk; ++i) {
                            /* centre c = c_lo + i; box [c - w/2, c - w/2 + w) ;
                               D index of time t is t - (c_lo - half) */
                            const int lo = half + i - w / 2;
                            const double b = P[lo + w] - P[lo];
                            if (b > mb) { mb = b; mt = i; }
                        }
                        const double snr = (mb - w * mu) / (sig * sqrt((double)w));
                        if (snr > bsnr) { bsnr = snr; bt = mt; bk = k; }
                    }
                }
                const size_t o = ((size_t)d * n_chunks + ch) * n_pix + p;
                best_snr[o] = bsnr;
                best_t[o] = bt < 0 ? -1 : c_lo + bt;
                best_k[o] = bk;
            }
        }
        free(D);
        free(P);
    }
    return status;
}
