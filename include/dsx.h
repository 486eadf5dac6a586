/*
 * dsx.h -- C ABI of libdsx.so, the MI355X (gfx950) search-node fast path:
 * 4-bit voltage unpack -> int8-MFMA correlator -> nearest-cell gridder ->
 * C2R FFT imager -> subband dedispersion fused with boxcar detection.
 *
 * Conventions
 *   - Every entry point returns 0 on success or a negative DSX_E* code;
 *     dsx_last_error() returns the message of the calling thread's last
 *     failure.  No entry point allocates or frees caller memory.
 *   - Pointers are device pointers (hipMalloc / torch CUDA tensors) unless
 *     the name says host_.  Work is enqueued on `stream` (0 = default);
 *     nothing synchronises unless stated.
 *   - Visibilities are int32 (re, im) pairs: exact integer correlations of
 *     the 4-bit samples, vis[bls] = sum conj(E_b) E_a with b <= a and
 *     bls = a (a + 1) / 2 + b.
 *
 * The reference (dsa110/dsa110-rt) has no native boundary on this path: its
 * Python stage functions call Triton / NVRTC kernels directly.  Each entry
 * point below names the reference function whose work it replaces.
 */
#ifndef DSX_H
#define DSX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t *dsx_stream_t; /* a hipStream_t */

enum {
    DSX_OK = 0,
    DSX_EINVAL = -1,   /* bad shape / argument */
    DSX_ELAUNCH = -2,  /* kernel launch failed */
    DSX_ERANGE = -3,   /* size beyond a documented limit */
    DSX_ENOTSUP = -4,  /* configuration not supported by this build */
};

enum { DSX_F32 = 0, DSX_F16 = 1, DSX_I8 = 2, DSX_I32 = 3 };

/* dsx_dedisp_plan / dsx_dedisp_detect layout flags */
enum {
    DSX_DD_DMA = 1,        /* windows copied HBM -> LDS by LDS-direct loads during compute */
    DSX_DD_ROWS_BF16 = 2,  /* subband-pair rows stored as BF16 */
};

/* Library version, (major << 16) | (minor << 8) | patch. */
int dsx_version(void);

/* Message of the calling thread's last failed call ("" if none). */
const char *dsx_last_error(void);

/* ---------------------------------------------------------------- unpack
 * On-wire bytes (n_pkt, n_ant, n_chan, 2 t, 2 pol) -> real and imag planes
 * in GEMM layout (n_chan, 2 t, 2 pol, n_pkt, n_ant).  out_dtype DSX_I8
 * writes the raw 4-bit integers (scale ignored); DSX_F16 / DSX_F32 write
 * value * scale.
 * Replaces unpack_int4_split, src/dsart/services/slow_corr_kernel.py:220-403
 * (Triton kernel services/triton_unpack_onwire.py:51-226). */
int dsx_unpack_i4(const uint8_t *raw, void *re, void *im, int n_pkt, int n_ant,
                  int n_chan, int out_dtype, float scale, dsx_stream_t stream);

/* ------------------------------------------------------------ correlator
 * Exact int8-MFMA correlation with int32 accumulation.
 *   input_layout 0: `raw` holds on-wire bytes (n_pkt, n_ant, n_chan, 2, 2);
 *                   `im` must be NULL.
 *   input_layout 1: `raw`/`im` hold int8 real/imag planes in GEMM layout
 *                   (n_chan, 2, 2, n_pkt, n_ant) (values in [-8, 7]).
 * Each output sums t_int native samples, csf adjacent channels and, when
 * fuse_pol != 0, both polarisations (Stokes I, no 1/2).
 *   out_layout 0: vis (n_steps, nbase, n_chan / csf, npol_out, 2)   (reference)
 *   out_layout 1: vis (n_steps, n_chan / csf, npol_out, nbase, 2)   (subband-major)
 * with n_steps = 2 n_pkt / t_int and npol_out = fuse_pol ? 1 : 2.
 * Replaces FastCorrKernel.compute_split + stokes_i_pol_sum,
 * src/dsart/services/corr_fast_kernel.py:364-792 (Triton
 * services/triton_corr_fused.py:44-209). */
int dsx_corr_i8(const void *raw, const void *im, int32_t *vis, int n_pkt, int n_ant,
                int n_chan, int t_int, int csf, int fuse_pol, int input_layout,
                int out_layout, dsx_stream_t stream);

/* Stage 1 of the coarse-DM plan folded into the correlator's operand load:
 * output step t of the block reads channel c's samples of step
 * t + chan_shift[c] - delay (device int32 [n_chan], each in [0, delay]);
 * steps before the block come from `tail`, the previous block's last `delay`
 * steps of raw bytes ((delay * t_int / 2, n_ant, n_chan, 2, 2), zeros before
 * the stream).  The block's output covers steps [T0 - delay, T0 + L - delay)
 * and stays exact int32.  The caller copies the block's last `delay` steps
 * into the tail afterwards.  Other arguments as dsx_corr_i8 (raw input).
 * Replaces Stage1MultiDMCoarseDM.dedisperse_from_vis (per-channel stage-1
 * shifts on the visibilities before gridding, with the sliding 2-block
 * window), src/dsart/services/corr_fast_integration.py:495-1040, with the
 * shifts of DMPlan.delay_bins_per_subband, coarse_dm/dm_plan.py:479-515. */
int dsx_corr_i8_stage1(const uint8_t *raw, const uint8_t *tail, const int32_t *chan_shift,
                       int delay, int32_t *vis, int n_pkt, int n_ant, int n_chan, int t_int,
                       int csf, int fuse_pol, int out_layout, dsx_stream_t stream);

/* ----------------------------------------------- calibrated correlator
 * Complex gains (and the RFI mask, folded in as zero gains) applied to every
 * voltage as it is loaded; fp16 MFMA with float32 accumulation.
 *   input_layout 0: volt_re holds on-wire bytes (n_pkt, n_ant, n_chan, 2, 2),
 *                   volt_im NULL; voltage v = fp16(code * scale).
 *   input_layout 1: volt_re / volt_im hold fp16 planes in GEMM layout
 *                   (n_chan, 2, 2, n_pkt, n_ant) (e.g. unpack_int4_split output).
 * cal_re / cal_im: float32 (n_chan, 2 pol, n_ant) gains, or both NULL (the
 * operands are the voltages).  time_keep: NULL or float32 (n_acc, n_chan,
 * 2 pol) factors multiplying the gain of every packet of accumulation
 * pkt / (n_pkt / n_acc) (n_acc divides n_pkt).  Operand
 * x = fp16(v_re g_re - v_im g_im, v_re g_im + v_im g_re), each product and
 * sum rounded in float32 (no fused multiply-add).  vis: float (re, im) pairs,
 * vis[bls] = sum conj(x_b) x_a, in the layouts of dsx_corr_i8 (out_layout
 * 0 / 1).  n_ant <= 128.
 * Replaces apply_cal_split + FastCorrKernel.compute_split
 * (src/dsart/services/slow_corr_kernel.py:459-569,
 * services/corr_fast_kernel.py:364-499) and the cal-on-load GEMM
 * src/dsart/x2/corr/gemm2x.py:158-168 (CPU twin gemm2x.py:346-390). */
int dsx_corr_cal(const void *volt_re, const void *volt_im, const float *cal_re,
                 const float *cal_im, const float *time_keep, int n_acc, float *vis,
                 int n_pkt, int n_ant, int n_chan, int t_int, int csf, int fuse_pol,
                 int input_layout, int out_layout, float scale, dsx_stream_t stream);

/* fp16 planes (n_chan, 2, 2, n_pkt, n_ant) times float32 gains
 * (n_chan, 2, n_ant) (cal_per_chan != 0) or (2, n_ant) (one gain for every
 * channel): out = fp16(v * g) with the operand arithmetic of dsx_corr_cal.
 * Replaces apply_cal_split, src/dsart/services/slow_corr_kernel.py:459-569. */
int dsx_apply_cal(const void *re, const void *im, const float *cal_re, const float *cal_im,
                  int cal_per_chan, void *out_re, void *out_im, int n_chan, int n_pkt,
                  int n_ant, dsx_stream_t stream);

/* --------------------------------------------------------------- gridder
 * Gather form of the nearest-cell gridder: for every step f and filled cell
 * c, out[f][c] = sum over sources i in [cell_ptr[c], cell_ptr[c+1]) of
 * vis[f][cell_src[i]] where vis is (n_steps, n_src) complex.  vis_dtype
 * DSX_I32 (int32 pairs, exact; out int32 pairs) or DSX_F32 (float pairs,
 * e.g. complex64; out float pairs, summed in CSR order).
 * Replaces FastVisGridder.compute, src/dsart/grid/kernel.py:508-611 (and the
 * CSR gather of services/triton_dedisp.py:50-95). */
int dsx_grid_gather(const void *vis, const int32_t *cell_ptr, const int32_t *cell_src,
                    void *out, int n_steps, int n_src, int n_cells, int vis_dtype,
                    dsx_stream_t stream);

/* ---------------------------------------------------------------- imager
 * Fused gridder + C2R FFT imager.  For every step t and subband s of the
 * subband-major visibilities vis (n_steps, n_sub * per_sub, nbase, 2) int32,
 * scatter taps [tap_off[s], tap_off[s+1]) (tap_off: device, n_sub + 1) into the Hermitian half plane
 * H[r][w] (r < n_rows), inverse-FFT, and write
 *     img[s][t][i][j] = mask01(i, j) * scale * C2R(H)[i][j] / (2 N^2)
 * which equals edge_mask * fftshift(Re(ifft2(grid))) of the reference grid.
 * tap_src indexes one (t, s) slab of per_sub * nbase pairs; tap_dst holds
 * cell r * N + w with bit 20 = conjugate and bit 21 = negate.  mask01 zeroes
 * the outer npad ring and the DC pixel (N/2, N/2).  img_dtype DSX_F16 or
 * DSX_F32; img is (n_sub, n_steps, N, N).
 * Replaces grid + dirty_image_from_uv_grid + apply_edge_mask,
 * src/dsart/image/imager.py:112-324 and GpuImager.process_cube,
 * src/dsart/image/imager_gpu.py:217-657. */
int dsx_image_subbands(const int32_t *vis, const int32_t *tap_src, const int32_t *tap_dst,
                       const int32_t *tap_off, void *img, int n_steps, int n_sub,
                       int per_sub, int nbase, int n_grid, int n_rows, int npad,
                       float scale, int img_dtype, dsx_stream_t stream);

/* dsx_image_subbands for float32 (re, im) visibilities (the calibrated
 * correlator's output, dsx_corr_cal out_layout 1): the half plane is summed
 * with float32 LDS atomics, so the order of a cell's additions (and its last
 * bits) is not fixed from run to run.
 * Replaces the same reference functions as dsx_image_subbands, fed the
 * calibrated visibilities of apply_cal_split (slow_corr_kernel.py:459-569). */
int dsx_image_subbands_f32(const float *vis, const int32_t *tap_src, const int32_t *tap_dst,
                           const int32_t *tap_off, void *img, int n_steps, int n_sub,
                           int per_sub, int nbase, int n_grid, int n_rows, int npad,
                           float scale, int img_dtype, dsx_stream_t stream);

/* Static-sky subtraction of a block of images img (n_sub, n_steps, n_pix)
 * (img_dtype DSX_F32 / DSX_F16), per (subband, pixel): when subtract != 0 and
 * count > 0 every step loses wsum / count (the mean of the previous `count`
 * blocks' means); then this block's mean (float64 sum over its steps) enters
 * the window: ring (window, n_sub, n_pix) float32 slot `head` (its old value
 * leaves wsum when count == window) and wsum (n_sub, n_pix) float64.  The
 * caller advances head and count.  Imaging is linear, so this is the image of
 * the per-cell uv subtraction of StaticSkyMean.apply,
 * src/dsart/services/corr_fast_integration.py:1049-1246. */
int dsx_static_sky(void *img, int img_dtype, int n_sub, int n_steps, int64_t n_pix, float *ring,
                   double *wsum, int window, int head, int count, int subtract,
                   dsx_stream_t stream);

/* Fused dequant + subband shift-and-add + C2R imager for a search cube.
 * For every trial f < n_fdm and cube time t in [t_lo, t_det):
 *   G[x][y]    = sum_g c_g(streams[g][t - shifts[f][g]][.][x][y])  (chgroups whose
 *                row t - shifts[f][g] lies in [0, t_stream); others add nothing)
 *   cube[t][f] = mask * fftshift(Re(ifft2(prescale * G)))           (numpy 1/N^2)
 * streams is (n_chg, t_stream, 2, N, N) int8 (real plane, then imaginary);
 * shifts is the device (n_fdm, n_chg) int32 table.  c_g(z) = z summed exactly
 * in int32 when cal_scale, cal_off_re and cal_off_im are all NULL, else
 * scale * z + offset in float32 with per-chgroup (cal_2d = 0: [n_chg]) or
 * per-(chgroup, stream row) (cal_2d = 1: [n_chg][t_stream]) scale and
 * offsets (NULL: 1 / 0).  mask is [N][N] float32 (or [n_fdm][N][N] with
 * mask_per_fdm), NULL for none.  cube is (t_det, n_fdm, N, N) of cube_dtype
 * DSX_F32 or DSX_F16; rows t < t_lo are not written.  The uv grid is never
 * written to HBM: each work-group folds the Hermitian half plane straight
 * from the int8 rows into LDS.  An fp16 cube needs acc, a float32 buffer of
 * the cube's shape for the partial images of the uv-row passes (N = 256:
 * four passes).
 * Replaces GpuImager.process_cube (dense path),
 * src/dsart/image/imager_gpu.py:217-657, with its combine
 * src/dsart/image/fused_combine_cuda.py:1104-1359. */
int dsx_image_cube(const int8_t *streams, const int32_t *shifts, int n_chg, int t_stream, int n_fdm,
                   int t_det, int t_lo, int n_grid, float prescale, const float *cal_scale,
                   const float *cal_off_re, const float *cal_off_im, int cal_2d, const float *mask,
                   int mask_per_fdm, void *cube, int cube_dtype, float *acc, dsx_stream_t stream);

/* Dense-grid imager: img[b] = fftshift(Re(ifft2(prescale * grid[b]))) for
 * n_images complex64 grids (n_images, N, N) -> float32 (n_images, N, N),
 * numpy's 1/N^2 normalisation, no mask.
 * Replaces dirty_image_from_uv_grid, src/dsart/image/imager.py:248-301. */
int dsx_dirty_image(const void *grid, void *img, int n_images, int n_grid, float prescale,
                    dsx_stream_t stream);

/* ------------------------------------------------------------ image ring
 * Time-major subband image history.  ring is (n_sub, n_pix, hist + cap)
 * with step T0 - hist + k at position k.  Copies the last `hist` steps of
 * ring_old (which held `prev_len` new steps after its history) into
 * ring_new[.., 0:hist), then transposes img (n_sub, n_new, n_pix) into
 * ring_new[.., hist:hist + n_new).  dtype DSX_F16 or DSX_F32 for both.
 * Replaces the per-chgroup stream history of combine_chgroups,
 * src/dsart/fine_dm/combiner.py:548-655 (T_stream = T_det + max shift). */
int dsx_ring_advance(const void *ring_old, void *ring_new, const void *img, int n_sub,
                     int n_pix, int hist, int cap, int prev_len, int n_new, int dtype,
                     dsx_stream_t stream);

/* f-fold time downsampling of (n_planes, n_in, n_inner) planes:
 * out[s][t][k] = sum_{j<f} in[s][f t + j][k], t < n_in / f (float32 sums,
 * dtype DSX_F16 or DSX_F32 for both).  Feeds the coarser-resolution image
 * histories of an uneven DM plan.
 * Same sum as the reference's float64 decimate(), src/dsart/x2/ref/corr2x_ref.py:234-237
 * (D consecutive samples -> one stream sample). */
int dsx_downsample_time(const void *in, void *out, int n_planes, int n_in, int64_t n_inner,
                        int factor, int dtype, dsx_stream_t stream);

/* ------------------------------------------------ dedispersion + detector
 * Layout of the LDS time windows for a DM plan (host function, no GPU work).
 * host_shift is [s][d] (host memory).  Picks the largest time tile (multiple
 * of 64 dividing n_centres, <= 2048) whose windows fit lds_budget bytes.
 * mode 0: pair rows when they fit, else unpaired; 1: pair rows only;
 * 2: unpaired only; 3: unpaired with two window buffers (float32 history
 * only): dsx_dedisp_detect copies the next tile's windows from HBM straight
 * into the second buffer (LDS-direct loads) while the current tile computes;
 * the plan marks the second buffer with q_total = r_total; 4: pair rows
 * built from one of two window buffers while the next tile's windows stream
 * into the other by LDS-direct loads (float32 history only; LDS holds the
 * rows, then 2 x r_total), row offsets padded so that the 64 DM lanes of a
 * wave reading one pair's rows at one step spread over the 64 LDS banks;
 * 5: pair rows as mode 1, stored as BF16 (the pair sums rounded to nearest
 * even; the reduced-precision partial-sum variant).  Outputs (host memory):
 *   sub_info [n_sub][3] = (window offset, max_d shift[s][d], window length)
 *   rows     [rows_cap][6] pair rows (Q offset, length, R index of subband a
 *            at e = 0, R index of subband b at e = 0, b's window [lo, hi))
 *   tbase    [n_terms][n_dm] LDS index of each term of D(t) at t = tile start
 *   info[8]  = (tile, n_rows, n_terms, r_total, q_total (4-byte words), paired,
 *              LDS bytes per work-group, DSX_DD_* flags)
 * LDS holds float32 subband windows and pair rows.  Pair rows: for subbands a = 2g, b = 2g + 1 and each distinct
 * delta = shift[b][d] - shift[a][d], Q[tau] = ring_a[tau] + ring_b[tau - delta]
 * over the span of steps the DMs with that delta read; D(t) is then the sum
 * of n_sub / 2 rows instead of n_sub subbands. */
int dsx_dedisp_plan(const int32_t *host_shift, int n_sub, int n_dm, int n_widths, int n_centres,
                    int ring_dtype, int lds_budget, int mode, int32_t *sub_info, int32_t *rows,
                    int rows_cap, int32_t *tbase, int32_t *info);

/* One block of centres [c_lo, c_lo + n_centres) (multiple of 64) for every
 * DM and pixel of the time-major ring (n_sub, n_pix, ring_len), where
 * position hist + tau - T0 holds absolute step tau:
 *   D[d][t][p] = sum_s ring[s][p][hist + (t - shift[s][d]) - T0]
 *   B_k(c)     = sum_{j<w_k} D[c - w_k/2 + j],   w_k = 2^k, k < n_widths
 *   snr_k      = (B_k - w_k mean[p][d]) / (sqrt(m2[p][d] / n_prev[d]) sqrt(w_k))
 * For each aligned chunk of 64 centres whose boxes are complete
 * (first centre - w_max/2 >= valid_from[d]) and each DM with
 * n_prev[d] >= min_count, the chunk maximum over widths (ties -> lowest k,
 * then earliest centre) is appended to cand (16-byte records
 * {int32 t - c_lo, int32 p, int16 d, int16 k, float snr}: the record's time
 * is relative to c_lo, so it fits int32 however long the stream runs) when
 * snr > threshold, counted by *n_cand (device int32 the caller zeroes;
 * entries past cand_cap are dropped but still counted; cand_cap 0 only
 * counts).  When merge != 0, samples t in [c_lo, c_lo + n_centres) with
 * t >= valid_from[d] feed the block's sums, merged (Chan) into the float64
 * running mean / m2 ([p][d] layout, n_prev[d] samples before) in place with
 * blk_count[d] samples; merge = 0 leaves the state unchanged, so a caller
 * whose records overflowed can detect the block again at a higher
 * threshold.  Samples are accumulated in float32 centred on the running
 * mean rounded to 11 significant bits (see dsx_dedisp_prime for a DM's
 * first block), so the sums keep their precision for pixels that hold a
 * steady source.  sub_info, rows, tbase, n_rows, n_terms, tile, r_total,
 * q_total and flags come from dsx_dedisp_plan (copied to device memory;
 * rows may be NULL when n_rows == 0; DSX_DD_DMA selects the LDS-direct
 * window loads of modes 3 and 4 and needs ring_dtype DSX_F32;
 * DSX_DD_ROWS_BF16 the BF16 pair rows of mode 5).
 * ring_dtype DSX_F16 or DSX_F32; hist >= 64 + w_max/2 + max shift.
 * Replaces combine_chgroups + boxcar_via_cumsum + multi_boxcar_argmax +
 * the Layer-1/Layer-2 normalisation, src/dsart/fine_dm/combiner.py:548-655,
 * src/dsart/detector/forward.py:163-504,1264-1854,
 * src/dsart/detector/triton_boxcar.py:193-403,
 * src/dsart/noise_norm/layer2.py:313-452. */
int dsx_dedisp_detect(const void *ring, int ring_dtype, int n_sub, int n_pix, int hist,
                      int ring_len, int64_t T0, const int32_t *sub_info, const int32_t *rows,
                      int n_rows, const int32_t *tbase, int n_terms, int n_dm,
                      const int32_t *valid_from, const double *n_prev, const double *blk_count,
                      double *mean, double *m2, int merge, int64_t c_lo, int n_centres,
                      int n_widths, int tile, int r_total, int q_total, int flags, float threshold,
                      double min_count, void *cand, int cand_cap, int32_t *n_cand,
                      dsx_stream_t stream);

/* Centre of each DM's first block: for every pixel p and DM d with
 * n_prev[d] == 0, mean[p][d] = the float64 average of its complete samples
 * D[d][t][p], max(valid_from[d], c_lo) <= t < c_lo + n_centres, from
 * per-subband prefix sums of the ring (same layout as dsx_dedisp_detect;
 * shift is the device [s][d] int32 table; the ring must hold those
 * samples: hist + c_lo - T0 >= max shift).  dsx_dedisp_detect then centres
 * the block's float32 sums on it; with n_prev[d] = 0 the merge gives it no
 * weight.  DMs with state are untouched.
 * Part of the Layer-2 burn-in the north-star statistics replace,
 * src/dsart/noise_norm/layer2.py:313-452 (Welford burn-in). */
int dsx_dedisp_prime(const void *ring, int ring_dtype, int n_sub, int n_pix, int hist, int ring_len,
                     int64_t T0, const int32_t *shift, int n_dm, const int32_t *valid_from,
                     const double *n_prev, int64_t c_lo, int n_centres, double *mean,
                     dsx_stream_t stream);

/* Exponential forgetting of the running statistics, applied after a block's
 * merge: m2[i] *= decay for i < n_m2 and n_prev[d] *= decay for d < n_dm
 * (0 < decay <= 1; the mean stays).  Each earlier block then weighs decay^age
 * in later merges, an EMA of the per-(pixel, DM) moments with
 * decay = exp(-block duration / tau) -- the forgetting of the reference's
 * Layer-2 sigma EMA, src/dsart/noise_norm/layer2.py:313-452
 * (gamma = 1 - exp(-cadence / tau)). */
int dsx_stats_decay(double *m2, int64_t n_m2, double *n_prev, int n_dm, double decay, dsx_stream_t stream);

/* ------------------------------------------------- detector over a cube
 * Time prefix of a search cube x (n_t, n_cell) (n_cell = F * H * W cells,
 * cube_dtype DSX_F16 / DSX_F32) into cs (n_t + 2h + 1, n_cell) float32:
 * cs[j][c] = sum_{t < j - h} x[t][c] (zeros before the cube, the total after
 * it), h = w_max / 2 -- the padded cumulative sum a boxcar bank reads.
 * Replaces precompute_padded_cumsum, src/dsart/detector/forward.py:250-336. */
int dsx_cube_cumsum(const void *cube, int cube_dtype, float *cs, int n_t, int64_t n_cell, int h,
                    dsx_stream_t stream);

/* Candidates of one cube from its prefix cs (dsx_cube_cumsum): for every
 * centre t in [t_base, t_base + n_out) and cell c, the centred boxcars
 * B_k = cs[t - w_k/2 + w_k + h][c] - cs[t - w_k/2 + h][c] (widths
 * host_widths[k], k < n_k <= 16, w_k/2 <= h), snr_k = B_k * host_sigma_inv[k]
 * (NULL: 1), the maximum over k (strict >, ties -> lowest k).  Cells whose
 * maximum beats threshold are queued in the workspace (256-byte counter head,
 * then 16-byte records; workspace_bytes sets the capacity); the
 * min(n_top, queued) strongest, ordered by (snr descending, flat index
 * t * n_cell + c ascending), are written to out as 16-byte records
 * {int32 t - t_base, int32 c, int32 k, float snr}, with their count in
 * *n_out_dev and the number of cells above threshold (kept or not) in
 * *n_above.  n_top <= 4096; n_out * n_cell < 2^32.  A count above the queue
 * capacity means the queue overflowed and the selection covers only the
 * queued cells.  Nothing is allocated; the counter is cleared by a kernel.
 * Replaces multi_boxcar_argmax_triton + decode_topk_argmax_lowmem's top-k,
 * src/dsart/detector/triton_boxcar.py:193-403,
 * src/dsart/detector/decoder.py:388-533 (the merge-radius suppression
 * follows on the host over the n_top records). */
int dsx_cube_candidates(const float *cs, int n_t, int64_t n_cell, int h, const int32_t *host_widths,
                        const float *host_sigma_inv, int n_k, float threshold, int t_base, int n_out,
                        int n_top, void *out, int32_t *n_out_dev, int32_t *n_above, void *workspace,
                        int64_t workspace_bytes, dsx_stream_t stream);

/* dsx_cube_candidates with no allocation and no workspace argument, for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  It has no
 * counterpart in the reference.  The counter and the queue live in a static
 * device array of this library, 16 MiB per GPU held until the process ends
 * (capacity dsx_cube_candidates_noalloc_capacity() = 2^20 queued cells;
 * *n_above beyond it reports an overflowed queue, as with a workspace of
 * that size); n_top beyond 4096 is rejected.  The counter is cleared by a
 * kernel at every call.  Every call on one GPU shares that GPU's array, so
 * calls that overlap on one GPU (concurrent streams or threads) race; the
 * caller serialises them.  Results are those of dsx_cube_candidates. */
int dsx_cube_candidates_noalloc(const float *cs, int n_t, int64_t n_cell, int h, const int32_t *host_widths,
                                const float *host_sigma_inv, int n_k, float threshold, int t_base, int n_out,
                                int n_top, void *out, int32_t *n_out_dev, int32_t *n_above,
                                dsx_stream_t stream);

/* Queue capacity (records) of dsx_cube_candidates_noalloc. */
int64_t dsx_cube_candidates_noalloc_capacity(void);

/* Raw boxcar sums of every width at n_s given cells (device int64 flat
 * indices (t - t_base) * n_cell + c of the centres, t in [t_base, t_base +
 * n_out)): out[k][i] (NaN for an index outside the centres).  The sample the
 * Layer-2 noise estimator clips, src/dsart/noise_norm/layer2.py:60-110. */
int dsx_cube_box_samples(const float *cs, int n_t, int64_t n_cell, int h, const int32_t *host_widths, int n_k,
                         int t_base, int n_out, const int64_t *cells, int n_s, float *out,
                         dsx_stream_t stream);

/* Shift-and-add combine of n_streams float32 streams (n_streams, t_stream,
 * n_inner): out[t][k] = sum_g streams[g][t0 + t - shift[g]][k] for
 * t < n_t, zero where t0 + t - shift[g] is outside [0, t_stream); shift is
 * a device int32 array.  Complex streams pass as 2 floats per cell.
 * Replaces combine_chgroups, src/dsart/fine_dm/combiner.py:548-655. */
int dsx_combine(const float *streams, const int32_t *shift, float *out, int n_streams,
                int t_stream, int64_t n_inner, int t0, int n_t, dsx_stream_t stream);

/* ------------------------------------------------------------- synthetic
 * Counter-based synthetic voltages for samples [sample0, sample0 + 2 n_pkt):
 * splitmix64(seed + i * 0x9E3779B97F4A7C15) of the flat component index,
 * top 16 bits through lut (65536 int32, 1/1024 units), plus burst phasors
 * (n_burst, n_ant, n_chan, 2) int32 for samples in
 * [burst_start[b][c], burst_start[b][c] + burst_len[b]), rounded half to
 * even, clipped to [-7, 7], packed (real low nibble) in on-wire layout.
 * Test-data generator; no reference counterpart on the path
 * (cf. src/dsart/inject/online.py for the reference's injector). */
int dsx_synth_block(uint8_t *raw, uint64_t seed, int64_t sample0, int n_pkt, int n_ant,
                    int n_chan, const int32_t *lut, int n_burst, const int64_t *burst_start,
                    const int32_t *burst_len, const int32_t *burst_phasor, dsx_stream_t stream);

/* ------------------------------------------------------ detector helpers
 * Centred boxcar along the leading axis of x (n_t, n_inner) float32:
 * out[i] = sum_{j<w} x[i - w/2 + j] with zeros outside.
 * Replaces boxcar_via_cumsum, src/dsart/detector/forward.py:163-247. */
int dsx_boxcar(const float *x, float *out, int n_t, int n_inner, int width,
               dsx_stream_t stream);

/* For each output cell (n_out, n_inner) compute K boxcar scores from a
 * padded cumulative sum cs (n_tpad, n_inner) (float32 or float16):
 * snr_k = (cs[t_base + t + off_k + w_k] - cs[t_base + t + off_k]) * sigma_inv_k,
 * keep the maximum (strict >, ties -> lowest k) and its k.
 * Replaces multi_boxcar_argmax_triton, src/dsart/detector/triton_boxcar.py:193-403. */
int dsx_multi_boxcar_argmax(const void *cs, int cs_dtype, const int32_t *widths,
                            const int32_t *offsets, const float *sigma_inv, int n_k,
                            void *out_max, int out_dtype, int16_t *out_win, int n_tpad,
                            int n_out, int n_inner, int t_base, dsx_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* DSX_H */
