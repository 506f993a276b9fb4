/*
 * bc250_stages.h - per-stage entry points of the gfx950 encoder pipeline.
 *
 * Each launcher enqueues one stage on `stream` (a hipStream_t, NULL for the
 * null stream) over device buffers in the layouts of gpu_compute.h and
 * returns 0, or -1 when the launch is rejected.  They replace the
 * reference's per-shader dispatches recorded by
 * gpu_compute_dispatch_encode_ext (approach1-compute-encoder/src/
 * gpu_compute.c:3078-3290); gpu_compute.hip calls them in that order, and
 * the parity tests call them one at a time.
 *
 * MV records are 4 x uint32 per block: {mvx, mvy, sad, pad} (quarter pel)
 * for refined vectors, {dx, dy, cost, sad} (whole pel) for the integer
 * search.  `*16` arrays hold one record per MB, `*8` arrays four per MB
 * (8x8 quadrants TL, TR, BL, BR).
 */
#ifndef BC250_STAGES_H
#define BC250_STAGES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* +-range integer full search, range in {16, 32, 64}, lambda8 in 0..15.
 * Replaces the motion_estimation.comp dispatch (gpu_compute.c:3093-3104),
 * integer stage (shaders/motion_estimation.comp:278-374). */
int bc250_launch_me_fullsearch(const uint8_t *cur, int cpitch, const uint8_t *ref, int rpitch, int wc, int hc,
                               int range, int lambda8, uint32_t *out16, uint32_t *out8, void *stream);
/* The same search restricted to the MBs list[0 .. *count), count read on the
 * device (the other MBs' records are left as they are). */
int bc250_launch_me_fullsearch_list(const uint8_t *cur, int cpitch, const uint8_t *ref, int rpitch, int wc, int hc,
                                    int range, int lambda8, uint32_t *out16, uint32_t *out8, const int *list,
                                    const int *count, void *stream);
/* The same records for a short list: each listed MB's candidate rows are
 * split over several workgroups whose keys meet in keys (5 * MBs of the
 * picture uint32 of work space, set on the stream first unless keys_ready:
 * the learned search's fallback list sets them itself), so a few hundred
 * MBs still fill the GPU. */
int bc250_launch_me_fullsearch_list_split(const uint8_t *cur, int cpitch, const uint8_t *ref, int rpitch, int wc,
                                          int hc, int range, int lambda8, uint32_t *out16, uint32_t *out8,
                                          const int *list, const int *count, uint32_t *keys, int keys_ready,
                                          void *stream);

/* Half- then quarter-pel refinement of the integer winners
 * (shaders/motion_estimation.comp:376-469). do8 = also the 8x8 quadrants. */
int bc250_launch_me_refine(const uint8_t *cur, int cpitch, const uint8_t *ref, int rpitch, int wc, int hc,
                           const uint32_t *in16, const uint32_t *in8, uint32_t *mv16, uint32_t *mv8, int do8,
                           void *stream);

/* P16x16 encode: MC, residual, 4x4 transform, inter quant, reconstruction.
 * Replaces residual_predict -> dct_transform -> quantize -> reconstruct
 * (gpu_compute.c:3109-3190). `o*` must not alias `r*`. */
int bc250_launch_p_encode(const uint8_t *cy, const uint8_t *cuv, int cpy, int cpuv, const uint8_t *ry,
                          const uint8_t *ruv, int rpy, int rpuv, uint8_t *oy, uint8_t *ouv, int opy, int opuv,
                          int wc, int hc, int qp, const uint32_t *mvs, int16_t *levels, int32_t *dc, uint32_t *nz,
                          uint32_t *modes, void *stream);

/* I16x16 wavefront in one persistent launch; replaces the 187 per-diagonal
 * intra_wavefront.comp dispatches (gpu_compute.c:3191-3255).  `sync` is a
 * workspace of bc250_intra_work_bytes() bytes whose first
 * 2 + coded_height/16 ints must be zero; sync[1] != 0 afterwards means a
 * dependency wait gave up after spin_limit polls. */
size_t bc250_intra_work_bytes(int wc, int hc);
int bc250_launch_i_encode(const uint8_t *cy, const uint8_t *cuv, int cpy, int cpuv, uint8_t *oy, uint8_t *ouv,
                          int opy, int opuv, int wc, int hc, int qp, int num_slices, int16_t *levels, int32_t *dc,
                          uint32_t *nz, uint32_t *modes, int *sync, unsigned spin_limit, void *stream);

/* H.264 8.7 deblocking in place (luma + chroma): a boundary-strength launch
 * and one persistent wavefront launch; replaces the two deblock_filter.comp
 * passes (gpu_compute.c:3257-3290).  `sync` is a workspace of
 * bc250_deblock_work_bytes() bytes whose first 2 ints must be zero (the
 * rest is cleared by the launch itself); sync[1] != 0 afterwards means a
 * dependency wait gave up after spin_limit polls. */
size_t bc250_deblock_work_bytes(int wc, int hc);
int bc250_launch_deblock(uint8_t *y, uint8_t *uv, int py, int puv, int wc, int hc, int qp, int is_intra,
                         const uint32_t *nz, const uint32_t *mvs, int num_slices, int idc, int *sync,
                         unsigned spin_limit, void *stream);

/* Intra 4x4 / 8x8 luma prediction (H.264 8.3.1.2, 8.3.2.2 with the 8x8
 * reference filter), batched over n blocks; the per-block functions they
 * replace are the reference decoder's h264d_pred4x4 / h264d_pred8x8_luma
 * (h264_pred.h:58-65).  The reference encoder codes only I16x16; these serve
 * Intra4x4 / Intra8x8 macroblocks.  Inputs per block:
 *   4x4: nbr[16*b] = p[-1,0..3], p[-1,-1], p[0..7,-1] (3 bytes unused)
 *   8x8: nbr[32*b] = p[-1,0..7], p[-1,-1], p[0..15,-1] (7 bytes unused)
 *   ctl[b] = mode (0..8: V, H, DC, DDL, DDR, VR, HD, VL, HU) | top<<4 |
 *            left<<5 | corner<<6 | top-right<<7 (availability)
 * out: 16 (64) samples per block, row-major. */
int bc250_h264_pred4x4_batch(const uint8_t *nbr, const uint8_t *ctl, int n, uint8_t *out, void *stream);
int bc250_h264_pred8x8_batch(const uint8_t *nbr, const uint8_t *ctl, int n, uint8_t *out, void *stream);

/* 8x8 transform (High profile), batched over n blocks of 64 int16:
 * residual -> levels (forward Y = C X C^T and a quantiser that is the exact
 * inverse of the decoder's scaling, rounding round_q12/12 of a step: 4 for
 * intra, 2 for inter as the 4x4 path's f) and levels -> residual (8.5.13.1
 * flat-scaling dequant + 8.5.13.2 inverse, (x+32)>>6), the decoder's
 * scale8 + h264d_idct8_add (h264_recon.c:87-93,192-210).  The reference
 * encoder codes only 4x4 transforms; these serve transform_size_8x8 MBs. */
int bc250_h264_tq8x8_batch(const int16_t *res, int n, int qp, int round_q12, int16_t *out, void *stream);
int bc250_h264_dqit8x8_batch(const int16_t *lv, int n, int qp, int16_t *out, void *stream);

/* Zero `n` ints with a fill kernel (no memset node, so capturable). */
int bc250_launch_zero_i32(int *p, int n, void *stream);

/* Replicate the last visible row / column of a plane pair into the coded
 * padding (w x h visible, wc x hc coded). */
int bc250_launch_pad_edges(uint8_t *y, uint8_t *uv, int py, int puv, int w, int h, int wc, int hc, void *stream);

/* The reference's Tier-2 motion search (src/cpu_simd_me.c:275-497,
 * cpu_simd_me_search_frame) on the GPU, bit-exact: per MB in raster order the
 * zero-vector SAD (<= 512 keeps it), the left neighbour's vector / 4 as a
 * predictor (taken when it costs < 768), the 8-point square search with steps
 * R/2 .. 1 and cost SAD + 5|mv|, then four half-sample candidates (rounded-up
 * two-sample averages) when 128 < cost < 3000.  R = search_radius clamped to
 * 2..16.  Planes MB-padded (pitch >= 16 * ceil(width / 16)); out_mvs:
 * gpu_mv_t per MB {mvx, mvy (quarter samples), cost, 0}.  One wave per MB
 * row (the rows are independent, MBs within a row are not). */
int bc250_launch_square_me(const uint8_t *cur, int cpitch, const uint8_t *ref, int rpitch, int width, int height,
                           int search_radius, uint32_t *out_mvs, void *stream);

/* The reference's GPU motion search (shaders/motion_estimation.comp:
 * 278-469) with exact integer SADs in place of its float-normalised ones:
 * zero-vector SAD (<= 512 keeps it), diamond (0,+-1) (+-1,0) at steps 8, 4,
 * 2, 1 within +-16 with cost 2 x even-parity SAD + lambda |mv| (stop after a
 * step below 768), then 8 half- and 8 quarter-sample candidates by plain SAD
 * of the 8.4.2.2.1 samples.  lambda 5, or 8 in the reference's fast mode.
 * Coded-size planes (edges replicated); out_mvs: gpu_mv_t per MB. */
int bc250_launch_diamond_me(const uint8_t *cur, int cpitch, const uint8_t *ref, int rpitch, int wc, int hc,
                            int lambda, uint32_t *out_mvs, void *stream);

#ifdef __cplusplus
}
#endif
#endif
