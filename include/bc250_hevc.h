/*
 * bc250_hevc.h - HEVC stages of the encoder compute pipeline on gfx950.
 *
 * The reference codes HEVC on the CPU (enc/src/hevc_enc_template.c with
 * enc/src/hevc_intra.c and the decoder's enc/src/hevc_mc_template.c); it has
 * no GPU interface for it.  These entry points move those per-block
 * functions onto the GPU as batched launches with the same argument meaning
 * and results, plus frame-level launches that code a whole picture:
 *
 *   per-block functions (batched)            reference function replaced
 *   bc250_hevc_tq4_batch                     hevc_transform_quant_4x4[_10]  (hevc_intra.h:132,155)
 *   bc250_hevc_dqit4_batch                   hevc_dequant_itransform_4x4[_10] (hevc_intra.h:139,157)
 *   bc250_hevc_tq8_batch                     hevc_transform_quant_8x8        (hevc_intra.h:162)
 *   bc250_hevc_dqit8_batch                   hevc_dequant_itransform_8x8     (hevc_intra.h:164)
 *   bc250_hevc_mc_uni_batch                  hevc_mc_uni_8 / hevc_mc_uni_10  (hevc_inter.h:26-31)
 *
 *   picture-level launches
 *   bc250_hevc_launch_refine8                quarter-sample refinement of the +-R full-search
 *                                            8x8 vectors with the HEVC interpolation
 *   bc250_hevc_launch_p_encode               predict_cu + inter_residual with the 8x8 luma
 *                                            transform (hevc_enc_template.c:669-680,817-933)
 *   bc250_hevc_intra_work_bytes /
 *   bc250_hevc_launch_i_encode               encode_ctu's I branch + intra_trial
 *                                            (hevc_enc_template.c:502-594,1409-1427)
 *   bc250_hevc_launch_i_encode_ctu /         the same two at CTU32 (no reference
 *   bc250_hevc_launch_deblock_ctu            counterpart)
 *
 * Picture layout: coded size (16-aligned), 8-bit, NV12 (luma plane, then
 * interleaved Cb/Cr), the same surfaces as the H.264 path.  CTUs are 16x16
 * in raster order, each four 8x8 CUs in z-order (CU index = ctu*4 + z).
 * Coefficient output per CTU, int16 [384]:
 *   I picture: 16 luma PUs (z-order) x 16, then 4 Cb CUs x 16, then 4 Cr x 16
 *   P picture: 4 luma CUs x 64 (8x8 row-major), then 4 Cb x 16, then 4 Cr x 16
 * Each block is row-major; levels as the reference's quantiser returns them.
 * All launchers return 0, or -1 on an invalid argument or launch failure.
 */
#ifndef BC250_HEVC_H
#define BC250_HEVC_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* n blocks of 16 (or 64) int16 each; qp is Qp' (QP + 12 at ten bits),
 * bd 8 or 10; round_q12 the rounding offset in twelfths of a step. */
int bc250_hevc_tq4_batch(const int16_t *res, int n, int qp, int use_dst, int round_q12, int bd, int16_t *out,
                         void *stream);
int bc250_hevc_dqit4_batch(const int16_t *coef, int n, int qp, int use_dst, int bd, int16_t *out, void *stream);
int bc250_hevc_tq8_batch(const int16_t *res, int n, int qp, int bd, int round_q12, int16_t *out, void *stream);
int bc250_hevc_dqit8_batch(const int16_t *coef, int n, int qp, int bd, int16_t *out, void *stream);

/* n blocks of N x N (N = 1 << log2n, 4..32) DCT-II: forward + quant and
 * dequant + inverse.  The inverse is the decoder's (hevcd_transform,
 * hevc_transform.c:274, with 8.6.2 flat-list scaling); the forward is the
 * encoder's choice and generalises the reference's 4x4 / 8x8 shifts:
 * (log2n + bd - 9, log2n + 6), quantiser shift bd + log2n - 5.  The
 * reference encoder codes no TU above 8x8; these serve 16x16 / 32x32 TUs. */
int bc250_hevc_tqn_batch(const int16_t *res, int n, int log2n, int qp, int bd, int round_q12, int16_t *out,
                         void *stream);
int bc250_hevc_dqitn_batch(const int16_t *coef, int n, int log2n, int qp, int bd, int16_t *out, void *stream);

/* The 32x32 pair of bc250_hevc_tqn_batch / bc250_hevc_dqitn_batch (log2n 5,
 * identical results) in a chosen form: BC250_T32_BUTTERFLY, register partial
 * butterflies on the vector ALUs, BC250_T32_MFMA, an exact split-int8
 * product on the matrix cores (int16 operand = 256 * signed high byte +
 * low byte, int32 accumulation), or BC250_T32_MATRIX.  No reference
 * counterpart: the reference codes no TU above 8x8; the inverse is the
 * decoder's hevcd_transform (hevc_transform.c:274) at nTbS 32. */
#define BC250_T32_BUTTERFLY 0
#define BC250_T32_MFMA 1
/* The plain matrix product of the N-point batches, kept for comparison. */
#define BC250_T32_MATRIX 2
/* The form bc250_hevc_tqn_batch / bc250_hevc_dqitn_batch use at log2n 5. */
#define BC250_T32_DEFAULT BC250_T32_MFMA
int bc250_hevc_tq32_batch(const int16_t *res, int n, int qp, int bd, int round_q12, int method, int16_t *out,
                          void *stream);
int bc250_hevc_dqit32_batch(const int16_t *coef, int n, int qp, int bd, int method, int16_t *out, void *stream);

/* n blocks of w x h (w, h <= 16) from one plane (uint8 when bd == 8, else
 * uint16; stride in samples): blk[4*i .. 4*i+3] = {x, y, mvx, mvy}, vectors
 * in quarter (luma) or eighth (chroma) samples.  dst: n * w * h samples. */
int bc250_hevc_mc_uni_batch(const void *ref, int stride, int w_pic, int h_pic, const int32_t *blk, int n, int w,
                            int h, int chroma, int bd, void *dst, void *stream);

/* Frame-stage entry points.  Planes are 8-bit NV12 (bd 8) or P010 (bd 10:
 * the sample in the high 10 bits of each 16-bit word); pitches in bytes. */

/* 8-bit luma for the integer search of a P010 picture: dst = the top 8 bits
 * of each 10-bit sample (w a multiple of 4). */
int bc250_hevc_p010_to_8bit(const void *src, int spitch, uint8_t *dst, int dpitch, int w, int h, void *stream);

/* Quarter-sample refinement of every 8x8 CU: imv8[4*cu] = {dx, dy, cost,
 * sad} integer vectors (bc250_launch_me_fullsearch's 8x8 output); out[4*cu]
 * = {mvx, mvy (quarter samples), sad, 0}.  Centre, then 8 half-sample
 * candidates (2,0) (-2,0) (0,2) (0,-2) (2,2) (2,-2) (-2,2) (-2,-2), then the
 * same offsets halved around the half winner, strictly lower SAD wins; every
 * candidate interpolated with the HEVC 8-tap filter at the bit depth. */
int bc250_hevc_launch_refine8(const void *cur, int cpitch, const void *ref, int rpitch, int wc, int hc,
                              const uint32_t *imv8, uint32_t *out, int bd, void *stream);

/* P picture: every CU inter with mvs[4*cu] = {mvx, mvy, ...}; luma 8x8 DCT,
 * chroma 4x4 DCT at QpC, rounding round_q12 (2 in the reference's default
 * dead-zone mode), QP' = qp + 6 (bd - 8).  Writes coef and the
 * reconstruction (o*) (inter_residual, hevc_enc_template.c:817-933). */
int bc250_hevc_launch_p_encode(const void *cy, const void *cuv, int cpy, int cpuv, const void *ry, const void *ruv,
                               int rpy, int rpuv, void *oy, void *ouv, int opy, int opuv, int wc, int hc, int qp,
                               int round_q12, const uint32_t *mvs, int16_t *coef, int bd, void *stream);

/* The same with the coding tree unit size 1 << ctu_log2 (no reference
 * counterpart beyond 4, which is the call above).  At CTU32 (5) each CTU32
 * chooses its CUs from the vectors: a 16x16 area whose four CU8 vectors are
 * equal is one CU16, a CTU32 inside the picture whose four areas are CU16s
 * with one vector is one CU32; a CU16 / CU32 codes one luma TU of its size
 * and one chroma TU of half its size per component (the N-point transforms
 * of bc250_hevc_tqn_batch / dqitn).  Levels stay in the per-16x16 layout:
 * a CU16's area holds its 16x16 luma TU (256, row-major) and 8x8 Cb / Cr
 * TUs (64 each); area q (z-order) of a CU32 holds levels [256 q, 256 q +
 * 256) of its 32x32 luma TU and [64 q, 64 q + 64) of each 16x16 chroma TU.
 * culog2 (nctu * 4 bytes; required at CTU32, may be NULL at CTU16) = log2
 * of the CU holding each CU8 (3, 4 or 5).  modes (may be NULL, nctu * 16
 * bytes): zero, then at CTU32 the CU sizes at [16 t + z].  cbf (may be NULL,
 * nctu * 4 bytes): the coded-luma flag of each CU8's 8x8 levels (the
 * deblocking input at CTU16; at CTU32 use bc250_hevc_cu_cbf_ctu). */
int bc250_hevc_launch_p_encode_ctu(const void *cy, const void *cuv, int cpy, int cpuv, const void *ry,
                                   const void *ruv, int rpy, int rpuv, void *oy, void *ouv, int opy, int opuv, int wc,
                                   int hc, int qp, int round_q12, const uint32_t *mvs, int16_t *coef, int ctu_log2,
                                   uint8_t *culog2, uint8_t *modes, uint8_t *cbf, int bd, void *stream);

/* bc250_hevc_launch_refine8 and bc250_hevc_launch_p_encode_ctu in one pass
 * (identical results): the wave that refines a CU8's vector (imv8 -> mv8,
 * the refine8 records) codes the CU with it, on the same planes (cy / ry at
 * the full bit depth).  No reference counterpart: the fusion saves the
 * launch and the vector round trip between the two stages. */
int bc250_hevc_launch_refine_p_encode(const void *cy, const void *cuv, int cpy, int cpuv, const void *ry,
                                      const void *ruv, int rpy, int rpuv, void *oy, void *ouv, int opy, int opuv,
                                      int wc, int hc, int qp, int round_q12, const uint32_t *imv8, uint32_t *mv8,
                                      int16_t *coef, int ctu_log2, uint8_t *culog2, uint8_t *modes, uint8_t *cbf,
                                      int bd, void *stream);

/* bc250_hevc_launch_refine_p_encode at CTU16 that also packs the levels for
 * a readback: each 16-level group of coef (24 per CTU16: 16 luma, 4 Cb, 4
 * Cr) holding a nonzero level goes to pack_data[k][16] with its group index
 * in pack_idx[k], k taken from *pack_count (which must start at zero) by one
 * atomic add per CU; the order of the groups is not fixed.  No reference
 * counterpart. */
int bc250_hevc_launch_refine_p_encode_pack(const void *cy, const void *cuv, int cpy, int cpuv, const void *ry,
                                           const void *ruv, int rpy, int rpuv, void *oy, void *ouv, int opy, int opuv,
                                           int wc, int hc, int qp, int round_q12, const uint32_t *imv8, uint32_t *mv8,
                                           int16_t *coef, uint8_t *modes, uint8_t *cbf, int bd, uint32_t *pack_idx,
                                           int16_t *pack_data, int *pack_count, void *stream);
/* The same with the learned search's full-search keys (bc250_learned_search_
 * picture_keys: 5 per MB, the 16x16 key set only for MBs searched in full):
 * such an MB's 8x8 records come from its keys instead of imv8, and its 16x16
 * record is written to out16 (the next picture's proposer input). */
int bc250_hevc_launch_refine_p_encode_pack_keys(const void *cy, const void *cuv, int cpy, int cpuv, const void *ry,
                                                const void *ruv, int rpy, int rpuv, void *oy, void *ouv, int opy,
                                                int opuv, int wc, int hc, int qp, int round_q12, const uint32_t *imv8,
                                                uint32_t *mv8, int16_t *coef, uint8_t *modes, uint8_t *cbf, int bd,
                                                uint32_t *pack_idx, int16_t *pack_data, int *pack_count,
                                                const uint32_t *fs_keys, uint32_t *out16, int range, int lambda8,
                                                void *stream);

/* I picture: CTU16 -> 4 CU8 -> 4 PU4 intra (Planar/DC/H/V by SAD, DST),
 * chroma DC per CU (DCT, QpC), slices as CTU-row ranges
 * [floor(rows*s/ns), floor(rows*(s+1)/ns)), QP' = qp + 6 (bd - 8) (the
 * reference's encode_ctu -> intra_trial, hevc_enc_template.c:502-594).  One
 * persistent wavefront launch; `sync` is a workspace of
 * bc250_hevc_intra_work_bytes() bytes whose first 2 ints must be zero;
 * sync[1] != 0 afterwards means a dependency wait gave up after spin_limit
 * polls.  modes: uint8 [nctu*16] (0 Planar, 1 DC, 10 H, 26 V), PU z-order.
 * Pitches multiples of 4 (bd 8) or 8 (bd 10) bytes. */
size_t bc250_hevc_intra_work_bytes(int wc, int hc);
int bc250_hevc_launch_i_encode(const void *cy, const void *cuv, int cpy, int cpuv, void *oy, void *ouv, int opy,
                               int opuv, int wc, int hc, int qp, int num_slices, int round_q12, uint8_t *modes,
                               int16_t *coef, int *sync, unsigned spin_limit, int bd, void *stream);

/* The same with the coding tree unit size 1 << ctu_log2: 4 is the call
 * above; 5 codes CTU32s (no reference counterpart: the reference encoder
 * codes CTU16 only): the four 16x16 quadrants of a CTU32 in z-order, each
 * as above, neighbours available by the CTU32 z-scan order (6.5.2: a
 * bottom-right quadrant's top-right neighbour is not yet coded, the next
 * CTU32's top-left quadrant has its bottom-left neighbour), slices as
 * CTU32-row ranges; outputs keep the per-16x16 layout.  Picture sizes as
 * above (partial CTU32s at the right and bottom edges code the quadrants
 * inside the picture). */
int bc250_hevc_launch_i_encode_ctu(const void *cy, const void *cuv, int cpy, int cpuv, void *oy, void *ouv, int opy,
                                   int opuv, int wc, int hc, int qp, int num_slices, int round_q12, uint8_t *modes,
                                   int16_t *coef, int *sync, unsigned spin_limit, int ctu_log2, int bd,
                                   void *stream);

/* Per-CU coded-luma flags of a P picture's levels (bc250_hevc_launch_p_encode
 * layout): cbf[4*t + z] = 1 when any of CU z's 64 luma levels is non-zero,
 * the bS 1 condition of the deblocking filter. */
int bc250_hevc_cu_cbf(const int16_t *coef, int nctu, uint8_t *cbf, void *stream);
/* The same for a P picture of bc250_hevc_launch_p_encode_ctu (coded size
 * wc x hc): cbf[4 t + z] = whether the transform block holding CU8 z of area
 * t (8x8, 16x16 or 32x32 luma, culog2[4 t + z] = 3, 4 or 5) has a non-zero
 * level; culog2 NULL is bc250_hevc_cu_cbf. */
int bc250_hevc_cu_cbf_ctu(const int16_t *coef, int wc, int hc, const uint8_t *culog2, uint8_t *cbf, void *stream);

/* H.265 8.7.2 deblocking of a reconstructed picture in place (wc, hc
 * multiples of 16; pitches in bytes, multiples of 8): bd 8 on NV12 planes,
 * bd 10 on P010 planes (the sample in the high 10 bits of each 16-bit word),
 * for the pipeline's coding shape: 8x8 CUs in CTU-raster, z-order (cu =
 * ctu*4 + (y/8&1)*2 + (x/8&1)), one PU each.  Vertical edges of the 8x8 grid
 * first, then horizontal edges; picture and slice borders (CTU-row slices as
 * in i_encode) are not filtered.  bS 2 when is_intra, else 1 if cbf[cp] |
 * cbf[cq] or the vectors mvs[4*cu] = {mvx, mvy, ...} differ by >= 4 quarter
 * samples, else 0 (approach1-compute-encoder/src/hevc_filter_template.c:
 * 221-290, the reference decoder); chroma on the 8x8 chroma grid at bS 2;
 * beta and tC scaled by 2^(bd-8).  The reference encoder's PPS disables this
 * filter (src/encoder_h265.c:330), so the pipeline applies it only when
 * asked. */
int bc250_hevc_launch_deblock(void *y, void *uv, int py, int puv, int wc, int hc, int qp, int is_intra,
                              const uint8_t *cbf, const uint32_t *mvs, int num_slices, int bd, void *stream);
/* The same with slices as ranges of CTU rows of 1 << ctu_log2 (4 or 5)
 * luma rows (the CTU32 pictures of bc250_hevc_launch_i_encode_ctu), and,
 * when culog2 (per CU8, see bc250_hevc_launch_p_encode_ctu) is not NULL, no
 * filtering of the 8x8-grid edges inside a CU16 / CU32 (they are neither
 * TU nor PU edges; cbf then per transform block: bc250_hevc_cu_cbf_ctu). */
int bc250_hevc_launch_deblock_ctu(void *y, void *uv, int py, int puv, int wc, int hc, int qp, int is_intra,
                                  const uint8_t *cbf, const uint32_t *mvs, const uint8_t *culog2, int num_slices,
                                  int ctu_log2, int bd, void *stream);

#ifdef __cplusplus
}
#endif
#endif
