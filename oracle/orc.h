/*
 * orc.h - CPU oracle for the encoder compute pipeline (TEST INFRASTRUCTURE).
 *
 * This library is the checker, never the product: only tests/, the smoke()
 * entry and bench.py's cpu_baseline leg may load it. It restates, in plain
 * C, the semantics of the reference's compute stages (cited per function in
 * the .c files) on the coded-size frame layout used by the HIP path:
 *
 *   luma   : wc x hc bytes, wc = 16*mb_w, hc = 16*mb_h, row pitch `py`
 *   chroma : NV12 interleaved Cb/Cr, wc bytes x hc/2 rows, pitch `puv`
 *   levels : int16 [num_mbs*24*16], block b of MB m at (m*24+b)*16,
 *            blocks 0..15 luma raster (by*4+bx), 16..19 Cb, 20..23 Cr,
 *            position p = row*4+col inside a block
 *   dc     : int32 [num_mbs*24]   pre-quant coefficient 0 of each block
 *   nz     : uint32 [num_mbs*24]  bit p set iff levels[p] != 0
 *   modes  : uint32 [num_mbs]     bits 0-1 I16x16 mode, bits 2-3 chroma mode
 *   mvs    : orc_mv_t [num_mbs]   quarter-pel MV + SAD (16 bytes each)
 */
#ifndef ORC_H
#define ORC_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { int32_t mvx, mvy; uint32_t sad, pad; } orc_mv_t;

/* Integer full-search result for one block: MV in whole pels + its cost. */
typedef struct { int32_t dx, dy; uint32_t cost, sad; } orc_imv_t;

int  orc_chroma_qp(int qpy);

/* --- 4x4 transform / quant primitives (H.264 8.5) ------------------------ */
void orc_fwd4x4(const int32_t in[16], int32_t out[16]);
void orc_inv4x4(const int32_t in[16], int32_t out[16]);
int  orc_quant_ac(int c, int pos, int qp, int is_intra);
int  orc_dequant_ac(int level, int pos, int qp);
void orc_luma_dc_hadamard_fwd(const int32_t in[16], int32_t out[16]);
void orc_luma_dc_hadamard_inv(const int32_t in[16], int32_t out[16]);
void orc_chroma_dc_hadamard(const int32_t in[4], int32_t out[4]);
int  orc_quant_dc_luma(int v, int qp);
int  orc_quant_dc_chroma(int v, int qpc);
int  orc_dequant_dc_luma(int f, int qp);
int  orc_dequant_dc_chroma(int f, int qpc);
void orc_intra16_luma_dc_levels(const int32_t dc_in[16], int qp, int32_t out[16]);

/* --- interpolation (H.264 8.4.2.2) --------------------------------------- */
int  orc_luma_qpel(const uint8_t *ref, int py, int wc, int hc, int xq, int yq);
int  orc_chroma_8th(const uint8_t *uv, int puv, int wcc, int hcc, int comp,
                    int x8, int y8);
void orc_pred_luma_block(const uint8_t *ref, int py, int wc, int hc,
                         int x0, int y0, int w, int h, int mvx, int mvy,
                         uint8_t *dst, int dpitch);

/* --- motion estimation ----------------------------------------------------- */
/* the reference's Tier-2 square search (cpu_simd_me_search_frame); out: 4 int32 per MB */
void orc_square_me(const uint8_t *src, int sp, const uint8_t *ref, int rp, int width, int height, int search_radius,
                   int32_t *out);
/* the reference's diamond search (motion_estimation.comp) with integer SADs; out: 4 int32 per MB */
void orc_diamond_me(const uint8_t *cur, const uint8_t *ref, int py, int wc, int hc, int lambda, int32_t *out);
void orc_me_fullsearch(const uint8_t *cur, const uint8_t *ref, int py,
                       int wc, int hc, int range, int lambda8,
                       orc_imv_t *int16, orc_imv_t *int8);
void orc_me_fullsearch_mb(const uint8_t *cur, const uint8_t *ref, int py,
                          int wc, int hc, int range, int lambda8, int mbx,
                          int mby, orc_imv_t *int16, orc_imv_t *int8);
void orc_me_refine(const uint8_t *cur, const uint8_t *ref, int py, int wc,
                   int hc, const orc_imv_t *int16, const orc_imv_t *int8,
                   orc_mv_t *mv16, orc_mv_t *mv8);
void orc_me_refine_mb(const uint8_t *cur, const uint8_t *ref, int py, int wc, int hc, int mbx, int mby,
                      const orc_imv_t *int16, const orc_imv_t *int8, orc_mv_t *mv16, orc_mv_t *mv8);

/* --- macroblock encode ----------------------------------------------------- */
void orc_h264_encode_p(const uint8_t *cy, const uint8_t *cuv,
                       const uint8_t *ry, const uint8_t *ruv, int py, int puv,
                       int wc, int hc, int qp, const orc_mv_t *mvs,
                       int16_t *levels, int32_t *dc, uint32_t *nz,
                       uint32_t *modes, uint8_t *oy, uint8_t *ouv);
void orc_h264_encode_i(const uint8_t *cy, const uint8_t *cuv, int py, int puv,
                       int wc, int hc, int qp, int num_slices,
                       int16_t *levels, int32_t *dc, uint32_t *nz,
                       uint32_t *modes, uint8_t *oy, uint8_t *ouv);
int  orc_slice_of(int mb, int total_mbs, int num_slices);
void orc_pred16x16(int mode, const int top[16], const int left[16], int corner, int ta, int la,
                   uint8_t out[256]);
void orc_pred_chroma_dc(const int top[8], const int left[8], int ta, int la, uint8_t out[64]);

/* --- deblocking (H.264 8.7) ------------------------------------------------ */
void orc_h264_deblock(uint8_t *y, uint8_t *uv, int py, int puv, int wc, int hc,
                      int qp, int is_intra, const uint32_t *nz,
                      const orc_mv_t *mvs, int num_slices, int disable_idc);
void orc_h264_filter_line_luma(uint8_t *q0, int step, int bs, int qp);
void orc_h264_filter_line_chroma(uint8_t *q0, int step, int bs, int qpc);
int  orc_h264_alpha(int index_a);
int  orc_h264_beta(int index_b);
int  orc_h264_tc0(int index_a, int bs);

/* --- intra 4x4 / 8x8 luma prediction (H.264 8.3.1.2, 8.3.2.2) ---------------- */
void orc_h264_pred4x4(const uint8_t top8[8], const uint8_t left4[4], int corner, int mode, int avail_top,
                      int avail_left, uint8_t out[16]);
void orc_h264_pred8x8(const uint8_t top16[16], const uint8_t left8[8], int corner, int mode, int avail_top,
                      int avail_left, int avail_corner, int avail_top_right, uint8_t out[64]);

/* --- 8x8 transform (H.264 8.5.13; forward + quantiser inverting it) ------- */
int  orc_h264_class8(int i, int j);
int  orc_h264_v8(int qp_rem, int cls);
void orc_h264_fwd8x8(const int32_t in[64], int64_t out[64]);
void orc_h264_quant8x8(const int64_t y[64], int qp, int round_q12, int16_t out[64]);
void orc_h264_dequant8x8(const int16_t c[64], int qp, int32_t d[64]);
void orc_h264_inv8x8(const int32_t d[64], int32_t out[64]);
void orc_h264_tq8x8(const int16_t res[64], int qp, int round_q12, int16_t out[64]);
void orc_h264_dqit8x8(const int16_t lv[64], int qp, int16_t out[64]);

/* --- whole frame (stage order of the reference's dispatch) ----------------- */
void orc_h264_encode_frame(const uint8_t *cy, const uint8_t *cuv,
                           const uint8_t *ry, const uint8_t *ruv, int py,
                           int puv, int wc, int hc, int qp, int is_intra,
                           int num_slices, int range, int lambda8,
                           int deblock_idc, orc_mv_t *mvs, orc_mv_t *mv8,
                           int16_t *levels, int32_t *dc, uint32_t *nz,
                           uint32_t *modes, uint8_t *oy, uint8_t *ouv);

/* H.264 CAVLC (orc_cavlc.c): one residual block (raster coefficients,
 * zig-zag scanned from position `ac`), a chroma DC block, Exp-Golomb codes,
 * the two motion vector predictors and the whole-frame slice writer. */
int  orc_cavlc_block(const int16_t coef[16], int ac, int nc, uint8_t *buf, int cap, int *nbits);
int  orc_cavlc_chroma_dc(const int32_t c[4], uint8_t *buf, int cap, int *nbits);
int  orc_exp_golomb(int value, int is_signed, uint8_t *buf, int cap, int *nbits);
int  orc_h264_mb_header(int kind, int a, int b, int c, int d, uint8_t *buf, int cap, int *nbits);
void orc_h264_mv_pred(const orc_mv_t *mv, int wmb, int start, int mb, int *px, int *py);
void orc_h264_skip_mv(const orc_mv_t *mv, int wmb, int start, int mb, int *px, int *py);
int64_t orc_h264_cavlc_slices(const int16_t *levels, const int32_t *dc, const uint32_t *modes,
                              const orc_mv_t *mv, const int32_t *cfg, uint8_t *out, int64_t cap, int64_t *off,
                              int64_t *len);

/* --- HEVC stages (orc_hevc.c) --------------------------------------------- */
int  orc_hevc_chroma_qp(int qp_luma);
void orc_hevc_quant(const int32_t *v, int n, int shift, int q, int round_q12, int16_t *out);
void orc_hevc_dequant(const int16_t *c, int n, int q, int bd_shift, int16_t *out);
void orc_hevc_tq4(const int16_t res[16], int qp, int use_dst, int round_q12, int bd, int16_t out[16]);
void orc_hevc_dqit4(const int16_t coef[16], int qp, int use_dst, int bd, int16_t out[16]);
void orc_hevc_tq8(const int16_t res[64], int qp, int bd, int round_q12, int16_t out[64]);
void orc_hevc_dqit8(const int16_t coef[64], int qp, int bd, int16_t out[64]);
void orc_hevc_mc_uni(const void *ref, int stride, int w_pic, int h_pic, int x, int y, int w, int h,
                     int mvx, int mvy, int chroma, int bd, void *dst, int dst_stride);
void orc_hevc_predict4(const void *plane, int bd, int stride, int width, int height, int x0, int y0,
                       int mode, int is_luma, int y_min, int32_t pred[16]);
int  orc_hevc_choose_mode(const void *src, const void *recon, int bd, int stride, int width, int height,
                          int x0, int y0, int y_min, int32_t pred[16]);
int  orc_hevc_slice_y_min(int ctu_row, int ctu_rows, int ns);
int  orc_hevc_slice_y_min_ctu(int r16, int rows16, int ns, int ctu_log2);
long long orc_hevc_zrank(int x, int y, int width, int is_luma, int ctu_log2);
void orc_hevc_encode_i_ctu(const void *sy, const void *scb, const void *scr, void *ry, void *rcb, void *rcr,
                           int bd, int wc, int hc, int qp, int num_slices, int round_q12, int ctu_log2,
                           uint8_t *modes, int16_t *coef);
void orc_hevc_deblock_ctu(void *y, void *uv, int py, int puv, int wc, int hc, int qp, int is_intra,
                          const uint8_t *cbf, const int32_t *mv4, int num_slices, int ctu_log2, int bd,
                          const uint8_t *culog2);
void orc_hevc_encode_p_ctu(const void *sy, const void *scb, const void *scr, const void *fy, const void *fcb,
                           const void *fcr, void *ry, void *rcb, void *rcr, int bd, int wc, int hc, int qp,
                           int round_q12, const int32_t *mv, int ctu_log2, int16_t *coef, uint8_t *culog2);
void orc_hevc_encode_i(const void *sy, const void *scb, const void *scr, void *ry, void *rcb, void *rcr,
                       int bd, int wc, int hc, int qp, int num_slices, int round_q12, uint8_t *modes,
                       int16_t *coef);
void orc_hevc_encode_p(const void *sy, const void *scb, const void *scr, const void *fy, const void *fcb,
                       const void *fcr, void *ry, void *rcb, void *rcr, int bd, int wc, int hc, int qp,
                       int round_q12, const int32_t *mv, int16_t *coef);
void orc_hevc_dct_matrix(int N, int *M);
void orc_hevc_tqn(const int16_t *res, int log2n, int qp, int bd, int round_q12, int16_t *out);
void orc_hevc_dqitn(const int16_t *coef, int log2n, int qp, int bd, int16_t *out);
int  orc_hevc_beta(int q);
int  orc_hevc_tc(int q);
int  orc_hevc_bs(int is_intra, const uint8_t *cbf, const int32_t *mv4, int cp, int cq);
void orc_hevc_deblock(uint8_t *y, uint8_t *uv, int py, int puv, int wc, int hc, int qp, int is_intra,
                      const uint8_t *cbf, const int32_t *mv4, int num_slices);
void orc_hevc_deblock_bd(void *y, void *uv, int py, int puv, int wc, int hc, int qp, int is_intra,
                         const uint8_t *cbf, const int32_t *mv4, int num_slices, int bd);
uint32_t orc_hevc_refine8(const void *cur, const void *ref, int bd, int wc, int hc, int x0, int y0,
                          int *mvx, int *mvy);

#ifdef __cplusplus
}
#endif
#endif
