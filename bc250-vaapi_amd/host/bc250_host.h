/*
 * bc250_host.h - host-side C of the MI355X encoder pipeline: the seeded
 * synthetic frame generator, the H.264 host DC stages that the entropy
 * coder runs on the GPU's `dc` readback, and the frame-sequence driver
 * built on the gpu_compute_* ABI.
 */
#ifndef BC250_HOST_H
#define BC250_HOST_H

#include <stdint.h>
#include "bc250_rc.h"

#include "gpu_compute.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- seeded synthetic frames --------------------------------------------
 * luma = 128 + sum of 8 sinusoids (periods 8..128 px, sum of |amplitude| =
 * 40) translated per frame by a global quarter-pel motion uniform in
 * [-4,4] px, then `nrect` rectangles (32..128 px) of uniform-noise texture
 * moving by a random +-8 px per frame, then Gaussian noise sigma 2, clamped
 * to [16,235]; chroma = 2x2 average of luma re-centred at 128 +-20.  Frame
 * f depends only on (seed, f): frames may be generated in any order and on
 * any number of threads. */
typedef struct bc250_gen bc250_gen_t;
bc250_gen_t *bc250_gen_create(uint64_t seed, int width, int height, int nrect);
void bc250_gen_destroy(bc250_gen_t *g);
/* NV12 into host planes (uv interleaved Cb/Cr). Returns 0. */
int bc250_gen_frame(bc250_gen_t *g, int frame, uint8_t *y, int y_pitch, uint8_t *uv, int uv_pitch);
/* Global offset of frame f in quarter pels (for tests). */
void bc250_gen_offset_qpel(bc250_gen_t *g, int frame, int *ox, int *oy);

/* ---- H.264 host DC stages (reference encoder_h264.c:653-797) ------------ */
int h264_chroma_qp(int qpy);
/* Reference encoder_h264.c:775: transmitted I16x16 luma DC levels in the
 * natural row/column order (Hadamard, threshold-5 quantiser, transpose). */
void h264_intra16_luma_dc_transform(const int dc_in[4][4], int qp, int dc_out[16], int dc_out_pretranspose[16]);
/* Transmitted chroma DC levels of one component (2x2 Hadamard + quantiser
 * at QPc), raster TL, TR, BL, BR. */
void h264_chroma_dc_levels(const int dc_in[4], int qpc, int out[4]);
/* Frame-level DC levels from the `dc` readback: luma [nmb*16] (I16x16
 * frames only, zero otherwise) and chroma [nmb*8] (Cb 4 then Cr 4). */
void h264_frame_dc_levels(const int32_t *dc, int nmb, int qp, int is_intra, int16_t *luma_dc, int16_t *chroma_dc);

/* NAL packaging of the GPU-written slices (the reference's
 * encoder_h264.c:3440-3470 and bitstream.c:232-250).
 * bc250_rbsp_to_ebsp: emulation prevention (7.4.1): after two zero bytes, a
 * byte <= 3 is preceded by 0x03; returns the bytes written (stops at cap).
 * bc250_h264_frame_nals: every slice of a frame (RBSPs back to back as
 * slice_info describes them, see bc250_bitstream.h) as an Annex B NAL unit:
 * 00 00 00 01, header 0x65 (IDR, nal_ref_idc 3) or 0x41 (P, nal_ref_idc 2),
 * the EBSP.  Returns the bytes written, or 0 when cap (2 * RBSP + 5 bytes
 * per slice is always enough) is too small. */
size_t bc250_rbsp_to_ebsp(uint8_t *dst, size_t cap, const uint8_t *src, size_t n);

/* Sequence and picture parameter sets, with the fields the reference's
 * h264_sps_t / h264_pps_t carry (bitstream.h:94-150; all int32 here), in the
 * syntax of 7.3.2.1.1 / 7.3.2.2 that its bs_write_sps / bs_write_pps emit
 * (bitstream.c:301-417): constraint flags 0, no scaling lists, VUI with an
 * extended SAR and optional timing, one slice group.  Written as Annex B NAL
 * units (nal_ref_idc 3); return the bytes written, 0 when cap is too small
 * or a field is outside its range in 7.4.2.1.1 / 7.4.2.2 (pic_order_cnt_type 1,
 * whose offset cycle is not written, included). */
typedef struct bc250_h264_sps {
    int32_t profile_idc, level_idc, sps_id, chroma_format_idc, bit_depth_luma, bit_depth_chroma;
    int32_t log2_max_frame_num_minus4, pic_order_cnt_type, log2_max_poc_lsb_minus4, max_num_ref_frames;
    int32_t width_in_mbs, height_in_mbs, frame_mbs_only, direct_8x8_inference;
    int32_t frame_cropping, crop_left, crop_right, crop_top, crop_bottom;
    int32_t vui_present, sar_width, sar_height, timing_info_present, num_units_in_tick, time_scale;
} bc250_h264_sps_t;
typedef struct bc250_h264_pps {
    int32_t pps_id, sps_id, entropy_coding_mode, pic_order_present, num_ref_idx_l0_default,
        num_ref_idx_l1_default, weighted_pred, weighted_bipred_idc, pic_init_qp_minus26, chroma_qp_offset,
        second_chroma_qp_offset, deblocking_filter_control, constrained_intra_pred, transform_8x8_mode;
} bc250_h264_pps_t;
size_t bc250_h264_write_sps(const bc250_h264_sps_t *sps, uint8_t *dst, size_t cap);
size_t bc250_h264_write_pps(const bc250_h264_pps_t *pps, uint8_t *dst, size_t cap);
size_t bc250_h264_frame_nals(const uint8_t *rbsp, const uint32_t *slice_info, int num_slices, int is_idr,
                             uint8_t *dst, size_t cap);

/* ---- sequence driver ------------------------------------------------------ */
typedef struct {
    int qp;          /* CQP */
    int gop;         /* I-frame period (0: only the first frame is I) */
    int num_slices;
    int pipelined;   /* submit frame n+1 before syncing frame n */
} bc250_seq_cfg_t;

/* Encodes frames[0..n) (device images already holding the pictures) as
 * I P P P ... through begin_picture / dispatch_encode_ext / end_picture /
 * sync_slot.  frame_ms[i] (optional) receives submit-to-sync latency of
 * frame i (CLOCK_MONOTONIC), *total_ms the wall time of the whole run. */
int bc250_encode_sequence(gpu_context_t *ctx, const gpu_image_t *frames, int n, int width, int height,
                          const bc250_seq_cfg_t *cfg, double *frame_ms, double *total_ms);

/* Per-frame consumer of the sequence driver, called right after frame
 * `frame` synced in `slot`: its staging (readback 1) and slices (entropy 1)
 * stay valid until the driver submits the frame after next into the same
 * slot.  A non-zero return stops the run (the driver then returns -1). */
typedef int (*bc250_frame_cb)(void *user, int frame, int slot, int is_intra);
/* bc250_encode_sequence with a consumer (on_frame may be NULL); a frame
 * whose sync fails stops the run before its consumer is called. */
int bc250_encode_sequence_cb(gpu_context_t *ctx, const gpu_image_t *frames, int n, int width, int height,
                             const bc250_seq_cfg_t *cfg, double *frame_ms, double *total_ms, bc250_frame_cb on_frame,
                             void *user);

/* A consumer that writes the H.264 elementary stream: each frame's CAVLC
 * slices (gpu_compute_get_slices_slot) as Annex B NAL units, appended to
 * buf.  frame_bytes[frame] (optional, n entries) receives each frame's size.
 * Stops the run when the slices are missing or buf is full. */
typedef struct {
    gpu_context_t *ctx;
    uint8_t *buf;
    size_t cap, used;
    size_t *frame_bytes;
} bc250_nal_sink_t;
int bc250_nal_sink_frame(void *user, int frame, int slot, int is_intra);

/* A rate-controlled sequence (H.264, the context created with entropy 1):
 * each picture's QP comes from rc at its submission (bc250_rc_get_frame_qp,
 * or bc250_rc_model_frame_qp when rc->model; clamped to [rc->qp_min, 51] as
 * encoder_h264.c:2265-2269 does), and once it has synced, its Annex B bytes
 * (written to sink) x 8 go back through bc250_rc_update_stats (and
 * bc250_rc_model_frame_coded).  Pipelined, a picture's QP sees the pictures
 * finished before its submission, like the reference's submit / finish
 * split (encoder_h264.c:2244-2392, 2513-3542).  qp_out[i] (required) and
 * bits_out[i] (optional) receive each picture's QP and bits. */
int bc250_encode_sequence_rc(gpu_context_t *ctx, const gpu_image_t *frames, int n, int width, int height,
                             const bc250_seq_cfg_t *cfg, bc250_rc_t *rc, bc250_nal_sink_t *sink, int *qp_out,
                             int *bits_out, double *frame_ms, double *total_ms);

#ifdef __cplusplus
}
#endif
#endif
