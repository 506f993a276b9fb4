/* bc250_bitstream.h - entropy coding of the staged frame on the GPU.
 *
 * H.264 CAVLC slices written from the device-resident staging buffers of
 * gpu_compute.h (levels, DC terms, modes, vectors), in place of the host
 * loop of the reference's src/encoder_h264.c:3031-3240 (slice header, the
 * I16x16 / P16x16 / P_Skip macroblock loop and its residual writers in
 * src/cavlc.c:75-796, rbsp trailing bits).  The output is each slice's RBSP
 * -- exactly the bytes the reference passes to bs_rbsp_to_ebsp() for its
 * NAL unit (src/encoder_h264.c:3440-3470) -- packed back to back.
 */
#ifndef BC250_BITSTREAM_H
#define BC250_BITSTREAM_H

#include <stddef.h>
#include <stdint.h>

#include "gpu_compute.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The slice parameters the reference's writer takes from its SPS / PPS and
 * frame state (src/encoder_h264.c:3001-3005, 3097-3134).  All fields int32
 * so that the structure is also an int32[14] array. */
typedef struct bc250_h264_slice_cfg {
    int32_t width_in_mbs, height_in_mbs;
    int32_t qp;                        /* the frame's QP; slice_qp_delta = qp - 26 - pic_init_qp_minus26 */
    int32_t is_idr;                    /* 1: I slice of an IDR picture (every MB I16x16); 0: P slice */
    int32_t num_slices;                /* 1..64; slice s covers MBs [s * total / n, (s + 1) * total / n) */
    int32_t pps_id;
    int32_t pic_init_qp_minus26;
    int32_t frame_num;
    int32_t log2_max_frame_num_minus4; /* 0..12 */
    int32_t idr_pic_id;
    int32_t poc;                       /* written modulo 2^(log2_max_poc_lsb_minus4 + 4) */
    int32_t log2_max_poc_lsb_minus4;   /* 0..12 */
    int32_t deblock_idc;               /* disable_deblocking_filter_idc 0..2 */
    int32_t relax_chroma_skip;         /* the reference's quality_level >= 5 P_Skip rule */
} bc250_h264_slice_cfg_t;

/* Device bytes of the work buffer bc250_h264_cavlc_slices needs. */
size_t bc250_h264_cavlc_work_size(int width_in_mbs, int height_in_mbs);

/* Writes all slices of one frame.  Device pointers: levels int16
 * [nmb][24][16], dc int32 [nmb][24], modes uint32 [nmb] (I16x16 mode in bits
 * 0-1, chroma mode in bits 2-3; read for IDR frames), mvs [nmb] (read for
 * P frames), out (4-byte aligned, out_cap bytes), slice_info uint32
 * [4 * num_slices + 2], work (bc250_h264_cavlc_work_size bytes).  cfg is a
 * host pointer.
 *
 * slice_info on completion: [4 s] byte offset of slice s in out, [4 s + 1]
 * its RBSP length in bytes, [4 s + 2] the bit position of its trailing
 * mb_skip_run / stop bit, [4 s + 3] that trailing run; [4 n] total bytes,
 * [4 n + 1] nonzero when the total exceeded out_cap (nothing past out_cap is
 * written; the caller retries with a larger buffer).
 *
 * Frames of up to 2^17 macroblocks; at most the first 512 MB of out are
 * used.  Everything is enqueued on `stream` (a hipStream_t; NULL = default);
 * returns 0, or -1 for invalid arguments or a failed launch. */
int bc250_h264_cavlc_slices(const int16_t *levels, const int32_t *dc, const uint32_t *modes, const gpu_mv_t *mvs,
                            const bc250_h264_slice_cfg_t *cfg, uint8_t *out, size_t out_cap, uint32_t *slice_info,
                            void *work, size_t work_size, void *stream);

/* The same writer with no work argument, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  Its
 * work buffer is a file-scope device array of BC250_CAVLC_SCRATCH_BYTES
 * (4 MiB) held by each GPU the library runs on until the process ends, which
 * bounds the frame to bc250_h264_cavlc_noalloc_max_mbs() MBs (34,944: 4K UHD
 * is 32,400); larger frames are refused (-1, with a message).  It calls no
 * allocator.  Every call on a device shares that device's array: calls that
 * overlap on one device (different streams, or threads) race and must be
 * serialised by the caller.  Output, arguments and return value are those of
 * bc250_h264_cavlc_slices. */
#define BC250_CAVLC_SCRATCH_BYTES (4 << 20)
int bc250_h264_cavlc_slices_noalloc(const int16_t *levels, const int32_t *dc, const uint32_t *modes,
                                    const gpu_mv_t *mvs, const bc250_h264_slice_cfg_t *cfg, uint8_t *out,
                                    size_t out_cap, uint32_t *slice_info, void *stream);
size_t bc250_h264_cavlc_noalloc_max_mbs(void);

/* Copies slice_info (4 * num_slices + 2 words) and the slices' bytes that
 * bc250_h264_cavlc_slices wrote (slice_info[4 n] of them, at most dst_cap) to
 * dst / dst_info -- any memory the device can write, e.g. hipHostMalloc'd
 * pinned host buffers -- so that a readback moves exactly the coded bytes
 * without the host knowing their number.  Enqueued on `stream`. */
int bc250_h264_cavlc_copy_out(const uint8_t *out, const uint32_t *slice_info, int num_slices, uint8_t *dst,
                              size_t dst_cap, uint32_t *dst_info, void *stream);

/* ---- in the pipeline (gpu_compute.h): gpu_encode_params_t.entropy = 1 ----
 * Every H.264 frame then ends with its CAVLC slices, written on the GPU from
 * the frame's staging and copied to the slot's pinned host buffers.  The
 * context numbers the frames the usual IPPP way: an intra frame is an IDR
 * (frame_num 0, idr_pic_id + 1), each P frame increments frame_num, and poc
 * is 2 * frame_num.  Not available for HEVC (the dispatch fails). */

/* Sets the slice-header fields the pipeline does not derive: pps_id,
 * pic_init_qp_minus26, log2_max_frame_num_minus4, log2_max_poc_lsb_minus4
 * and relax_chroma_skip.  The other fields of cfg are ignored: sizes, QP,
 * slices and the frame numbering come from the dispatch, deblock_idc from
 * the encode params.  Defaults are all 0.
 * Returns 0, or -1 for out-of-range values. */
int gpu_compute_set_h264_slice_header(gpu_context_t *ctx, const bc250_h264_slice_cfg_t *cfg);

/* The slices of the frame in `slot`, after gpu_compute_sync_slot: *rbsp
 * points at the RBSPs back to back, *info at slice_info as above (pinned host
 * memory owned by the context, valid until the slot is reused), *num_slices
 * at their number.  Returns -1 when the frame was coded without entropy = 1
 * or its slices did not fit the buffers. */
int gpu_compute_get_slices_slot(gpu_context_t *ctx, int slot, const uint8_t **rbsp, const uint32_t **info,
                                int *num_slices);

#ifdef __cplusplus
}
#endif
#endif
