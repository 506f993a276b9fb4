/*
 * gpu_compute.h - C ABI of the MI355X (gfx950) encoder compute pipeline.
 *
 * Drop-in for the reference's Vulkan orchestration layer
 * (approach1-compute-encoder/src/gpu_compute.h).  Every declaration keeps
 * the reference name, argument meaning, return convention (0 / -1) and the
 * double-buffered staging contract; the citation on each names the
 * reference declaration it replaces.  Entry points marked EXTENSION have no
 * reference counterpart.  No HIP or torch type appears here: device
 * memory is carried as plain pointers, so the host encoders compile with a
 * plain C compiler and link against libbc250_hip.so.
 *
 * Surface layout: an image is one device allocation holding the luma plane
 * (coded size: width and height rounded up to 16) followed by the
 * interleaved Cb/Cr plane (coded width x coded height/2).  Rows are padded
 * to a pitch that is a multiple of 256 bytes.  Uploads replicate the last
 * visible row/column into the coded padding.
 */
#ifndef BC250_GPU_COMPUTE_H
#define BC250_GPU_COMPUTE_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The reference's callers test the surface handles of gpu_image_t /
 * gpu_memory_t against VK_NULL_HANDLE (encoder_h264.c:2293,2314-2316,2424;
 * encoder_h265.c:1798,1842).  Those handles are plain pointers here, so the
 * null handle is a null pointer; defined only when no other header did. */
#ifndef VK_NULL_HANDLE
#define VK_NULL_HANDLE 0
#endif

/* reference gpu_compute.h:23-40 (gpu_image_t, passed by value) */
typedef struct {
    void *y_plane;          /* device pointer, luma, coded size */
    void *uv_plane;         /* device pointer, interleaved Cb/Cr */
    uint32_t width;         /* visible size */
    uint32_t height;
    uint32_t coded_width;   /* 16-aligned */
    uint32_t coded_height;
    uint32_t y_pitch;       /* bytes */
    uint32_t uv_pitch;
    int current_layout;     /* kept for the reference's struct shape; unused */
    int format;             /* GPU_IMAGE_NV12 / GPU_IMAGE_P010 */
} gpu_image_t;

/* reference gpu_compute.h:42-46 */
typedef struct {
    void *memory;           /* base of the device allocation */
    size_t size;
    void *mapped_ptr;       /* host staging while mapped, else NULL */
} gpu_memory_t;

/* reference gpu_compute.h:62-68 */
typedef struct {
    uint32_t y_pitch;
    uint64_t y_offset;
    uint32_t uv_pitch;
    uint64_t uv_offset;
    uint64_t total_size;
} gpu_nv12_layout_t;

struct bc250_hip_state;

/* reference gpu_compute.h:70-352 (bc250_gpu_context_t, caller-allocated).
 * The HIP device state lives behind `hip`; the caller zero-initialises the
 * struct and calls bc250_gpu_init(). */
typedef struct bc250_gpu_context {
    struct bc250_hip_state *hip;
    int device;
    uint32_t frames_submitted;
    uint32_t frame_width;
    uint32_t frame_height;
    /* reference gpu_compute.h:241-242: the newest reconstruction (the next
     * frame's reference picture), valid after the first dispatch.  Its
     * width/height are the coded size; read it with map_surface after
     * gpu_compute_cpu_read_safe() reports 1 (encoder_h264.c:2325-2340). */
    gpu_image_t recon_image;
    gpu_memory_t recon_memory;
    /* reference gpu_compute.h:243 and :333, read by encoder_h264.c:2315,2334:
     * a reconstruction exists; a surface was exported as dma-buf (never
     * here, export is unsupported). */
    bool has_recon_frame;
    bool any_surface_exported;
} bc250_gpu_context_t;

typedef bc250_gpu_context_t gpu_context_t;

/* reference gpu_compute.h:571-576: per-MB motion vector in quarter-pel
 * units plus the SAD of the final (sub-pel) position. */
typedef struct {
    int32_t mvx, mvy;
    uint32_t sad;
    uint32_t _pad;
} gpu_mv_t;

#define GPU_IMAGE_NV12 0 /* reference gpu_compute.h:371-372 */
#define GPU_IMAGE_P010 1

/* ---- lifecycle: reference gpu_compute.h:357-377 -------------------------- */
int bc250_gpu_init(bc250_gpu_context_t *ctx);
void bc250_gpu_destroy(bc250_gpu_context_t *ctx);
void gpu_compute_report_cu_topology(bc250_gpu_context_t *ctx);
int gpu_compute_init(gpu_context_t *ctx);
void gpu_compute_terminate(gpu_context_t *ctx);

/* ---- surfaces: reference gpu_compute.h:380-463 ---------------------------- */
int gpu_compute_create_image(gpu_context_t *ctx, int width, int height, int format, gpu_image_t *image,
                             gpu_memory_t *memory);
void gpu_compute_destroy_image(gpu_context_t *ctx, gpu_image_t image, gpu_memory_t memory);
int gpu_compute_get_nv12_layout(gpu_context_t *ctx, gpu_image_t *image, gpu_memory_t memory,
                                gpu_nv12_layout_t *layout);
/* Device memory is not host-visible: mapping hands out a pinned host mirror
 * (with the same layout) and unmap uploads it. */
uint8_t *gpu_compute_map_surface(gpu_context_t *ctx, gpu_image_t *image, gpu_memory_t memory,
                                 gpu_nv12_layout_t *layout, bool *unmap);
void gpu_compute_unmap_surface(gpu_context_t *ctx, gpu_memory_t memory, bool unmap);
int gpu_compute_upload_nv12(gpu_context_t *ctx, gpu_image_t *image, gpu_memory_t memory, const uint8_t *y_plane,
                            int y_pitch, const uint8_t *uv_plane, int uv_pitch, int width, int height);
int gpu_compute_upload_p010(gpu_context_t *ctx, gpu_image_t *image, gpu_memory_t memory, const uint16_t *y_plane,
                            int y_pitch, const uint16_t *uv_plane, int uv_pitch, int width, int height);
int gpu_compute_video_proc(gpu_context_t *ctx, gpu_image_t *src, const int src_rect[4], gpu_image_t *dst,
                           const int dst_rect[4]);
void gpu_compute_copy_from_wc(void *dst, const void *src, size_t n);
int gpu_compute_download_nv12(gpu_context_t *ctx, gpu_image_t *image, gpu_memory_t memory, uint8_t *y_plane,
                              int y_pitch, uint8_t *uv_plane, int uv_pitch, int width, int height);
int gpu_compute_export_nv12_dmabuf(gpu_context_t *ctx, gpu_memory_t memory, int *out_fd);
int gpu_compute_import_dmabuf_image(gpu_context_t *ctx, int dma_buf_fd, int width, int height, int format,
                                    uint32_t stride, uint32_t offset, gpu_image_t *image, gpu_memory_t *memory);
int gpu_compute_wait_for_image_ready(gpu_context_t *ctx, gpu_memory_t memory);
int gpu_compute_wait_for_image_ready_host(gpu_context_t *ctx, gpu_memory_t memory);
int gpu_compute_cpu_read_safe(gpu_context_t *ctx);

/* ---- encode pipeline: reference gpu_compute.h:582-612 ---------------------
 * dispatch_encode_ext records, for the current slot, the stages of one frame
 * in the reference's order: motion estimation (skipped when cpu_mvs is
 * given; those MVs are uploaded instead), then the I16x16 wavefront (intra)
 * or the P16x16 encode (inter), then deblocking, then the copies of levels,
 * dc, pred modes, MVs and nz masks into the slot's host staging.  The
 * reference picture is this context's previous reconstruction (the render
 * target itself for the first frame).  me_mode is accepted for signature
 * compatibility; ME is the +-R full search set by
 * gpu_compute_set_encode_params(). */
int gpu_compute_begin_picture(gpu_context_t *ctx, gpu_image_t render_target);
int gpu_compute_dispatch_encode(gpu_context_t *ctx, gpu_image_t render_target, int width, int height, int qp,
                                int is_intra, int num_slices);
int gpu_compute_dispatch_encode_ext(gpu_context_t *ctx, gpu_image_t render_target, int width, int height, int qp,
                                    int is_intra, int num_slices, int me_mode, const gpu_mv_t *cpu_mvs);
int gpu_compute_dispatch_me_only(gpu_context_t *ctx, gpu_image_t render_target, int width, int height);
double gpu_compute_get_last_latency_ms(const gpu_context_t *ctx);
int gpu_compute_end_picture(gpu_context_t *ctx);
int gpu_compute_sync(gpu_context_t *ctx);
int gpu_compute_submitted_slot(gpu_context_t *ctx);
int gpu_compute_sync_slot(gpu_context_t *ctx, int slot);

/* ---- readback: reference gpu_compute.h:613-648 ----------------------------
 * Pointers into pinned host buffers owned by ctx, valid until the slot is
 * reused.  Layouts: levels int16 [nmb*24*16]; dc int32 [nmb*24]; pred modes
 * uint32 [nmb]; mvs gpu_mv_t [nmb]; nz uint32 [nmb*24]. */
int gpu_compute_get_quant_staging_data_slot(gpu_context_t *ctx, int slot, void **data, size_t *size);
int gpu_compute_get_dc_staging_data_slot(gpu_context_t *ctx, int slot, void **data, size_t *size);
int gpu_compute_get_pred_mode_staging_data_slot(gpu_context_t *ctx, int slot, void **data, size_t *size);
int gpu_compute_get_mv_staging_data_slot(gpu_context_t *ctx, int slot, void **data, size_t *size);
int gpu_compute_get_nz_staging_data_slot(gpu_context_t *ctx, int slot, void **data, size_t *size);
int gpu_compute_get_staging_data(gpu_context_t *ctx, void **data, size_t *size);
int gpu_compute_release_staging_data(gpu_context_t *ctx);
int gpu_compute_get_quant_staging_data(gpu_context_t *ctx, void **data, size_t *size);
int gpu_compute_get_dc_staging_data(gpu_context_t *ctx, void **data, size_t *size);
int gpu_compute_get_pred_mode_staging_data(gpu_context_t *ctx, void **data, size_t *size);
int gpu_compute_get_mv_staging_data(gpu_context_t *ctx, void **data, size_t *size);
int gpu_compute_get_nz_staging_data(gpu_context_t *ctx, void **data, size_t *size);

/* reference gpu_compute.h:559: creates a diagnostic dump file (mode 0600, a
 * symbolic link at the path is refused) in $BC250_DUMP_DIR, else
 * $XDG_RUNTIME_DIR/bc250_dump_frames (created 0700); `what` names the switch
 * that asked, for the messages.  NULL on failure.  The encoders' own dump
 * switches call it (encoder_h264.c:2815,2821; encoder_h265.c:1745). */
FILE *bc250_debug_dump_open(const char *name, const char *what);

/* ---- debug hooks: reference gpu_compute.h:556-667 (no-ops here) ---------- */
void bc250_debug_dump_nv12_frame(const uint8_t *y_plane, int y_pitch, const uint8_t *uv_plane, int uv_pitch,
                                 int width, int height);
void gpu_compute_debug_dump_recon(gpu_context_t *ctx, int width, int height);
void gpu_compute_debug_dump_real_input(gpu_context_t *ctx, gpu_image_t *image, gpu_memory_t memory, int width,
                                       int height);

/* ---- EXTENSIONS (no reference counterpart) -------------------------------- */

/* Encoder knobs the reference fixes in code (its push constants,
 * gpu_compute.c:2982-2983) or in environment variables. */
typedef struct {
    int search_range;  /* +-R integer full search: 16, 32 or 64 (default 64) */
    int lambda8;       /* 0..15: cost8 = SAD8 + lambda8*|mv|, cost16 = SAD16 + 4*lambda8*|mv| (default 1) */
    int deblock_idc;   /* disable_deblocking_filter_idc: 0 on, 1 off, 2 not across slices (default 0) */
    int me_8x8;        /* also refine the four 8x8 quadrant MVs (default 1) */
    int readback;      /* copy results into host staging at end_picture (default 1) */
    int codec;         /* GPU_CODEC_H264 (default) or GPU_CODEC_HEVC, see below */
    int square_radius; /* H.264: 0 = the +-search_range full search (default); 2..16 = the reference's
                        * Tier-2 square search (cpu_simd_me.c:275-497) on the GPU with this radius,
                        * its vectors used as they come (no 8x8 vectors, no further refinement) */
    int diamond;       /* H.264: 1 = the reference's diamond search (motion_estimation.comp) with integer
                        * SADs in place of the full search: lambda 5, or 8 when dispatch_encode_ext's
                        * me_mode is 1 (the reference's fast mode); no 8x8 vectors */
    int hevc_ctu_log2; /* HEVC coding tree unit: 0 or 4 = CTU16 (default, the reference's), 5 = CTU32:
                        * I pictures code the four 16x16 quadrants of each CTU32 in z-order with the
                        * CTU32 z-scan availability, slices are CTU32-row ranges (P pictures and the
                        * coefficient / mode layouts, indexed by 16x16 area, are unchanged) */
    int entropy;       /* H.264: 1 = write each frame's CAVLC slices on the GPU after deblocking and copy
                        * them to the slot (include/bc250_bitstream.h); 0 = staging only */
} gpu_encode_params_t;

/* codec = GPU_CODEC_HEVC runs the HEVC stages of include/bc250_hevc.h
 * instead: I pictures as CTU16 / CU8 / PU4 intra (the reference's
 * intra_trial), P pictures as 8x8 inter CUs with the +-R full-search 8x8
 * vectors refined to quarter samples by the HEVC interpolation, 8x8 luma and
 * 4x4 chroma DCT, then the 8.7.2 deblocking filter unless deblock_idc is 1
 * (slices never filter across their borders, so 0 and 2 agree).  deblock_idc
 * 1 is the reference's HEVC encoder, whose PPS sets
 * pps_deblocking_filter_disabled_flag (encoder_h265.c:330).  Staging:
 * quant = int16 [nctu*384] coefficients in the bc250_hevc.h layout, mv8 =
 * the CU vectors (quarter samples), HEVC modes via
 * gpu_compute_get_hevc_mode_staging_data_slot. */
#define GPU_CODEC_H264 0
#define GPU_CODEC_HEVC 1

int gpu_compute_set_encode_params(gpu_context_t *ctx, const gpu_encode_params_t *params);
int gpu_compute_get_encode_params(gpu_context_t *ctx, gpu_encode_params_t *params);

/* The four 8x8 quadrant MVs of each MB (TL, TR, BL, BR), gpu_mv_t [nmb*4]. */
int gpu_compute_get_mv8_staging_data_slot(gpu_context_t *ctx, int slot, void **data, size_t *size);

/* HEVC I-picture PU modes, uint8 [nctu*16] in PU z-order (0 Planar, 1 DC,
 * 10 H, 26 V); zero after a P picture. */
int gpu_compute_get_hevc_mode_staging_data_slot(gpu_context_t *ctx, int slot, void **data, size_t *size);

/* Synchronously copies the newest reconstruction (deblocked; the reference
 * picture of the next frame) into host planes of the visible size. */
int gpu_compute_download_recon(gpu_context_t *ctx, uint8_t *y_plane, int y_pitch, uint8_t *uv_plane,
                               int uv_pitch);

/* Stage timing: with it on, every frame records a HIP event at each stage
 * boundary (a few microseconds of stream time each); off (the default,
 * unless the environment sets BC250_PERF_STATS when the context is created,
 * as the reference gates its timestamp queries, gpu_compute.c:3684-3708),
 * frames record none and the two getters below report nothing new. */
int gpu_compute_set_stage_timing(gpu_context_t *ctx, int on);

/* GPU time of the stages of the frame in `slot` (ms), from HIP events on
 * the context's stream: [0] integer search, [1] sub-pel refine, [2] intra
 * or inter encode, [3] deblock, [4] staging copies, [5] total; -1 when the
 * frame ran without stage timing. */
int gpu_compute_get_stage_times(gpu_context_t *ctx, int slot, float ms[6]);

/* Stage times summed over every frame synced since the last reset (the
 * reference's BC250_PERF_STATS timestamps, gpu_compute.c:3684-3708). */
typedef struct {
    double sum_ms[5];     /* stages as in gpu_compute_get_stage_times */
    double total_ms;
    int frames;
    int p_frames;
    double me_int_ms_p;   /* integer-search time summed over P frames */
} gpu_stage_stats_t;
int gpu_compute_get_stage_stats(gpu_context_t *ctx, gpu_stage_stats_t *stats, int reset);

/* Nonzero when a wavefront kernel gave up waiting on a dependency (its
 * results are then invalid); cleared by the call. */
int gpu_compute_take_error_flags(gpu_context_t *ctx);

#ifdef __cplusplus
}
#endif

#endif /* BC250_GPU_COMPUTE_H */
