/*
 * bc250_hevc_stream.h - the HEVC (H.265) bitstream of pictures the GPU
 * pipeline coded: VPS / SPS / PPS and, per picture, the slice segment NAL
 * units with their CABAC slice data, written on the host from the staging
 * that gpu_compute_get_staging / the HEVC launchers produce (include/
 * bc250_hevc.h: levels int16 [384] per 16x16 area, modes uint8 [16] per
 * area, 8x8 CU vectors {mvx, mvy, sad, 0} in quarter samples).
 *
 * It replaces the reference's host entropy stage for HEVC: the parameter
 * sets of src/encoder_h265.c:160-345, the slice loop of :1419-1486, the
 * CTU / CU syntax of src/hevc_enc_template.c:596-662,1063-1170,1409-1497
 * and the CABAC coder of src/hevc_cabac.c (Rec. ITU-T H.265 7.3 and 9.3).
 * The syntax is what the pipeline codes:
 *   I pictures  CTU -> 8x8 CUs, each PART_NxN with four 4x4 intra PUs
 *               (Planar / DC / H / V, most-probable-mode coded), chroma
 *               DC, one 4x4 TU per PU and one 4x4 chroma TU per CU;
 *   P pictures  8x8 (CTU16) or 8x8 / 16x16 / 32x32 (CTU32) inter CUs,
 *               PART_2Nx2N, AMVP with the spatial candidates of 8.5.3.2.7
 *               (no temporal candidate), one luma TU of the CU's size and
 *               one chroma TU per component of half its size.
 * Slices are the pipeline's CTU-row ranges [floor(rows s / ns),
 * floor(rows (s + 1) / ns)); the loop filter does not cross them.
 */
#ifndef BC250_HEVC_STREAM_H
#define BC250_HEVC_STREAM_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int32_t width, height;            /* coded size, multiples of 16 */
    int32_t crop_width, crop_height;  /* displayed size (0: the coded size); even */
    int32_t bit_depth;                /* 8 (Main) or 10 (Main 10) */
    int32_t ctu_log2;                 /* 4 (CTU16, the reference's) or 5 (CTU32) */
    int32_t num_slices;               /* CTU-row slices per picture, 1..CTU rows */
    int32_t deblock;                  /* 1: the in-loop deblocking filter on; 0: off */
    int32_t init_qp;                  /* the PPS's init_qp (0..51) */
} bc250_hevc_stream_cfg_t;

/* VPS, SPS and PPS as Annex B NAL units (each after a 4-byte start code).
 * Returns the bytes written, or 0 for an invalid configuration or when cap
 * is too small. */
size_t bc250_hevc_write_parameter_sets(const bc250_hevc_stream_cfg_t *cfg, uint8_t *out, size_t cap);

/* One picture: num_slices slice segment NAL units (Annex B), IDR_W_RADL
 * (POC 0) when is_idr, else TRAIL_R predicting from the picture before it
 * (POC poc, whose low 8 bits are sent).  qp is the picture's QP (slice_qp_
 * delta = qp - init_qp).  coef / modes / mv8 are the staging of the
 * picture at the configuration's coded size and CTU size: mv8 (4 uint32 per
 * 8x8 CU) is read for P pictures only; modes holds the I picture's PU modes
 * or, at CTU32, the P picture's CU sizes.  Returns the bytes written, or 0
 * on an invalid argument or when cap is too small. */
size_t bc250_hevc_write_picture(const bc250_hevc_stream_cfg_t *cfg, int is_idr, int poc, int qp,
                                const int16_t *coef, const uint8_t *modes, const uint32_t *mv8, uint8_t *out,
                                size_t cap);

#ifdef __cplusplus
}
#endif
#endif
