/*
 * bc250_cpu_me.h - the reference's Tier-2 ("CPU SIMD") motion-search
 * interface, src/cpu_simd_me.h:21-53, with the search itself run by the
 * gfx950 kernel behind bc250_launch_square_me (bit-exact to
 * src/cpu_simd_me.c:275-497).  Same names, types and semantics, so the H.264
 * encoder's Tier-2 route (encoder_h264.c:2293-2382) links against it
 * unchanged: host planes in, gpu_mv_t vectors out, ready for
 * gpu_compute_dispatch_encode_ext(..., cpu_mvs).
 */
#ifndef BC250_CPU_ME_H
#define BC250_CPU_ME_H

#include <stdint.h>

#include "gpu_compute.h"

#ifdef __cplusplus
extern "C" {
#endif

/* cpu_simd_me.h:21-29 */
typedef struct {
    uint32_t width;
    uint32_t height;
    uint32_t width_in_mbs;
    uint32_t height_in_mbs;
    uint32_t search_radius; /* clamped to 2..16 by the search */
    int num_threads;        /* kept for the caller; the GPU search ignores it */
    int core_ids[2];
} cpu_simd_me_config_t;

/* cpu_simd_me.h:35-36: SAD of two 16x16 blocks (host code). */
uint32_t cpu_simd_sad_16x16(const uint8_t *src, int src_stride, const uint8_t *ref, int ref_stride);

/* cpu_simd_me.h:41, cpu_simd_me.c:199-244: fills width, height and the MB
 * counts, num_threads = 4 (BC250_MAX_CPU_THREADS / BC250_THREADS /
 * BC250_CPU_THREADS override it), core_ids = -1; search_radius is left as the
 * caller set it, as in the reference. */
void cpu_simd_me_config_init(cpu_simd_me_config_t *cfg, uint32_t width, uint32_t height);

/* cpu_simd_me.h:48-53: one P frame's vectors.  src_y / ref_y are host luma
 * planes holding at least the MB-padded frame (16 * ceil(height / 16) rows);
 * out_mvs has one gpu_mv_t per MB.  The planes are copied to the GPU, searched
 * there, and the vectors copied back.  0 on success, -1 on error. */
int cpu_simd_me_search_frame(const uint8_t *src_y, int src_pitch, const uint8_t *ref_y, int ref_pitch,
                             uint32_t width, uint32_t height, gpu_mv_t *out_mvs, const cpu_simd_me_config_t *cfg);

#ifdef __cplusplus
}
#endif
#endif
