/*
 * bc250_rc.h - host-side rate control for the sequence driver and the
 * serving loop: the reference's CBR / VBR / low-latency / CQP controller
 * (approach1-compute-encoder/src/rate_control.h:13-90), restated.
 *
 * Two ways to pick a picture's QP, as in the reference:
 *   - the buffer loop (bc250_rc_get_frame_qp, rate_control.c:199-266): a
 *     leaky bucket of buffer_size bits (one second, two pictures' budget in
 *     low-latency mode) drained by target_bits_per_frame per picture; QP =
 *     base QP + a proportional term (buffer error / half the buffer * 6) + an
 *     integral term (windowed to 30 pictures, up to 40 QP), stepped at most
 *     2..5 per picture;
 *   - the model (bc250_rc_model_frame_qp, rate_control.c:443-545): per
 *     picture kind a complexity c = bits * 2^((qp - 12) / 5), and the QP at
 *     which c costs one picture's share less part of the stream's debt.
 * The base QP comes from bits per pixel against one calibration point
 * (26 at 0.036169 bit/pixel, -6 QP per doubling; rate_control.c:48-68).
 *
 * Environment, read by bc250_rc_init as the reference reads it:
 * BC250_QP_MIN (1..51, default 12).  The bucket drains by the nominal
 * per-picture quota unless bc250_rc_set_wallclock_drain(rc, 1) asks for the
 * reference's real-time drain (target bitrate x elapsed seconds, clamped to
 * 1..250 ms), which the reference selects for live-streaming processes.
 */
#ifndef BC250_RC_H
#define BC250_RC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum { BC250_RC_CBR = 0, BC250_RC_VBR = 1, BC250_RC_LOW_LATENCY = 2, BC250_RC_CQP = 3 } bc250_rc_mode_t;

typedef struct {
    int32_t mode;                 /* bc250_rc_mode_t */
    uint32_t target_bitrate, max_bitrate;
    int32_t qp_min, qp_max;
    double framerate;
    uint32_t target_bits_per_frame;
    int64_t buffer_fullness, buffer_size;
    int32_t base_qp, current_qp;
    uint64_t prev_frame_sad;
    int64_t error_integral;
    uint32_t max_frame_bits;      /* 0: no limit */
    uint32_t quality_level;       /* 1 (quality) .. 7 (speed), 4 balanced */
    int32_t wallclock_drain;      /* 0: nominal per-picture drain */
    uint64_t last_frame_ns;
    double measured_fps;
    /* the model (set `model` to 1 after bc250_rc_init to use it) */
    int32_t model;
    double cplx[2];               /* P, I complexities; 0: none seen */
    double debt;                  /* bits produced - bits allowed since init */
    int32_t model_last_p_qp;
    double cplx_still;
    int32_t still_last_qp;
    double still_share;
    double pixels;
} bc250_rc_t;

/* rc_init (rate_control.c:70-132): a fresh bucket half full; the model's
 * complexities and `model` are left as they are. */
void bc250_rc_init(bc250_rc_t *rc, bc250_rc_mode_t mode, uint32_t bitrate, double fps, uint32_t width,
                   uint32_t height);
/* rc_update_bitrate (:134-171): the bucket rescaled, QP moved by the base
 * QP's change. */
void bc250_rc_update_bitrate(bc250_rc_t *rc, uint32_t bitrate, uint32_t width, uint32_t height);
/* rc_get_frame_qp (:199-266); est_sad only matters in VBR. */
int bc250_rc_get_frame_qp(bc250_rc_t *rc, uint64_t est_sad);
/* rc_update_stats (:268-381): the picture's coded bits into the bucket. */
void bc250_rc_update_stats(bc250_rc_t *rc, int bits_used);
/* rc_model_frame_qp / rc_model_frame_coded (:443-573). */
int bc250_rc_model_frame_qp(bc250_rc_t *rc, int intra, int still);
void bc250_rc_model_frame_coded(bc250_rc_t *rc, int intra, int still, int qp, int bits);

void bc250_rc_set_max_frame_size(bc250_rc_t *rc, uint32_t max_frame_bits);
uint32_t bc250_rc_get_max_frame_size(const bc250_rc_t *rc);
void bc250_rc_set_quality_level(bc250_rc_t *rc, uint32_t quality_level);
uint32_t bc250_rc_get_quality_level(const bc250_rc_t *rc);
void bc250_rc_set_wallclock_drain(bc250_rc_t *rc, int on);

#ifdef __cplusplus
}
#endif
#endif
