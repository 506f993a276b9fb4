/*
 * bc250_rc.c - the rate controller (bc250_rc.h).  The arithmetic follows
 * approach1-compute-encoder/src/rate_control.c operation for operation,
 * in the same floating-point order, so that a picture gets the QP the
 * reference would give it (tests/test_rate_control.py drives both).
 */
#define _POSIX_C_SOURCE 200809L
#include "bc250_rc.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* Starting QP from bits per pixel (rate_control.c:48-68): 26 at 0.036169
 * bit/pixel (1 Mbit/s of 1280x720 at 30 fps), 6 QP per doubling. */
#define BPP_ANCHOR 0.036169
#define QP_ANCHOR 26.0
#define QP_PER_DOUBLING 6.0
/* the buffer loop's integral term (:186-197) */
#define INT_WINDOW 30.0
#define INT_QP_RANGE 40.0
/* the model (:394-398) */
#define M_QP_HALF 5.0
#define M_MAX_STEP 2
#define M_I_BOOST 2
#define M_I_OVER_P 4.0
#define M_STILL_RANGE 6

static int env_qp_min(void)
{
    const char *e = getenv("BC250_QP_MIN");
    if (e && *e) {
        const int q = atoi(e);
        if (q >= 1 && q <= 51) return q;
    }
    return 12;
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }

static int base_qp_for(uint32_t bitrate, double fps, uint32_t width, uint32_t height)
{
    const double px_per_s = (double)width * (double)height * (fps > 0 ? fps : 60.0);
    if (px_per_s <= 0.0) return (int)QP_ANCHOR;
    const double bpp = (double)bitrate / px_per_s;
    if (bpp <= 0.0) return 51;
    const double qp = QP_ANCHOR - QP_PER_DOUBLING * (log(bpp / BPP_ANCHOR) / log(2.0));
    return clampi((int)lround(qp), env_qp_min(), 51);
}

/* the bucket of a mode: two pictures' budget in low latency, else a second */
static int64_t bucket_size(const bc250_rc_t *rc)
{
    int64_t s = rc->mode == BC250_RC_LOW_LATENCY ? (int64_t)(uint32_t)(rc->target_bits_per_frame * 2)
                                                  : (int64_t)rc->target_bitrate;
    return s < 1000 ? 1000 : s;
}

void bc250_rc_init(bc250_rc_t *rc, bc250_rc_mode_t mode, uint32_t bitrate, double fps, uint32_t width,
                   uint32_t height)
{
    if (!rc) return;
    rc->mode = mode;
    rc->target_bitrate = bitrate > 0 ? bitrate : 5000000;
    rc->max_bitrate = rc->target_bitrate * 3 / 2;
    rc->qp_min = env_qp_min();
    rc->qp_max = 51;
    rc->framerate = fps > 0 ? fps : 60.0;
    rc->target_bits_per_frame = (uint32_t)(rc->target_bitrate / rc->framerate);
    if (rc->target_bits_per_frame < 100) rc->target_bits_per_frame = 100;
    rc->buffer_size = bucket_size(rc);
    rc->buffer_fullness = rc->buffer_size / 2;
    int b = base_qp_for(rc->target_bitrate, rc->framerate, width, height);
    /* constant QP keeps a QP the caller already put in current_qp */
    if (mode == BC250_RC_CQP && rc->current_qp >= 12 && rc->current_qp <= 51) b = rc->current_qp;
    rc->base_qp = rc->current_qp = b;
    rc->prev_frame_sad = 0;
    rc->error_integral = 0;
    rc->max_frame_bits = 0;
    rc->quality_level = 4;
    rc->last_frame_ns = 0;
    rc->measured_fps = 0.0;
    rc->debt = 0.0;
    rc->pixels = (double)width * (double)height;
}

void bc250_rc_update_bitrate(bc250_rc_t *rc, uint32_t bitrate, uint32_t width, uint32_t height)
{
    if (!rc || bitrate == 0 || bitrate == rc->target_bitrate) return;
    const double ratio = (double)bitrate / (double)(rc->target_bitrate > 0 ? rc->target_bitrate : bitrate);
    rc->target_bitrate = bitrate;
    rc->target_bits_per_frame = (uint32_t)(rc->target_bitrate / (rc->framerate > 0.0 ? rc->framerate : 30.0));
    if (rc->target_bits_per_frame < 100) rc->target_bits_per_frame = 100;
    const int64_t old_size = rc->buffer_size;
    rc->buffer_size = bucket_size(rc);
    if (old_size > 0) {
        int64_t f = (int64_t)((double)rc->buffer_fullness * ratio);
        rc->buffer_fullness = f > rc->buffer_size ? rc->buffer_size : f < 0 ? 0 : f;
    } else {
        rc->buffer_fullness = rc->buffer_size / 2;
    }
    const int nb = base_qp_for(rc->target_bitrate, rc->framerate, width, height);
    rc->current_qp = clampi(rc->current_qp + (nb - rc->base_qp), rc->qp_min, rc->qp_max);
    rc->base_qp = nb;
}

int bc250_rc_get_frame_qp(bc250_rc_t *rc, uint64_t est_sad)
{
    if (!rc) return 26;
    if (rc->mode == BC250_RC_CQP) return rc->current_qp;
    int64_t half = rc->buffer_size / 2;
    if (half <= 0) half = 1;
    const int64_t err = rc->buffer_fullness - half;
    const double p = (double)err / (double)half * 6.0;
    int64_t cap = (int64_t)(half * INT_WINDOW);
    if (cap < 1) cap = 1;
    /* no integration in a direction QP can no longer follow */
    const int pinned = (rc->current_qp >= rc->qp_max && err > 0) || (rc->current_qp <= rc->qp_min && err < 0);
    if (!pinned) {
        rc->error_integral += err;
        if (rc->error_integral > cap) rc->error_integral = cap;
        if (rc->error_integral < -cap) rc->error_integral = -cap;
    }
    const double i = ((double)rc->error_integral / (double)cap) * INT_QP_RANGE;
    int adj = (int)round(p + i);
    if (rc->mode == BC250_RC_VBR && rc->prev_frame_sad > 0) {
        const double r = (double)est_sad / (double)rc->prev_frame_sad;
        adj += r > 1.3 ? 2 : r < 0.7 ? -2 : 0;
    }
    if (rc->max_frame_bits > 0 && rc->target_bits_per_frame > 0) {
        if (rc->target_bits_per_frame > rc->max_frame_bits) adj += 2;
        else if (rc->max_frame_bits < rc->target_bits_per_frame * 3 / 2) adj += 1;
    }
    int step = (rc->mode == BC250_RC_LOW_LATENCY || rc->quality_level >= 5) ? 3 : 2;
    int d = rc->base_qp + adj - rc->current_qp;
    step += abs(d) >= 8 ? 2 : abs(d) >= 4 ? 1 : 0;
    d = clampi(d, -step, step);
    rc->current_qp = clampi(rc->current_qp + d, rc->qp_min, rc->qp_max);
    rc->prev_frame_sad = est_sad;
    return rc->current_qp;
}

void bc250_rc_update_stats(bc250_rc_t *rc, int bits_used)
{
    if (!rc || rc->mode == BC250_RC_CQP) return;
    rc->buffer_fullness += bits_used;
    int64_t drain = rc->target_bits_per_frame;
    if (rc->wallclock_drain) {
        struct timespec t;
        uint64_t now = 0;
        if (clock_gettime(CLOCK_MONOTONIC, &t) == 0) now = (uint64_t)t.tv_sec * 1000000000ull + (uint64_t)t.tv_nsec;
        if (now && rc->last_frame_ns && now > rc->last_frame_ns) {
            double el = (double)(now - rc->last_frame_ns) / 1e9;
            el = el < 0.001 ? 0.001 : el > 0.250 ? 0.250 : el;
            drain = (int64_t)((double)rc->target_bitrate * el);
            rc->measured_fps = rc->measured_fps > 0.0 ? rc->measured_fps * 0.95 + (1.0 / el) * 0.05 : 1.0 / el;
        }
        if (now) rc->last_frame_ns = now;
    }
    rc->buffer_fullness -= drain;
    if (rc->model) {
        /* overspending is owed; CBR and low latency bank nothing unspent,
         * VBR up to half a second */
        rc->debt += (double)bits_used - (double)drain;
        const double lo = rc->mode == BC250_RC_VBR ? -(double)rc->target_bitrate / 2.0 : 0.0;
        if (rc->debt < lo) rc->debt = lo;
    }
    if (rc->buffer_fullness < 0) rc->buffer_fullness = 0;
    else if (rc->buffer_fullness > rc->buffer_size) rc->buffer_fullness = rc->buffer_size;
}

void bc250_rc_set_max_frame_size(bc250_rc_t *rc, uint32_t max_frame_bits)
{
    if (rc) rc->max_frame_bits = max_frame_bits;
}
uint32_t bc250_rc_get_max_frame_size(const bc250_rc_t *rc) { return rc ? rc->max_frame_bits : 0; }
void bc250_rc_set_quality_level(bc250_rc_t *rc, uint32_t quality_level)
{
    if (rc) rc->quality_level = quality_level < 1 ? 1 : quality_level > 7 ? 7 : quality_level;
}
uint32_t bc250_rc_get_quality_level(const bc250_rc_t *rc) { return rc ? rc->quality_level : 4; }
void bc250_rc_set_wallclock_drain(bc250_rc_t *rc, int on)
{
    if (rc) rc->wallclock_drain = on != 0;
}

/* ---- the model */

static double m_qp_exact(double cplx, double bits) { return 12.0 + M_QP_HALF * log2(cplx / bits); }
static double m_bits(double cplx, int qp) { return cplx * pow(2.0, -(qp - 12) / M_QP_HALF); }

/* pictures over which a debt is paid back */
static double m_horizon(const bc250_rc_t *rc)
{
    const double h = rc->mode == BC250_RC_LOW_LATENCY ? rc->framerate / 4.0
                     : rc->mode == BC250_RC_CBR      ? rc->framerate / 2.0
                                                     : rc->framerate;
    return h < 2.0 ? 2.0 : h;
}

int bc250_rc_model_frame_qp(bc250_rc_t *rc, int intra, int still)
{
    if (!rc) return 26;
    if (rc->mode == BC250_RC_CQP) return rc->current_qp;
    /* a moving P picture takes the still pictures' shares too (up to twice
     * its own), a still one what is left over */
    const double share = (double)rc->target_bitrate / rc->framerate;
    const double s = rc->still_share, sm = s < 0.5 ? s : 0.5;
    double bpf = share;
    if (!intra && !still) {
        bpf = share / (1.0 - sm);
    } else if (!intra && s > 0.0) {
        bpf = (share - (1.0 - s) * share / (1.0 - sm)) / s;
        if (bpf < share / 64.0) bpf = share / 64.0;
    }
    double want = bpf - rc->debt / m_horizon(rc);
    if (want < bpf / 8.0) want = bpf / 8.0;
    if (want > bpf * 4.0) want = bpf * 4.0;

    int qp;
    double cplx;
    if (rc->cplx[0] > 0.0) {
        const int st = still && !intra && rc->model_last_p_qp > 0;
        const double k = st && rc->cplx_still > 0.0 ? rc->cplx_still : rc->cplx[0];
        const int last = st && rc->still_last_qp > 0 ? rc->still_last_qp : rc->model_last_p_qp;
        const double ex = m_qp_exact(k, want);
        qp = (int)lround(ex);
        if (last > 0 && fabs(ex - last) < 1.0) qp = last;  /* hold until a whole step away */
        if (last > 0) qp = clampi(qp, last - M_MAX_STEP, last + M_MAX_STEP);
        if (st) {
            if (qp > rc->model_last_p_qp) qp = rc->model_last_p_qp;
            if (qp < rc->model_last_p_qp - M_STILL_RANGE) qp = rc->model_last_p_qp - M_STILL_RANGE;
        }
        if (intra) qp -= M_I_BOOST;
        cplx = intra ? (rc->cplx[1] > 0.0 ? rc->cplx[1] : rc->cplx[0] * M_I_OVER_P) : k;
    } else if (rc->cplx[1] > 0.0) {
        qp = (int)lround(m_qp_exact(rc->cplx[1] / M_I_OVER_P, want));
        if (intra) qp -= M_I_BOOST;
        cplx = intra ? rc->cplx[1] : rc->cplx[1] / M_I_OVER_P;
    } else {
        /* first guess: QP 36 at 0.1 bit per pixel */
        const double bpp = rc->pixels > 0.0 ? bpf / rc->pixels : 0.1;
        qp = (int)lround(36.0 - M_QP_HALF * log2(bpp / 0.1)) - (intra ? M_I_BOOST : 0);
        cplx = 0.0;
    }
    if (rc->max_frame_bits > 0 && cplx > 0.0)
        while (qp < rc->qp_max && m_bits(cplx, qp) > (double)rc->max_frame_bits) qp++;
    qp = clampi(qp, rc->qp_min, rc->qp_max);
    rc->current_qp = qp;
    return qp;
}

void bc250_rc_model_frame_coded(bc250_rc_t *rc, int intra, int still, int qp, int bits)
{
    if (!rc || bits <= 0) return;
    const double c = (double)bits * pow(2.0, (qp - 12) / M_QP_HALF);
    const int st = still && !intra && rc->model_last_p_qp > 0;
    double *k = intra ? &rc->cplx[1] : st ? &rc->cplx_still : &rc->cplx[0];
    if (!intra) rc->still_share = rc->still_share * (15.0 / 16.0) + (st ? 1.0 / 16.0 : 0.0);
    /* a log-average with weight 1/4 on the newest, restarted by a 16x jump */
    if (*k > 0.0 && c < *k * 16.0 && c > *k / 16.0) *k = exp(0.75 * log(*k) + 0.25 * log(c));
    else *k = c;
    if (st) {
        rc->still_last_qp = qp;
    } else if (!intra) {
        rc->model_last_p_qp = qp;
        rc->still_last_qp = 0;
    }
}
