/*
 * h264_host.c - the H.264 host DC stages and the sequence driver.
 *
 * The DC functions are what the host entropy coder applies to the GPU's
 * `dc` readback before CAVLC/CABAC (reference encoder_h264.c:653-797); the
 * GPU reconstruction (intra and inter kernels) applies the same quantiser,
 * so the transmitted DC levels and the reconstruction agree.
 */
#define _POSIX_C_SOURCE 199309L
#include <stdio.h>
#include <string.h>
#include <time.h>

#include "bc250_bitstream.h"
#include "bc250_host.h"
#include "bc250_rc.h"

static const int DC_LS0[6] = {160, 176, 208, 224, 256, 288};

int h264_chroma_qp(int qpy)
{
    static const int t[22] = {29, 30, 31, 32, 32, 33, 34, 34, 35, 35, 36,
                              36, 37, 37, 37, 38, 38, 38, 39, 39, 39, 39};
    if (qpy < 0) qpy = 0;
    if (qpy > 51) qpy = 51;
    return qpy < 30 ? qpy : t[qpy - 30];
}

/* Round-to-nearest DC quantiser, threshold 5 for luma and chroma alike
 * (reference encoder_h264.c:653-670, :687-704). */
static int quant_dc(int v, int qp)
{
    const int per = qp / 6;
    const int64_t ls = DC_LS0[qp % 6];
    const int64_t av = v < 0 ? -(int64_t)v : v;
    int64_t num, den;
    if (per >= 5) {
        num = av;
        den = ls << (per - 5);
    } else {
        num = av << (5 - per);
        den = ls;
    }
    const int64_t l = (num + den / 2) / den;
    return (int)(v < 0 ? -l : l);
}

/* Two-pass 4x4 Hadamard whose second pass rounds (x+1)>>1; its output is
 * the transpose of the natural 2-D transform (reference
 * encoder_h264.c:711-724). */
static void luma_dc_hadamard(const int in[4][4], int out[16])
{
    int t[16];
    for (int i = 0; i < 4; i++) {
        const int s01 = in[i][0] + in[i][1], d01 = in[i][0] - in[i][1];
        const int s23 = in[i][2] + in[i][3], d23 = in[i][2] - in[i][3];
        t[i] = s01 + s23;
        t[4 + i] = s01 - s23;
        t[8 + i] = d01 - d23;
        t[12 + i] = d01 + d23;
    }
    for (int i = 0; i < 4; i++) {
        const int s01 = t[4 * i] + t[4 * i + 1], d01 = t[4 * i] - t[4 * i + 1];
        const int s23 = t[4 * i + 2] + t[4 * i + 3], d23 = t[4 * i + 2] - t[4 * i + 3];
        out[4 * i] = (s01 + s23 + 1) >> 1;
        out[4 * i + 1] = (s01 - s23 + 1) >> 1;
        out[4 * i + 2] = (d01 - d23 + 1) >> 1;
        out[4 * i + 3] = (d01 + d23 + 1) >> 1;
    }
}

void h264_intra16_luma_dc_transform(const int dc_in[4][4], int qp, int dc_out[16], int dc_out_pretranspose[16])
{
    int raw[16], q[16];
    luma_dc_hadamard(dc_in, raw);
    for (int i = 0; i < 16; i++) q[i] = quant_dc(raw[i], qp);
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) dc_out[r * 4 + c] = q[c * 4 + r];
    if (dc_out_pretranspose) memcpy(dc_out_pretranspose, q, sizeof(q));
}

void h264_chroma_dc_levels(const int c[4], int qpc, int out[4])
{
    const int h[4] = {c[0] + c[1] + c[2] + c[3], c[0] - c[1] + c[2] - c[3], c[0] + c[1] - c[2] - c[3],
                      c[0] - c[1] - c[2] + c[3]};
    for (int i = 0; i < 4; i++) out[i] = quant_dc(h[i], qpc);
}

void h264_frame_dc_levels(const int32_t *dc, int nmb, int qp, int is_intra, int16_t *luma_dc, int16_t *chroma_dc)
{
    const int qpc = h264_chroma_qp(qp);
    for (int m = 0; m < nmb; m++) {
        const int32_t *d = dc + (size_t)m * 24;
        if (luma_dc) {
            if (is_intra) {
                int in[4][4], out[16];
                for (int b = 0; b < 16; b++) in[b >> 2][b & 3] = d[b];
                h264_intra16_luma_dc_transform(in, qp, out, NULL);
                for (int b = 0; b < 16; b++) luma_dc[(size_t)m * 16 + b] = (int16_t)out[b];
            } else {
                memset(luma_dc + (size_t)m * 16, 0, 16 * sizeof(int16_t));
            }
        }
        if (chroma_dc) {
            for (int comp = 0; comp < 2; comp++) {
                int in[4], out[4];
                for (int s = 0; s < 4; s++) in[s] = d[16 + comp * 4 + s];
                h264_chroma_dc_levels(in, qpc, out);
                for (int s = 0; s < 4; s++) chroma_dc[(size_t)m * 8 + comp * 4 + s] = (int16_t)out[s];
            }
        }
    }
}

static double now_ms(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec * 1e3 + t.tv_nsec * 1e-6;
}

int bc250_encode_sequence(gpu_context_t *ctx, const gpu_image_t *frames, int n, int width, int height,
                          const bc250_seq_cfg_t *cfg, double *frame_ms, double *total_ms)
{
    return bc250_encode_sequence_cb(ctx, frames, n, width, height, cfg, frame_ms, total_ms, NULL, NULL);
}

int bc250_nal_sink_frame(void *user, int frame, int slot, int is_intra)
{
    bc250_nal_sink_t *k = (bc250_nal_sink_t *)user;
    const uint8_t *rbsp = NULL;
    const uint32_t *info = NULL;
    int ns = 0;
    if (!k || !k->buf || gpu_compute_get_slices_slot(k->ctx, slot, &rbsp, &info, &ns) != 0) return -1;
    const size_t n = bc250_h264_frame_nals(rbsp, info, ns, is_intra, k->buf + k->used, k->cap - k->used);
    if (n == 0) return -1;
    k->used += n;
    if (k->frame_bytes) k->frame_bytes[frame] = n;
    return 0;
}

/* The driver loop; qp_fn (optional) picks each picture's QP at its
 * submission, else cfg->qp. */
typedef int (*seq_qp_fn)(void *user, int frame, int is_intra);
static int seq_run(gpu_context_t *ctx, const gpu_image_t *frames, int n, int width, int height,
                   const bc250_seq_cfg_t *cfg, double *frame_ms, double *total_ms, bc250_frame_cb on_frame,
                   void *user, seq_qp_fn qp_fn, void *qp_user)
{
    if (!ctx || !frames || n <= 0 || !cfg) return -1;
    int slot_intra[2] = {0, 0};
    double t_submit[2] = {0, 0};
    int slot_frame[2] = {-1, -1};
    const double t0 = now_ms();
    int rc = 0;
    for (int i = 0; i < n && rc == 0; i++) {
        const int is_intra = i == 0 || (cfg->gop > 0 && i % cfg->gop == 0);
        if (gpu_compute_begin_picture(ctx, frames[i]) != 0) { rc = -1; break; }
        const double ts = now_ms();
        const int qp = qp_fn ? qp_fn(qp_user, i, is_intra) : cfg->qp;
        if (gpu_compute_dispatch_encode_ext(ctx, frames[i], width, height, qp, is_intra, cfg->num_slices, 0,
                                            NULL) != 0 ||
            gpu_compute_end_picture(ctx) != 0) {
            rc = -1;
            break;
        }
        const int slot = gpu_compute_submitted_slot(ctx);
        t_submit[slot] = ts;
        slot_frame[slot] = i;
        slot_intra[slot] = is_intra;
        /* pipelined: finish the frame submitted before this one; else this one */
        const int fin = cfg->pipelined ? slot ^ 1 : slot;
        if (slot_frame[fin] >= 0) {
            if (gpu_compute_sync_slot(ctx, fin) != 0) rc = -1;
            if (frame_ms) frame_ms[slot_frame[fin]] = now_ms() - t_submit[fin];
            if (rc == 0 && on_frame && on_frame(user, slot_frame[fin], fin, slot_intra[fin]) != 0) rc = -1;
            slot_frame[fin] = -1;
        }
    }
    /* the last frames, oldest first */
    for (int k = 0; k < 2 && rc == 0; k++) {
        int s = -1;
        for (int t = 0; t < 2; t++)
            if (slot_frame[t] >= 0 && (s < 0 || slot_frame[t] < slot_frame[s])) s = t;
        if (s < 0) break;
        if (gpu_compute_sync_slot(ctx, s) != 0) rc = -1;
        if (frame_ms) frame_ms[slot_frame[s]] = now_ms() - t_submit[s];
        if (rc == 0 && on_frame && on_frame(user, slot_frame[s], s, slot_intra[s]) != 0) rc = -1;
        slot_frame[s] = -1;
    }
    if (total_ms) *total_ms = now_ms() - t0;
    return rc;
}

int bc250_encode_sequence_cb(gpu_context_t *ctx, const gpu_image_t *frames, int n, int width, int height,
                             const bc250_seq_cfg_t *cfg, double *frame_ms, double *total_ms, bc250_frame_cb on_frame,
                             void *user)
{
    return seq_run(ctx, frames, n, width, height, cfg, frame_ms, total_ms, on_frame, user, NULL, NULL);
}

/* ---- rate-controlled sequences */

typedef struct {
    bc250_nal_sink_t *sink;
    bc250_rc_t *rc;
    int *qp;
    int *bits;
} rc_run_t;

static int rc_pick_qp(void *u, int frame, int is_intra)
{
    rc_run_t *r = (rc_run_t *)u;
    int qp = r->rc->model ? bc250_rc_model_frame_qp(r->rc, is_intra, 0) : bc250_rc_get_frame_qp(r->rc, 0);
    /* the encoder's own clamp (encoder_h264.c:2265-2269) */
    if (qp < r->rc->qp_min) qp = r->rc->qp_min;
    if (qp > 51) qp = 51;
    r->qp[frame] = qp;
    return qp;
}

static int rc_frame_done(void *u, int frame, int slot, int is_intra)
{
    rc_run_t *r = (rc_run_t *)u;
    const size_t before = r->sink->used;
    if (bc250_nal_sink_frame(r->sink, frame, slot, is_intra) != 0) return -1;
    const int bits = (int)((r->sink->used - before) * 8);
    if (r->bits) r->bits[frame] = bits;
    bc250_rc_update_stats(r->rc, bits);
    if (r->rc->model) bc250_rc_model_frame_coded(r->rc, is_intra, 0, r->qp[frame], bits);
    return 0;
}

int bc250_encode_sequence_rc(gpu_context_t *ctx, const gpu_image_t *frames, int n, int width, int height,
                             const bc250_seq_cfg_t *cfg, bc250_rc_t *rc, bc250_nal_sink_t *sink, int *qp_out,
                             int *bits_out, double *frame_ms, double *total_ms)
{
    if (!rc || !sink || !qp_out) return -1;
    rc_run_t r = {sink, rc, qp_out, bits_out};
    return seq_run(ctx, frames, n, width, height, cfg, frame_ms, total_ms, rc_frame_done, &r, rc_pick_qp, &r);
}

size_t bc250_rbsp_to_ebsp(uint8_t *dst, size_t cap, const uint8_t *src, size_t n)
{
    size_t o = 0;
    int zeros = 0;
    for (size_t i = 0; i < n && o < cap; i++) {
        const uint8_t b = src[i];
        if (zeros >= 2 && b <= 3) {
            dst[o++] = 3;
            zeros = 0;
            if (o >= cap) break;
        }
        dst[o++] = b;
        zeros = b == 0 ? zeros + 1 : 0;
    }
    return o;
}

size_t bc250_h264_frame_nals(const uint8_t *rbsp, const uint32_t *slice_info, int num_slices, int is_idr,
                             uint8_t *dst, size_t cap)
{
    if (!rbsp || !slice_info || !dst || num_slices < 1) return 0;
    size_t o = 0;
    for (int s = 0; s < num_slices; s++) {
        const size_t len = slice_info[4 * s + 1];
        /* worst case: one 0x03 per two input bytes */
        if (o + 5 + len + len / 2 + 1 > cap) return 0;
        dst[o] = dst[o + 1] = dst[o + 2] = 0;
        dst[o + 3] = 1;
        dst[o + 4] = is_idr ? (3 << 5 | 5) : (2 << 5 | 1);
        o += 5 + bc250_rbsp_to_ebsp(dst + o + 5, cap - o - 5, rbsp + slice_info[4 * s], len);
    }
    return o;
}

/* ---- parameter sets: a list of syntax elements, then one bit writer ---- */
enum { EL_U, EL_UE, EL_SE };
typedef struct { int kind, bits; int64_t v; } ps_el;
typedef struct { ps_el e[64]; int n; } ps_list;

static void el(ps_list *l, int kind, int bits, int64_t v)
{
    if (l->n < 64) l->e[l->n++] = (ps_el){kind, bits, v};
}

/* the RBSP of the element list (with rbsp_trailing_bits), then the NAL */
static size_t ps_nal(const ps_list *l, int nal_type, uint8_t *dst, size_t cap)
{
    uint8_t rbsp[256] = {0};
    size_t bit = 0;
#define PUT(nb, val)                                                                   \
    for (int b_ = (nb) - 1; b_ >= 0; b_--, bit++)                                     \
        if (((uint64_t)(val) >> b_) & 1) rbsp[bit >> 3] |= (uint8_t)(0x80 >> (bit & 7));
    for (int i = 0; i < l->n; i++) {
        const ps_el *x = &l->e[i];
        if (x->kind == EL_U) {
            PUT(x->bits, x->v);
        } else {
            const uint64_t k = x->kind == EL_UE ? (uint64_t)x->v : (x->v > 0 ? (uint64_t)(2 * x->v - 1) : (uint64_t)(-2 * x->v));
            const uint64_t cn = k + 1;
            int z = 0;
            while ((cn >> (z + 1)) != 0) z++;
            PUT(z, 0);
            PUT(z + 1, cn);
        }
    }
    PUT(1, 1);
#undef PUT
    const size_t n = (bit + 7) >> 3;
    if (cap < 5 + n + n / 2 + 1) return 0;
    dst[0] = dst[1] = dst[2] = 0;
    dst[3] = 1;
    dst[4] = (uint8_t)(3 << 5 | nal_type);
    return 5 + bc250_rbsp_to_ebsp(dst + 5, cap - 5, rbsp, n);
}

static int in_range(int32_t v, int32_t lo, int32_t hi) { return v >= lo && v <= hi; }

/* The value ranges of 7.4.2.1.1 for the fields this writer emits (and
 * pic_order_cnt_type 1, whose offsets it does not write, refused). */
static int sps_fields_ok(const bc250_h264_sps_t *p)
{
    if (!in_range(p->profile_idc, 0, 255) || !in_range(p->level_idc, 0, 255) || !in_range(p->sps_id, 0, 31)) return 0;
    const int cf = p->profile_idc == 100 ? p->chroma_format_idc : 1;
    if (p->profile_idc == 100 &&
        (!in_range(cf, 0, 3) || !in_range(p->bit_depth_luma, 8, 14) || !in_range(p->bit_depth_chroma, 8, 14)))
        return 0;
    if (!in_range(p->log2_max_frame_num_minus4, 0, 12) || !in_range(p->pic_order_cnt_type, 0, 2) ||
        p->pic_order_cnt_type == 1)
        return 0;
    if (p->pic_order_cnt_type == 0 && !in_range(p->log2_max_poc_lsb_minus4, 0, 12)) return 0;
    if (!in_range(p->max_num_ref_frames, 0, 16) || !in_range(p->width_in_mbs, 1, 1 << 16) ||
        !in_range(p->height_in_mbs, 1, 1 << 16))
        return 0;
    if (p->frame_cropping) {
        /* CropUnitX / CropUnitY of 7.4.2.1.1 (equations 7-19 .. 7-22) */
        const int fm = p->frame_mbs_only ? 1 : 2;
        const int ux = cf == 1 || cf == 2 ? 2 : 1, uy = (cf == 1 ? 2 : 1) * fm;
        if (p->crop_left < 0 || p->crop_right < 0 || p->crop_top < 0 || p->crop_bottom < 0) return 0;
        if ((int64_t)ux * ((int64_t)p->crop_left + p->crop_right) >= 16 * (int64_t)p->width_in_mbs ||
            (int64_t)uy * ((int64_t)p->crop_top + p->crop_bottom) >= 16 * (int64_t)fm * p->height_in_mbs)
            return 0;
    }
    if (p->vui_present) {
        if (!in_range(p->sar_width, 0, 65535) || !in_range(p->sar_height, 0, 65535)) return 0;
        if (p->timing_info_present && (p->num_units_in_tick <= 0 || p->time_scale <= 0)) return 0;
    }
    return 1;
}

/* The value ranges of 7.4.2.2 (pic_init_qp_minus26 down to -(26 + 36), the
 * deepest QpBdOffsetY). */
static int pps_fields_ok(const bc250_h264_pps_t *p)
{
    return in_range(p->pps_id, 0, 255) && in_range(p->sps_id, 0, 31) && in_range(p->num_ref_idx_l0_default, 0, 31) &&
           in_range(p->num_ref_idx_l1_default, 0, 31) && in_range(p->weighted_bipred_idc, 0, 2) &&
           in_range(p->pic_init_qp_minus26, -62, 25) && in_range(p->chroma_qp_offset, -12, 12) &&
           (!p->transform_8x8_mode || in_range(p->second_chroma_qp_offset, -12, 12));
}

size_t bc250_h264_write_sps(const bc250_h264_sps_t *p, uint8_t *dst, size_t cap)
{
    if (!p || !dst || !sps_fields_ok(p)) return 0;
    ps_list l = {.n = 0};
    el(&l, EL_U, 8, p->profile_idc);
    el(&l, EL_U, 8, 0); /* constraint_set0..5_flag, reserved_zero_2bits */
    el(&l, EL_U, 8, p->level_idc);
    el(&l, EL_UE, 0, p->sps_id);
    if (p->profile_idc == 100) {
        el(&l, EL_UE, 0, p->chroma_format_idc);
        el(&l, EL_UE, 0, p->bit_depth_luma - 8);
        el(&l, EL_UE, 0, p->bit_depth_chroma - 8);
        el(&l, EL_U, 2, 0); /* qpprime_y_zero_transform_bypass, seq_scaling_matrix_present */
    }
    el(&l, EL_UE, 0, p->log2_max_frame_num_minus4);
    el(&l, EL_UE, 0, p->pic_order_cnt_type);
    if (p->pic_order_cnt_type == 0) el(&l, EL_UE, 0, p->log2_max_poc_lsb_minus4);
    el(&l, EL_UE, 0, p->max_num_ref_frames);
    el(&l, EL_U, 1, 0); /* gaps_in_frame_num_value_allowed */
    el(&l, EL_UE, 0, p->width_in_mbs - 1);
    el(&l, EL_UE, 0, p->height_in_mbs - 1);
    el(&l, EL_U, 1, p->frame_mbs_only != 0);
    if (!p->frame_mbs_only) el(&l, EL_U, 1, 0); /* mb_adaptive_frame_field */
    el(&l, EL_U, 1, p->direct_8x8_inference != 0);
    el(&l, EL_U, 1, p->frame_cropping != 0);
    if (p->frame_cropping) {
        el(&l, EL_UE, 0, p->crop_left);
        el(&l, EL_UE, 0, p->crop_right);
        el(&l, EL_UE, 0, p->crop_top);
        el(&l, EL_UE, 0, p->crop_bottom);
    }
    el(&l, EL_U, 1, p->vui_present != 0);
    if (p->vui_present) {
        el(&l, EL_U, 1, 1);   /* aspect_ratio_info_present */
        el(&l, EL_U, 8, 255); /* Extended_SAR */
        el(&l, EL_U, 16, p->sar_width);
        el(&l, EL_U, 16, p->sar_height);
        el(&l, EL_U, 3, 0); /* overscan, video_signal_type, chroma_loc_info present flags */
        el(&l, EL_U, 1, p->timing_info_present != 0);
        if (p->timing_info_present) {
            el(&l, EL_U, 32, (uint32_t)p->num_units_in_tick);
            el(&l, EL_U, 32, (uint32_t)p->time_scale);
            el(&l, EL_U, 1, 0); /* fixed_frame_rate */
        }
        el(&l, EL_U, 4, 0); /* nal / vcl hrd, pic_struct, bitstream_restriction flags */
    }
    return ps_nal(&l, 7, dst, cap);
}

size_t bc250_h264_write_pps(const bc250_h264_pps_t *p, uint8_t *dst, size_t cap)
{
    if (!p || !dst || !pps_fields_ok(p)) return 0;
    ps_list l = {.n = 0};
    el(&l, EL_UE, 0, p->pps_id);
    el(&l, EL_UE, 0, p->sps_id);
    el(&l, EL_U, 1, p->entropy_coding_mode != 0);
    el(&l, EL_U, 1, p->pic_order_present != 0);
    el(&l, EL_UE, 0, 0); /* num_slice_groups_minus1 */
    el(&l, EL_UE, 0, p->num_ref_idx_l0_default);
    el(&l, EL_UE, 0, p->num_ref_idx_l1_default);
    el(&l, EL_U, 1, p->weighted_pred != 0);
    el(&l, EL_U, 2, p->weighted_bipred_idc);
    el(&l, EL_SE, 0, p->pic_init_qp_minus26);
    el(&l, EL_SE, 0, 0); /* pic_init_qs_minus26 */
    el(&l, EL_SE, 0, p->chroma_qp_offset);
    el(&l, EL_U, 1, p->deblocking_filter_control != 0);
    el(&l, EL_U, 1, p->constrained_intra_pred != 0);
    el(&l, EL_U, 1, 0); /* redundant_pic_cnt_present */
    if (p->transform_8x8_mode) {
        el(&l, EL_U, 1, 1);
        el(&l, EL_U, 1, 0); /* pic_scaling_matrix_present */
        el(&l, EL_SE, 0, p->second_chroma_qp_offset);
    }
    return ps_nal(&l, 8, dst, cap);
}
