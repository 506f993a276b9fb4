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

#include "bc250_host.h"

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
    if (!ctx || !frames || n <= 0 || !cfg) return -1;
    double t_submit[2] = {0, 0};
    int slot_frame[2] = {-1, -1};
    const double t0 = now_ms();
    int rc = 0;
    for (int i = 0; i < n && rc == 0; i++) {
        const int is_intra = i == 0 || (cfg->gop > 0 && i % cfg->gop == 0);
        if (gpu_compute_begin_picture(ctx, frames[i]) != 0) { rc = -1; break; }
        const double ts = now_ms();
        if (gpu_compute_dispatch_encode_ext(ctx, frames[i], width, height, cfg->qp, is_intra, cfg->num_slices, 0,
                                            NULL) != 0 ||
            gpu_compute_end_picture(ctx) != 0) {
            rc = -1;
            break;
        }
        const int slot = gpu_compute_submitted_slot(ctx);
        t_submit[slot] = ts;
        slot_frame[slot] = i;
        /* pipelined: finish the frame submitted before this one; else this one */
        const int fin = cfg->pipelined ? slot ^ 1 : slot;
        if (slot_frame[fin] >= 0) {
            if (gpu_compute_sync_slot(ctx, fin) != 0) rc = -1;
            if (frame_ms) frame_ms[slot_frame[fin]] = now_ms() - t_submit[fin];
            slot_frame[fin] = -1;
        }
    }
    for (int s = 0; s < 2 && rc == 0; s++)
        if (slot_frame[s] >= 0) {
            if (gpu_compute_sync_slot(ctx, s) != 0) rc = -1;
            if (frame_ms) frame_ms[slot_frame[s]] = now_ms() - t_submit[s];
            slot_frame[s] = -1;
        }
    if (total_ms) *total_ms = now_ms() - t0;
    return rc;
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
