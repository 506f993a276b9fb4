/*
 * orc_hevc.c - CPU oracle for the HEVC stages (TEST INFRASTRUCTURE; only
 * tests/, smoke() and bench.py's cpu_baseline leg load it).
 *
 * Restates, per function, the arithmetic of the reference's HEVC encoder
 * core and of the decoder functions it borrows:
 *   transforms / quant  enc/src/hevc_intra.c:144-795 (4x4 DCT-II / DST-VII,
 *                       8x8 DCT-II, levels, dequant, chroma QP)
 *   intra prediction    enc/src/hevc_intra_template.c:28-213, z-scan
 *                       availability enc/src/hevc_intra.c:75-130
 *   motion compensation enc/src/hevc_mc_template.c:143-256,492-514
 *   I-picture CU coding enc/src/hevc_enc_template.c:502-594 (intra_trial),
 *                       1409-1427 (encode_ctu, I branch)
 *   inter CU residual   enc/src/hevc_enc_template.c:669-680,817-933 with the
 *                       8x8 luma transform only (tu8 without inter_tu4)
 * Planes are separate (Y, Cb, Cr), samples uint8 (8-bit) or uint16 (10-bit),
 * in coded size (width and height multiples of 16), row stride in samples.
 */
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

/* ---------------------------------------------------------------- matrices */
/* H.265 8.6.4.2 transMatrix coefficients (the standard's constants). */
static const int kDct4[4][4] = {{64, 64, 64, 64}, {83, 36, -36, -83}, {64, -64, -64, 64}, {36, -83, 83, -36}};
static const int kDst4[4][4] = {{29, 55, 74, 84}, {74, 74, 0, -74}, {84, -29, -74, 55}, {55, -84, 74, -29}};
static const int kDct8[8][8] = {
    {64, 64, 64, 64, 64, 64, 64, 64},     {89, 75, 50, 18, -18, -50, -75, -89},
    {83, 36, -36, -83, -83, -36, 36, 83}, {75, -18, -89, -50, 50, 89, 18, -75},
    {64, -64, -64, 64, 64, -64, -64, 64}, {50, -89, 18, 75, -75, -18, 89, -50},
    {36, -83, 83, -36, -36, 83, -83, 36}, {18, -50, 75, -89, 89, -75, 50, -18}};
static const int kLevelScale[6] = {40, 45, 51, 57, 64, 72};

static int sat16(int64_t v) { return v > 32767 ? 32767 : (v < -32768 ? -32768 : (int)v); }
static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* Separable transform of an n x n block: first along columns (vertical),
 * then along rows, y = M . x with rounding shift and int16 saturation after
 * each pass (the reference keeps both passes in sixteen bits).  inverse=1
 * applies M transposed.  in/out row-major; out(i,j): i vertical, j
 * horizontal frequency (or spatial row/column for the inverse). */
static void transform2d(const int32_t *in, int n, const int *M, int inverse, int s1, int s2, int32_t *out)
{
    int32_t t[32 * 32];
    for (int i = 0; i < n; i++)
        for (int c = 0; c < n; c++) {
            int64_t s = 0;
            for (int r = 0; r < n; r++) s += (int64_t)(inverse ? M[r * n + i] : M[i * n + r]) * in[r * n + c];
            t[i * n + c] = sat16((s + (1 << (s1 - 1))) >> s1);
        }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            int64_t s = 0;
            for (int c = 0; c < n; c++) s += (int64_t)(inverse ? M[c * n + j] : M[j * n + c]) * t[i * n + c];
            out[i * n + j] = sat16((s + (1 << (s2 - 1))) >> s2);
        }
}

int orc_hevc_chroma_qp(int qp_luma)
{
    static const int mid[14] = {29, 30, 31, 32, 33, 33, 34, 34, 35, 35, 36, 36, 37, 37};  /* Table 8-10, 30..43 */
    const int qpi = clampi(qp_luma, 0, 57);
    return qpi < 30 ? qpi : (qpi > 43 ? qpi - 6 : mid[qpi - 30]);
}

/* level = sign * min(32767, floor(((|v| << shift) + off) / (ls[q%6] << (q/6+4)))),
 * off = floor((ls << (q/6+4)) * round_q12 / 12). */
void orc_hevc_quant(const int32_t *v, int n, int shift, int q, int round_q12, int16_t *out)
{
    const int p = q / 6 + 4;
    const uint64_t d = (uint64_t)kLevelScale[q % 6];
    const uint64_t off = ((d << p) * (uint64_t)round_q12) / 12;
    for (int i = 0; i < n; i++) {
        const uint64_t mag = (uint64_t)(v[i] < 0 ? -(int64_t)v[i] : v[i]);
        uint64_t l = (((mag << shift) + off) >> p) / d;
        if (l > 32767) l = 32767;
        out[i] = (int16_t)(v[i] < 0 ? -(int64_t)l : (int64_t)l);
    }
}

/* 8.6.3 with the flat scaling list: clip16((c * 16 * ls << (q/6) + 2^(b-1)) >> b). */
void orc_hevc_dequant(const int16_t *c, int n, int q, int bd_shift, int16_t *out)
{
    const int64_t scale = (int64_t)kLevelScale[q % 6] << (q / 6 + 4);
    for (int i = 0; i < n; i++) out[i] = (int16_t)sat16(((int64_t)c[i] * scale + (1 << (bd_shift - 1))) >> bd_shift);
}

static int all_zero16(const int16_t *a, int n)
{
    for (int i = 0; i < n; i++)
        if (a[i]) return 0;
    return 1;
}

/* 4x4: forward shifts (1, 8) at 8 bits, (3, 8) at 10; dequant bdShift 5 / 7;
 * inverse (7, 12) / (7, 10).  qp is Qp' (QP + 12 at ten bits); the 8-bit
 * forward path clamps it to 0..51, the 10-bit paths to 0..63, the 8-bit
 * inverse takes it as given (as the reference does). */
void orc_hevc_tq4(const int16_t res[16], int qp, int use_dst, int round_q12, int bd, int16_t out[16])
{
    if (all_zero16(res, 16)) { memset(out, 0, 32); return; }
    int32_t x[16], y[16];
    for (int i = 0; i < 16; i++) x[i] = res[i];
    transform2d(x, 4, use_dst ? &kDst4[0][0] : &kDct4[0][0], 0, bd == 8 ? 1 : 3, 8, y);
    const int q = bd == 8 ? clampi(qp, 0, 51) : clampi(qp, 0, 63);
    orc_hevc_quant(y, 16, bd == 8 ? 5 : 7, q, round_q12, out);
}

void orc_hevc_dqit4(const int16_t coef[16], int qp, int use_dst, int bd, int16_t out[16])
{
    if (all_zero16(coef, 16)) { memset(out, 0, 32); return; }
    int16_t d[16];
    const int q = bd == 8 ? qp : clampi(qp, 0, 63);
    orc_hevc_dequant(coef, 16, q, bd == 8 ? 5 : 7, d);
    int32_t x[16], y[16];
    for (int i = 0; i < 16; i++) x[i] = d[i];
    transform2d(x, 4, use_dst ? &kDst4[0][0] : &kDct4[0][0], 1, 7, 20 - bd, y);
    for (int i = 0; i < 16; i++) out[i] = (int16_t)y[i];
}

/* 8x8 DCT-II: forward shifts (bd - 6, 9), quant shift bd - 2, inverse (7, 20 - bd). */
void orc_hevc_tq8(const int16_t res[64], int qp, int bd, int round_q12, int16_t out[64])
{
    if (all_zero16(res, 64)) { memset(out, 0, 128); return; }
    int32_t x[64], y[64];
    for (int i = 0; i < 64; i++) x[i] = res[i];
    transform2d(x, 8, &kDct8[0][0], 0, bd - 6, 9, y);
    orc_hevc_quant(y, 64, bd - 2, clampi(qp, 0, 63), round_q12, out);
}

void orc_hevc_dqit8(const int16_t coef[64], int qp, int bd, int16_t out[64])
{
    if (all_zero16(coef, 64)) { memset(out, 0, 128); return; }
    int16_t d[64];
    orc_hevc_dequant(coef, 64, clampi(qp, 0, 63), bd - 2, d);
    int32_t x[64], y[64];
    for (int i = 0; i < 64; i++) x[i] = d[i];
    transform2d(x, 8, &kDct8[0][0], 1, 7, 20 - bd, y);
    for (int i = 0; i < 64; i++) out[i] = (int16_t)y[i];
}

/* N-point DCT-II of H.265 8.6.4.2 for N = 4..32: row k of the N-point
 * matrix is row k*32/N of the 32-point one, whose entries are the standard's
 * integer magnitudes of 64*sqrt(2)*cos(pi*b/64) (64 for the first row) with
 * the cosine's sign, b = ((2n+1)k mod 128) folded into 0..32. */
void orc_hevc_dct_matrix(int N, int *M)
{
    static const int odd32[16] = {90, 90, 88, 85, 82, 78, 73, 67, 61, 54, 46, 38, 31, 22, 13, 4};
    static const int odd16[8] = {90, 87, 80, 70, 57, 43, 25, 9};
    static const int odd8[4] = {89, 75, 50, 18};
    static const int odd4[2] = {83, 36};
    int mag[33];
    for (int b = 0; b <= 32; b++) {
        if (b & 1) mag[b] = odd32[b >> 1];
        else if (b & 2) mag[b] = odd16[b >> 2];
        else if (b & 4) mag[b] = odd8[b >> 3];
        else if (b & 8) mag[b] = odd4[b >> 4];
        else mag[b] = 64;  /* b = 0, 16, 32 (32 only as cos(pi/2), never reached) */
    }
    for (int k = 0; k < N; k++)
        for (int n = 0; n < N; n++) {
            const int kk = k * (32 / N);
            int a = ((2 * n + 1) * kk) % 128, sign = 1;
            if (a > 64) a = 128 - a;
            if (a > 32) { a = 64 - a; sign = -1; }
            M[k * N + n] = kk == 0 ? 64 : sign * mag[a];
        }
}

/* N x N DCT-II (log2n 2..5) with the shifts the reference uses at 4 and 8
 * generalised: forward (log2n + bd - 9, log2n + 6), quant shift bd + log2n -
 * 5, dequant bdShift bd + log2n - 5, inverse (7, 20 - bd). */
void orc_hevc_tqn(const int16_t *res, int log2n, int qp, int bd, int round_q12, int16_t *out)
{
    const int N = 1 << log2n, n2 = N * N;
    static int M[32 * 32];
    orc_hevc_dct_matrix(N, M);
    int32_t x[1024], y[1024];
    for (int i = 0; i < n2; i++) x[i] = res[i];
    transform2d(x, N, M, 0, log2n + bd - 9, log2n + 6, y);
    orc_hevc_quant(y, n2, bd + log2n - 5, clampi(qp, 0, bd == 8 ? 51 : 63), round_q12, out);
}

void orc_hevc_dqitn(const int16_t *coef, int log2n, int qp, int bd, int16_t *out)
{
    const int N = 1 << log2n, n2 = N * N;
    static int M[32 * 32];
    orc_hevc_dct_matrix(N, M);
    int16_t d[1024];
    orc_hevc_dequant(coef, n2, clampi(qp, 0, 63), bd + log2n - 5, d);
    int32_t x[1024], y[1024];
    for (int i = 0; i < n2; i++) x[i] = d[i];
    transform2d(x, N, M, 1, 7, 20 - bd, y);
    for (int i = 0; i < n2; i++) out[i] = (int16_t)y[i];
}

/* ---------------------------------------------------- motion compensation */
/* H.265 8.5.3.3.3 interpolation filters (the standard's constants). */
static const int kQpel[4][8] = {
    {0, 0, 0, 0, 0, 0, 0, 0}, {-1, 4, -10, 58, 17, -5, 1, 0}, {-1, 4, -11, 40, 40, -11, 4, -1}, {0, 1, -5, 17, 58, -10, 4, -1}};
static const int kEpel[8][4] = {{0, 0, 0, 0},    {-2, 58, 10, -2}, {-4, 54, 16, -2}, {-6, 46, 28, -4},
                                {-4, 36, 36, -4}, {-4, 28, 46, -6}, {-2, 16, 54, -4}, {-2, 10, 58, -2}};

static int px_at(const void *p, int bd, int stride, int w, int h, int x, int y)
{
    x = clampi(x, 0, w - 1);
    y = clampi(y, 0, h - 1);
    return bd == 8 ? ((const uint8_t *)p)[(size_t)y * stride + x] : ((const uint16_t *)p)[(size_t)y * stride + x];
}

/* One w x h block at (x, y) displaced by (mvx, mvy) (quarter luma / eighth
 * chroma samples), edge samples repeated; 14-bit intermediates stored as
 * int16 between the passes; then (v + round) >> (14 - bd), clipped. */
void orc_hevc_mc_uni(const void *ref, int stride, int w_pic, int h_pic, int x, int y, int w, int h, int mvx,
                     int mvy, int chroma, int bd, void *dst, int dst_stride)
{
    const int steps = chroma ? 3 : 2, ntap = chroma ? 4 : 8, back = chroma ? 1 : 3;
    const int fx = mvx & ((1 << steps) - 1), fy = mvy & ((1 << steps) - 1);
    const int ix = x + (mvx >> steps), iy = y + (mvy >> steps);
    const int *fh = chroma ? kEpel[fx] : kQpel[fx];
    const int *fv = chroma ? kEpel[fy] : kQpel[fy];
    const int maxv = (1 << bd) - 1;
    for (int r = 0; r < h; r++)
        for (int c = 0; c < w; c++) {
            int v;
            if (!fx && !fy) {
                v = px_at(ref, bd, stride, w_pic, h_pic, ix + c, iy + r);
            } else {
                int16_t p14;
                if (!fy) {
                    int s = 0;
                    for (int k = 0; k < ntap; k++) s += fh[k] * px_at(ref, bd, stride, w_pic, h_pic, ix + c + k - back, iy + r);
                    p14 = (int16_t)(s >> (bd - 8));
                } else if (!fx) {
                    int s = 0;
                    for (int k = 0; k < ntap; k++) s += fv[k] * px_at(ref, bd, stride, w_pic, h_pic, ix + c, iy + r + k - back);
                    p14 = (int16_t)(s >> (bd - 8));
                } else {
                    int s = 0;
                    for (int k = 0; k < ntap; k++) {
                        int t = 0;
                        for (int j = 0; j < ntap; j++)
                            t += fh[j] * px_at(ref, bd, stride, w_pic, h_pic, ix + c + j - back, iy + r + k - back);
                        s += fv[k] * (int16_t)(t >> (bd - 8));
                    }
                    p14 = (int16_t)(s >> 6);
                }
                v = clampi((p14 + (1 << (13 - bd))) >> (14 - bd), 0, maxv);
            }
            if (bd == 8) ((uint8_t *)dst)[(size_t)r * dst_stride + c] = (uint8_t)v;
            else ((uint16_t *)dst)[(size_t)r * dst_stride + c] = (uint16_t)v;
        }
}

/* --------------------------------------------------------- intra 4x4 */
/* z-scan rank of the 4x4 block holding sample (x, y) (6.5.2): CTUs of
 * 1 << ctu_log2 luma samples (16 or 32; chroma CTUs half that) in raster
 * order, inside a CTU the Morton order of the 4x4 blocks (y bit above x bit
 * at every level).  ctu_log2 4 is the reference's fixed shape CTU16 -> 4 CU8
 * -> 4 PU4 (hevc_intra.c:75-130). */
long long orc_hevc_zrank(int x, int y, int width, int is_luma, int ctu_log2)
{
    const int lg = is_luma ? ctu_log2 : ctu_log2 - 1, cs = 1 << lg;
    const int wctu = (width + cs - 1) / cs;
    const int rx = (x % cs) >> 2, ry = (y % cs) >> 2;
    long long r = (long long)(y / cs) * wctu + x / cs;
    for (int b = lg - 3; b >= 0; b--) r = r * 4 + ((ry >> b) & 1) * 2 + ((rx >> b) & 1);
    return r;
}

static int zavail(int nx, int ny, int width, int height, int is_luma, int y_min, long long cur, int ctu_log2)
{
    if (nx < 0 || ny < y_min || nx >= width || ny >= height) return 0;
    return orc_hevc_zrank(nx, ny, width, is_luma, ctu_log2) < cur;
}

/* Neighbours of the 4x4 block at (x0, y0): left[0..4] = p[-1][0..4],
 * top[0..4] = p[0..4][-1], *corner = p[-1][-1], with the substitution scan
 * over (p[-1][3..0], p[-1][-1], p[0..3][-1], p[4][-1]) and p[-1][4] read
 * when available, else copied from left[3]. */
static void gather(const void *plane, int bd, int stride, int width, int height, int x0, int y0, int is_luma,
                   int y_min, int ctu_log2, int left[5], int top[5], int *corner)
{
    const long long cur = orc_hevc_zrank(x0, y0, width, is_luma, ctu_log2);
    const int aL = zavail(x0 - 1, y0, width, height, is_luma, y_min, cur, ctu_log2);
    const int aT = zavail(x0, y0 - 1, width, height, is_luma, y_min, cur, ctu_log2);
    const int aC = zavail(x0 - 1, y0 - 1, width, height, is_luma, y_min, cur, ctu_log2);
    const int aTR = zavail(x0 + 4, y0 - 1, width, height, is_luma, y_min, cur, ctu_log2);
    int v[10], a[10];
    for (int i = 0; i < 4; i++) {
        a[i] = aL;
        v[i] = aL ? px_at(plane, bd, stride, width, height, x0 - 1, y0 + 3 - i) : 0;
        a[5 + i] = aT;
        v[5 + i] = aT ? px_at(plane, bd, stride, width, height, x0 + i, y0 - 1) : 0;
    }
    a[4] = aC;
    v[4] = aC ? px_at(plane, bd, stride, width, height, x0 - 1, y0 - 1) : 0;
    a[9] = aTR;
    v[9] = aTR ? px_at(plane, bd, stride, width, height, x0 + 4, y0 - 1) : 0;
    int first = -1;
    for (int i = 0; i < 10 && first < 0; i++)
        if (a[i]) first = i;
    if (first < 0) {
        for (int i = 0; i < 10; i++) v[i] = 1 << (bd - 1);
    } else {
        for (int i = 0; i < first; i++) v[i] = v[first];
        for (int i = first + 1; i < 10; i++)
            if (!a[i]) v[i] = v[i - 1];
    }
    for (int i = 0; i < 4; i++) {
        left[i] = v[3 - i];
        top[i] = v[5 + i];
    }
    *corner = v[4];
    top[4] = v[9];
    left[4] = zavail(x0 - 1, y0 + 4, width, height, is_luma, y_min, cur, ctu_log2)
                  ? px_at(plane, bd, stride, width, height, x0 - 1, y0 + 4)
                  : left[3];
}

/* 8.4.4.2.5-7 for Planar (0), DC (1), Horizontal (10), Vertical (26); luma
 * gets the DC / H / V boundary filters. */
static void pred_refs(const int left[5], const int top[5], int corner, int mode, int is_luma, int bd, int out[16])
{
    const int maxv = (1 << bd) - 1;
    if (mode == 0) {
        for (int y = 0; y < 4; y++)
            for (int x = 0; x < 4; x++)
                out[y * 4 + x] = ((3 - x) * left[y] + (x + 1) * top[4] + (3 - y) * top[x] + (y + 1) * left[4] + 4) >> 3;
    } else if (mode == 1) {
        const int dc = (left[0] + left[1] + left[2] + left[3] + top[0] + top[1] + top[2] + top[3] + 4) >> 3;
        for (int i = 0; i < 16; i++) out[i] = dc;
        if (is_luma) {
            out[0] = (left[0] + 2 * dc + top[0] + 2) >> 2;
            for (int k = 1; k < 4; k++) {
                out[k] = (top[k] + 3 * dc + 2) >> 2;
                out[k * 4] = (left[k] + 3 * dc + 2) >> 2;
            }
        }
    } else if (mode == 10) {
        for (int y = 0; y < 4; y++)
            for (int x = 0; x < 4; x++) out[y * 4 + x] = left[y];
        if (is_luma)
            for (int x = 0; x < 4; x++) out[x] = clampi(left[0] + ((top[x] - corner) >> 1), 0, maxv);
    } else if (mode == 26) {
        for (int y = 0; y < 4; y++)
            for (int x = 0; x < 4; x++) out[y * 4 + x] = top[x];
        if (is_luma)
            for (int y = 0; y < 4; y++) out[y * 4] = clampi(top[0] + ((left[y] - corner) >> 1), 0, maxv);
    } else {
        for (int i = 0; i < 16; i++) out[i] = 1 << (bd - 1);
    }
}

static void predict4(const void *plane, int bd, int stride, int width, int height, int x0, int y0, int mode,
                     int is_luma, int y_min, int ctu_log2, int32_t pred[16])
{
    int l[5], t[5], c, o[16];
    gather(plane, bd, stride, width, height, x0, y0, is_luma, y_min, ctu_log2, l, t, &c);
    pred_refs(l, t, c, mode, is_luma, bd, o);
    for (int i = 0; i < 16; i++) pred[i] = o[i];
}

void orc_hevc_predict4(const void *plane, int bd, int stride, int width, int height, int x0, int y0, int mode,
                       int is_luma, int y_min, int32_t pred[16])
{
    predict4(plane, bd, stride, width, height, x0, y0, mode, is_luma, y_min, 4, pred);
}

/* SAD mode decision in the order Planar, DC, H, V with strict '<'. */
static int choose_mode(const void *src, const void *recon, int bd, int stride, int width, int height, int x0,
                       int y0, int y_min, int ctu_log2, int32_t pred[16])
{
    static const int modes[4] = {0, 1, 10, 26};
    int l[5], t[5], c;
    gather(recon, bd, stride, width, height, x0, y0, 1, y_min, ctu_log2, l, t, &c);
    int best = -1, best_sad = 0;
    for (int m = 0; m < 4; m++) {
        int o[16], sad = 0;
        pred_refs(l, t, c, modes[m], 1, bd, o);
        for (int i = 0; i < 16; i++) sad += abs(px_at(src, bd, stride, width, height, x0 + (i & 3), y0 + (i >> 2)) - o[i]);
        if (best < 0 || sad < best_sad) {
            best = modes[m];
            best_sad = sad;
            for (int i = 0; i < 16; i++) pred[i] = o[i];
        }
    }
    return best;
}

int orc_hevc_choose_mode(const void *src, const void *recon, int bd, int stride, int width, int height, int x0,
                         int y0, int y_min, int32_t pred[16])
{
    return choose_mode(src, recon, bd, stride, width, height, x0, y0, y_min, 4, pred);
}

static void put(void *p, int bd, int stride, int x, int y, int v)
{
    if (bd == 8) ((uint8_t *)p)[(size_t)y * stride + x] = (uint8_t)v;
    else ((uint16_t *)p)[(size_t)y * stride + x] = (uint16_t)v;
}

static int get(const void *p, int bd, int stride, int x, int y)
{
    return bd == 8 ? ((const uint8_t *)p)[(size_t)y * stride + x] : ((const uint16_t *)p)[(size_t)y * stride + x];
}

/* first luma row of the slice containing CTU row r: slices are CTU-row
 * ranges [floor(rows*s/ns), floor(rows*(s+1)/ns)). */
int orc_hevc_slice_y_min(int ctu_row, int ctu_rows, int ns)
{
    if (ns < 1) ns = 1;
    if (ns > ctu_rows) ns = ctu_rows;
    for (int s = ns - 1; s >= 0; s--) {
        const int r0 = (int)(((long long)ctu_rows * s) / ns);
        if (ctu_row >= r0) return r0 * 16;
    }
    return 0;
}

/* First luma row of the slice containing CTU16 row r16 when slices are
 * ranges of CTUs of 1 << ctu_log2 luma rows (16 or 32), split as above. */
int orc_hevc_slice_y_min_ctu(int r16, int rows16, int ns, int ctu_log2)
{
    const int s = ctu_log2 - 4, rows = (rows16 + (1 << s) - 1) >> s;
    return orc_hevc_slice_y_min(r16 >> s, rows, ns) << s;
}

/* I picture: CTUs of 1 << ctu_log2 (16 or 32) luma samples in raster order,
 * a CTU32 split into its four 16x16 quadrants in z-order (those inside the
 * picture), every 16x16 into 4 CU8, every luma CU into 4 PU4 (intra, DST, SAD
 * mode choice), chroma one 4x4 DC block per CU and component (DCT, QpC),
 * neighbours available by the z-scan order of that CTU size.  Outputs per
 * 16x16 area in raster order (z-order inside): modes[16] (PU z-order),
 * coef[384] = 16 luma PUs x 16, then 4 Cb CUs x 16, then 4 Cr CUs x 16 (each
 * 4x4 row-major); recon planes written in place (start them as any value). */
void orc_hevc_encode_i_ctu(const void *sy, const void *scb, const void *scr, void *ry, void *rcb, void *rcr, int bd,
                           int wc, int hc, int qp, int num_slices, int round_q12, int ctu_log2, uint8_t *modes,
                           int16_t *coef)
{
    static const int ox[4] = {0, 4, 0, 4}, oy[4] = {0, 0, 4, 4};
    const int wctu = wc / 16, hctu = hc / 16, cwc = wc / 2, chc = hc / 2, maxv = (1 << bd) - 1;
    const int qoff = bd == 8 ? 0 : 12;
    const int cqp = orc_hevc_chroma_qp(qp) + qoff;
    const int sub = ctu_log2 == 5 ? 2 : 1;  /* 16x16 areas per CTU side */
    for (int cty = 0; cty * sub < hctu; cty++)
        for (int ctx = 0; ctx * sub < wctu; ctx++)
            for (int q = 0; q < sub * sub; q++) {
                const int tx = ctx * sub + (q & 1), ty = cty * sub + (q >> 1);
                if (tx >= wctu || ty >= hctu) continue;
                const int y_min = orc_hevc_slice_y_min_ctu(ty, hctu, num_slices, ctu_log2);
                const int t = ty * wctu + tx;
                for (int cu = 0; cu < 4; cu++) {
                    const int cx = tx * 16 + ox[cu] * 2, cy = ty * 16 + oy[cu] * 2;
                    for (int pu = 0; pu < 4; pu++) {
                        const int px = cx + ox[pu], py = cy + oy[pu], k = cu * 4 + pu;
                        int32_t pred[16];
                        modes[t * 16 + k] = (uint8_t)choose_mode(sy, ry, bd, wc, wc, hc, px, py, y_min, ctu_log2, pred);
                        int16_t res[16], rr[16], *lv = coef + (size_t)t * 384 + k * 16;
                        for (int i = 0; i < 16; i++) res[i] = (int16_t)(get(sy, bd, wc, px + (i & 3), py + (i >> 2)) - pred[i]);
                        orc_hevc_tq4(res, qp + qoff, 1, round_q12, bd, lv);
                        orc_hevc_dqit4(lv, qp + qoff, 1, bd, rr);
                        for (int i = 0; i < 16; i++) put(ry, bd, wc, px + (i & 3), py + (i >> 2), clampi(pred[i] + rr[i], 0, maxv));
                    }
                    const int ux = cx / 2, uy = cy / 2;
                    for (int comp = 0; comp < 2; comp++) {
                        const void *s = comp ? scr : scb;
                        void *r = comp ? rcr : rcb;
                        int32_t pred[16];
                        predict4(r, bd, cwc, cwc, chc, ux, uy, 1, 0, y_min / 2, ctu_log2, pred);
                        int16_t res[16], rr[16], *lv = coef + (size_t)t * 384 + 256 + comp * 64 + cu * 16;
                        for (int i = 0; i < 16; i++) res[i] = (int16_t)(get(s, bd, cwc, ux + (i & 3), uy + (i >> 2)) - pred[i]);
                        orc_hevc_tq4(res, cqp, 0, round_q12, bd, lv);
                        orc_hevc_dqit4(lv, cqp, 0, bd, rr);
                        for (int i = 0; i < 16; i++) put(r, bd, cwc, ux + (i & 3), uy + (i >> 2), clampi(pred[i] + rr[i], 0, maxv));
                    }
                }
            }
}

void orc_hevc_encode_i(const void *sy, const void *scb, const void *scr, void *ry, void *rcb, void *rcr, int bd,
                       int wc, int hc, int qp, int num_slices, int round_q12, uint8_t *modes, int16_t *coef)
{
    orc_hevc_encode_i_ctu(sy, scb, scr, ry, rcb, rcr, bd, wc, hc, qp, num_slices, round_q12, 4, modes, coef);
}

/* P picture, every CU8 inter with the given vector (mv[cu] quarter luma
 * samples, CUs indexed [ctu][z-order cu]): luma 8x8 DCT, chroma 4x4 DCT at
 * QpC, both with round_q12.  coef per CTU: 4 CUs x 64 luma (8x8 row-major),
 * then 4 Cb x 16, then 4 Cr x 16.  Recon planes get the reconstruction. */
void orc_hevc_encode_p(const void *sy, const void *scb, const void *scr, const void *fy, const void *fcb,
                       const void *fcr, void *ry, void *rcb, void *rcr, int bd, int wc, int hc, int qp,
                       int round_q12, const int32_t *mv, int16_t *coef)
{
    static const int ox[4] = {0, 8, 0, 8}, oy[4] = {0, 0, 8, 8};
    const int wctu = wc / 16, hctu = hc / 16, cwc = wc / 2, chc = hc / 2, maxv = (1 << bd) - 1;
    const int qoff = bd == 8 ? 0 : 12;
    const int cqp = orc_hevc_chroma_qp(qp) + qoff;
    uint16_t pbuf[64];
    for (int t = 0; t < wctu * hctu; t++)
        for (int cu = 0; cu < 4; cu++) {
            const int cx = (t % wctu) * 16 + ox[cu], cy = (t / wctu) * 16 + oy[cu];
            const int mvx = mv[(t * 4 + cu) * 2], mvy = mv[(t * 4 + cu) * 2 + 1];
            int16_t res[64], rr[64], *lv = coef + (size_t)t * 384 + cu * 64;
            orc_hevc_mc_uni(fy, wc, wc, hc, cx, cy, 8, 8, mvx, mvy, 0, bd, pbuf, 8);
            for (int i = 0; i < 64; i++) {
                const int p = bd == 8 ? ((uint8_t *)pbuf)[i] : pbuf[i];
                res[i] = (int16_t)(get(sy, bd, wc, cx + (i & 7), cy + (i >> 3)) - p);
            }
            orc_hevc_tq8(res, qp + qoff, bd, round_q12, lv);
            orc_hevc_dqit8(lv, qp + qoff, bd, rr);
            for (int i = 0; i < 64; i++) {
                const int p = bd == 8 ? ((uint8_t *)pbuf)[i] : pbuf[i];
                put(ry, bd, wc, cx + (i & 7), cy + (i >> 3), clampi(p + rr[i], 0, maxv));
            }
            for (int comp = 0; comp < 2; comp++) {
                const void *s = comp ? scr : scb, *f = comp ? fcr : fcb;
                void *r = comp ? rcr : rcb;
                int16_t cres[16], crr[16], *clv = coef + (size_t)t * 384 + 256 + comp * 64 + cu * 16;
                orc_hevc_mc_uni(f, cwc, cwc, chc, cx / 2, cy / 2, 4, 4, mvx, mvy, 1, bd, pbuf, 4);
                for (int i = 0; i < 16; i++) {
                    const int p = bd == 8 ? ((uint8_t *)pbuf)[i] : pbuf[i];
                    cres[i] = (int16_t)(get(s, bd, cwc, cx / 2 + (i & 3), cy / 2 + (i >> 2)) - p);
                }
                orc_hevc_tq4(cres, cqp, 0, round_q12, bd, clv);
                orc_hevc_dqit4(clv, cqp, 0, bd, crr);
                for (int i = 0; i < 16; i++) {
                    const int p = bd == 8 ? ((uint8_t *)pbuf)[i] : pbuf[i];
                    put(r, bd, cwc, cx / 2 + (i & 3), cy / 2 + (i >> 2), clampi(p + crr[i], 0, maxv));
                }
            }
        }
}

/* One inter CU of 16x16 or 32x32 luma (log2n 4 or 5) at (x0, y0), vector
 * (mvx, mvy): one luma TU of its size and one chroma TU of half its size per
 * component (N-point DCT-II, quant / dequant as orc_hevc_tqn / dqitn).  Its
 * levels go to the 16x16 areas it covers (t0 = the first, wctu areas per
 * row): area q (z-order inside a CU32) holds luma levels [256 q, 256 q + 256)
 * of the TU (row-major) and chroma levels [64 q, 64 q + 64) of each chroma
 * TU, in the per-area slots of orc_hevc_encode_p (luma 0..255, Cb 256..319,
 * Cr 320..383). */
static void code_large_cu(const void *sy, const void *scb, const void *scr, const void *fy, const void *fcb,
                          const void *fcr, void *ry, void *rcb, void *rcr, int bd, int wc, int hc, int qp,
                          int round_q12, int x0, int y0, int log2n, int mvx, int mvy, int16_t *coef)
{
    const int n = 1 << log2n, cn = n / 2, wctu = wc / 16, cwc = wc / 2, chc = hc / 2, maxv = (1 << bd) - 1;
    const int qoff = bd == 8 ? 0 : 12, cqp = orc_hevc_chroma_qp(qp) + qoff;
    static uint16_t pbuf[1024];
    static int16_t res[1024], lv[1024], rr[1024];
    for (int comp = -1; comp < 2; comp++) {
        const int chroma = comp >= 0, m = chroma ? cn : n, lg = chroma ? log2n - 1 : log2n;
        const int bx = chroma ? x0 / 2 : x0, by = chroma ? y0 / 2 : y0, pw = chroma ? cwc : wc;
        const void *s = comp < 0 ? sy : (comp ? scr : scb), *f = comp < 0 ? fy : (comp ? fcr : fcb);
        void *r = comp < 0 ? ry : (comp ? rcr : rcb);
        orc_hevc_mc_uni(f, pw, pw, chroma ? chc : hc, bx, by, m, m, mvx, mvy, chroma, bd, pbuf, m);
        for (int i = 0; i < m * m; i++) {
            const int pv = bd == 8 ? ((uint8_t *)pbuf)[i] : pbuf[i];
            res[i] = (int16_t)(get(s, bd, pw, bx + i % m, by + i / m) - pv);
        }
        const int q = chroma ? cqp : qp + qoff;
        orc_hevc_tqn(res, lg, q, bd, round_q12, lv);
        orc_hevc_dqitn(lv, lg, q, bd, rr);
        for (int i = 0; i < m * m; i++) {
            const int pv = bd == 8 ? ((uint8_t *)pbuf)[i] : pbuf[i];
            put(r, bd, pw, bx + i % m, by + i / m, clampi(pv + rr[i], 0, maxv));
        }
        const int per = chroma ? 64 : 256, base = chroma ? 256 + comp * 64 : 0;
        for (int a = 0; a < (log2n == 5 ? 4 : 1); a++) {
            const int t = (y0 / 16 + (a >> 1)) * wctu + x0 / 16 + (a & 1);
            memcpy(coef + (size_t)t * 384 + base, lv + a * per, (size_t)per * 2);
        }
    }
}

/* P picture at CTU16 (ctu_log2 4: orc_hevc_encode_p) or CTU32 (5).  At
 * CTU32 each CTU32 chooses its CUs from the 8x8 vectors: a 16x16 area whose
 * four CU8 vectors are equal is one CU16, a CTU32 inside the picture whose
 * four areas are CU16s with one vector is one CU32; every other area stays
 * four CU8s.  culog2[4 t + z] (may be NULL) = log2 of the CU holding CU8 z of
 * area t (3, 4 or 5). */
void orc_hevc_encode_p_ctu(const void *sy, const void *scb, const void *scr, const void *fy, const void *fcb,
                           const void *fcr, void *ry, void *rcb, void *rcr, int bd, int wc, int hc, int qp,
                           int round_q12, const int32_t *mv, int ctu_log2, int16_t *coef, uint8_t *culog2)
{
    const int wctu = wc / 16, hctu = hc / 16;
    orc_hevc_encode_p(sy, scb, scr, fy, fcb, fcr, ry, rcb, rcr, bd, wc, hc, qp, round_q12, mv, coef);
    if (culog2) memset(culog2, 3, (size_t)wctu * hctu * 4);
    if (ctu_log2 != 5) return;
    for (int Y = 0; 2 * Y < hctu; Y++)
        for (int X = 0; 2 * X < wctu; X++) {
            int cu16[4], all = 1;
            for (int a = 0; a < 4; a++) {
                const int tx = 2 * X + (a & 1), ty = 2 * Y + (a >> 1);
                cu16[a] = 0;
                if (tx < wctu && ty < hctu) {
                    const int32_t *v = mv + (size_t)(ty * wctu + tx) * 8;
                    cu16[a] = 1;
                    for (int z = 1; z < 4; z++) cu16[a] &= v[2 * z] == v[0] && v[2 * z + 1] == v[1];
                }
                all &= cu16[a];
            }
            const int32_t *v0 = mv + (size_t)(2 * Y * wctu + 2 * X) * 8;
            for (int a = 1; a < 4 && all; a++) {
                const int32_t *v = mv + (size_t)((2 * Y + (a >> 1)) * wctu + 2 * X + (a & 1)) * 8;
                all = v[0] == v0[0] && v[1] == v0[1];
            }
            for (int a = 0; a < 4; a++) {
                const int tx = 2 * X + (a & 1), ty = 2 * Y + (a >> 1), t = ty * wctu + tx;
                if (!cu16[a]) continue;
                if (culog2) memset(culog2 + (size_t)t * 4, all ? 5 : 4, 4);
                if (!all)
                    code_large_cu(sy, scb, scr, fy, fcb, fcr, ry, rcb, rcr, bd, wc, hc, qp, round_q12, tx * 16, ty * 16,
                                  4, mv[(size_t)t * 8], mv[(size_t)t * 8 + 1], coef);
            }
            if (all)
                code_large_cu(sy, scb, scr, fy, fcb, fcr, ry, rcb, rcr, bd, wc, hc, qp, round_q12, X * 32, Y * 32, 5,
                              v0[0], v0[1], coef);
        }
}

/* Quarter-pel refinement of one 8x8 luma block with the HEVC interpolation:
 * centre = the integer vector, then 8 half-sample candidates in the order
 * (2,0) (-2,0) (0,2) (0,-2) (2,2) (2,-2) (-2,2) (-2,-2) around it, then the
 * same 8 quarter-sample offsets (halved) around the half winner, each kept
 * only when its SAD is strictly lower.  Returns the SAD; the vector (quarter samples) is read from
 * and written back to *mvx, *mvy. */
uint32_t orc_hevc_refine8(const void *cur, const void *ref, int bd, int wc, int hc, int x0, int y0, int *mvx,
                          int *mvy)
{
    static const int dx[8] = {2, -2, 0, 0, 2, 2, -2, -2}, dy[8] = {0, 0, 2, -2, 2, -2, 2, -2};
    uint16_t p[64];
    uint32_t best = 0;
    int bx = *mvx, by = *mvy;
    for (int stage = 0; stage < 3; stage++) {
        const int n = stage == 0 ? 1 : 8, cx = bx, cy = by;
        for (int k = 0; k < n; k++) {
            const int mx = stage == 0 ? cx : cx + (stage == 1 ? dx[k] : dx[k] / 2);
            const int my = stage == 0 ? cy : cy + (stage == 1 ? dy[k] : dy[k] / 2);
            orc_hevc_mc_uni(ref, wc, wc, hc, x0, y0, 8, 8, mx, my, 0, bd, p, 8);
            uint32_t sad = 0;
            for (int i = 0; i < 64; i++) {
                const int pv = bd == 8 ? ((uint8_t *)p)[i] : p[i];
                sad += (uint32_t)abs(get(cur, bd, wc, x0 + (i & 7), y0 + (i >> 3)) - pv);
            }
            if (stage == 0 || sad < best) {
                best = sad;
                bx = mx;
                by = my;
            }
        }
    }
    *mvx = bx;
    *mvy = by;
    return best;
}

/* ------------------------------------------------------------------------
 * Deblocking (H.265 8.7.2) for the fixed coding shape of this pipeline: 8x8
 * CUs, each one PU and (P) one 8x8 TU or (I) 4x4 TUs, one reference picture,
 * uni-prediction.  Edges on the 8x8 luma grid; all vertical edges of the
 * picture first, then the horizontal ones on that result; picture borders
 * and slice borders (CTU-row slices, loop filter not across slices) are not
 * filtered.  bS: 2 intra, 1 if either CU has coded luma or the vectors
 * differ by >= 4 quarter samples, else 0 (the reference decoder's
 * strength(), hevc_filter_template.c:221-290).  Chroma edges on the 8x8
 * chroma grid, bS 2 only.  Tables: Table 8-12 (hevcd_beta / hevcd_tc).
 * ---------------------------------------------------------------------- */
static const int kBeta[52] = {0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  6,  7,
                              8,  9,  10, 11, 12, 13, 14, 15, 16, 17, 18, 20, 22, 24, 26, 28, 30, 32,
                              34, 36, 38, 40, 42, 44, 46, 48, 50, 52, 54, 56, 58, 60, 62, 64};
static const int kTc[54] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,  0,  1,  1,  1,  1,  1,  1,  1,  1, 1,
                            2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 5, 5, 6, 6, 7, 8, 9, 10, 11, 13, 14, 16, 18, 20, 22, 24};

int orc_hevc_beta(int q) { return kBeta[clampi(q, 0, 51)]; }
int orc_hevc_tc(int q) { return kTc[clampi(q, 0, 53)]; }

/* Sample access by linear index into a uint8 (bd 8) or uint16 (bd > 8) plane. */
static int gs(const void *p, int bd, ptrdiff_t i) { return bd == 8 ? ((const uint8_t *)p)[i] : ((const uint16_t *)p)[i]; }
static void ps(void *p, int bd, ptrdiff_t i, int v)
{
    if (bd == 8) ((uint8_t *)p)[i] = (uint8_t)v;
    else ((uint16_t *)p)[i] = (uint16_t)v;
}

/* Luma: 4 lines across one edge whose q0 sample of line 0 is at index q;
 * PX(k, i) is sample i on side p (i = 0 nearest the edge) of line k, QX(k, i)
 * the q side.  beta and tc are already scaled to the bit depth. */
static void luma_segment(void *pl, int bd, ptrdiff_t q, ptrdiff_t along, ptrdiff_t across, int beta, int tc)
{
    const int maxv = (1 << bd) - 1;
#define PX(k, i) gs(pl, bd, q + (k) * along - ((i) + 1) * across)
#define QX(k, i) gs(pl, bd, q + (k) * along + (i) * across)
#define SET_P(k, i, v) ps(pl, bd, q + (k) * along - ((i) + 1) * across, (v))
#define SET_Q(k, i, v) ps(pl, bd, q + (k) * along + (i) * across, (v))
    const int dp0 = abs(PX(0, 2) - 2 * PX(0, 1) + PX(0, 0)), dp3 = abs(PX(3, 2) - 2 * PX(3, 1) + PX(3, 0));
    const int dq0 = abs(QX(0, 2) - 2 * QX(0, 1) + QX(0, 0)), dq3 = abs(QX(3, 2) - 2 * QX(3, 1) + QX(3, 0));
    const int d = dp0 + dq0 + dp3 + dq3;
    if (d >= beta) return;
    int strong = 1;
    for (int k = 0; k < 4; k += 3) {
        const int dd = k ? dp3 + dq3 : dp0 + dq0;
        if (!(2 * dd < (beta >> 2) && abs(PX(k, 3) - PX(k, 0)) + abs(QX(k, 0) - QX(k, 3)) < (beta >> 3) &&
              abs(PX(k, 0) - QX(k, 0)) < ((5 * tc + 1) >> 1)))
            strong = 0;
    }
    const int dEp = (dp0 + dp3) < ((beta + (beta >> 1)) >> 3), dEq = (dq0 + dq3) < ((beta + (beta >> 1)) >> 3);
    for (int k = 0; k < 4; k++) {
        const int p0 = PX(k, 0), p1 = PX(k, 1), p2 = PX(k, 2), p3 = PX(k, 3);
        const int q0 = QX(k, 0), q1 = QX(k, 1), q2 = QX(k, 2), q3 = QX(k, 3);
        if (strong) {
            SET_P(k, 0, clampi((p2 + 2 * p1 + 2 * p0 + 2 * q0 + q1 + 4) >> 3, p0 - 2 * tc, p0 + 2 * tc));
            SET_P(k, 1, clampi((p2 + p1 + p0 + q0 + 2) >> 2, p1 - 2 * tc, p1 + 2 * tc));
            SET_P(k, 2, clampi((2 * p3 + 3 * p2 + p1 + p0 + q0 + 4) >> 3, p2 - 2 * tc, p2 + 2 * tc));
            SET_Q(k, 0, clampi((p1 + 2 * p0 + 2 * q0 + 2 * q1 + q2 + 4) >> 3, q0 - 2 * tc, q0 + 2 * tc));
            SET_Q(k, 1, clampi((p0 + q0 + q1 + q2 + 2) >> 2, q1 - 2 * tc, q1 + 2 * tc));
            SET_Q(k, 2, clampi((p0 + q0 + q1 + 3 * q2 + 2 * q3 + 4) >> 3, q2 - 2 * tc, q2 + 2 * tc));
        } else {
            int delta = (9 * (q0 - p0) - 3 * (q1 - p1) + 8) >> 4;
            if (abs(delta) >= tc * 10) continue;
            delta = clampi(delta, -tc, tc);
            SET_P(k, 0, clampi(p0 + delta, 0, maxv));
            SET_Q(k, 0, clampi(q0 - delta, 0, maxv));
            if (dEp) SET_P(k, 1, clampi(p1 + clampi((((p2 + p0 + 1) >> 1) - p1 + delta) >> 1, -(tc >> 1), tc >> 1), 0, maxv));
            if (dEq) SET_Q(k, 1, clampi(q1 + clampi((((q2 + q0 + 1) >> 1) - q1 - delta) >> 1, -(tc >> 1), tc >> 1), 0, maxv));
        }
    }
#undef PX
#undef QX
#undef SET_P
#undef SET_Q
}

static void chroma_segment(void *pl, int bd, ptrdiff_t q, ptrdiff_t along, ptrdiff_t across, int tc)
{
    const int maxv = (1 << bd) - 1;
    for (int k = 0; k < 4; k++) {
        const ptrdiff_t iq = q + k * along, ip = iq - across;
        const int P1 = gs(pl, bd, ip - across), P0 = gs(pl, bd, ip), Q0 = gs(pl, bd, iq), Q1 = gs(pl, bd, iq + across);
        const int delta = clampi((((Q0 - P0) * 4) + P1 - Q1 + 4) >> 3, -tc, tc);
        ps(pl, bd, ip, clampi(P0 + delta, 0, maxv));
        ps(pl, bd, iq, clampi(Q0 - delta, 0, maxv));
    }
}

/* cu index of luma sample (x, y): CTU raster, CU z-order */
static int cu_at(int x, int y, int wctu) { return ((y >> 4) * wctu + (x >> 4)) * 4 + ((y >> 3) & 1) * 2 + ((x >> 3) & 1); }

int orc_hevc_bs(int is_intra, const uint8_t *cbf, const int32_t *mv4, int cp, int cq)
{
    if (is_intra) return 2;
    if (cbf[cp] || cbf[cq]) return 1;
    const int32_t *a = mv4 + 4 * cp, *b = mv4 + 4 * cq;
    return (abs(a[0] - b[0]) >= 4 || abs(a[1] - b[1]) >= 4) ? 1 : 0;
}

/* y: luma (pitch py samples), uv: interleaved Cb/Cr (pitch puv samples),
 * uint8 at bd 8, else uint16 with the sample in the low bits; mv4: {mvx,
 * mvy, ., .} per CU; cbf: per CU.  beta and tC scale by 2^(bd-8) (8.7.2.5.3,
 * 8.7.2.5.5). */
/* Whether the 8x8-grid edge at coordinate e between CU8 p and q lies inside
 * one CU16 / CU32 (neither a TU nor a PU edge). */
static int inside_cu(const uint8_t *culog2, int cp, int cq, int e)
{
    if (!culog2) return 0;
    const int L = culog2[cq];
    return L >= 4 && culog2[cp] == L && (e % (1 << L)) != 0;
}

/* culog2 (may be NULL): log2 CU size per CU8 (orc_hevc_encode_p_ctu); edges
 * inside a CU16 / CU32 are not filtered. */
void orc_hevc_deblock_ctu(void *y, void *uv, int py, int puv, int wc, int hc, int qp, int is_intra,
                          const uint8_t *cbf, const int32_t *mv4, int num_slices, int ctu_log2, int bd,
                          const uint8_t *culog2)
{
    const int wctu = wc / 16, hctu = hc / 16;
    const int beta = kBeta[clampi(qp, 0, 51)] << (bd - 8);
    const int qpc = orc_hevc_chroma_qp(qp);
    for (int dir = 0; dir < 2; dir++) {       /* 0: vertical edges, 1: horizontal edges */
        for (int e = 8; e < (dir ? hc : wc); e += 8) {
            if (dir && (e & 15) == 0 && orc_hevc_slice_y_min_ctu(e >> 4, hctu, num_slices, ctu_log2) == e) continue;
            for (int s = 0; s < (dir ? wc : hc); s += 4) {   /* 4-sample segments along the edge */
                const int x = dir ? s : e, yy = dir ? e : s;
                const int cq = cu_at(x, yy, wctu), cp = cu_at(dir ? x : x - 1, dir ? yy - 1 : yy, wctu);
                const int bs = inside_cu(culog2, cp, cq, e) ? 0 : orc_hevc_bs(is_intra, cbf, mv4, cp, cq);
                if (!bs) continue;
                const int tc = kTc[clampi(qp + 2 * (bs - 1), 0, 53)] << (bd - 8);
                luma_segment(y, bd, (ptrdiff_t)yy * py + x, dir ? 1 : py, dir ? py : 1, beta, tc);
            }
            /* chroma: edges on the 8x8 chroma grid (every 16 luma), bS 2 only */
            if ((e & 15) == 0) {
                const int ctc = kTc[clampi(qpc + 2, 0, 53)] << (bd - 8);
                for (int s = 0; s < (dir ? wc : hc) / 2; s += 4) {
                    const int cx = dir ? s : e / 2, cy = dir ? e / 2 : s;
                    const int lx = 2 * cx, ly = 2 * cy;
                    const int cq = cu_at(lx, ly, wctu), cp = cu_at(dir ? lx : lx - 1, dir ? ly - 1 : ly, wctu);
                    if (inside_cu(culog2, cp, cq, e) || orc_hevc_bs(is_intra, cbf, mv4, cp, cq) != 2) continue;
                    for (int comp = 0; comp < 2; comp++)
                        chroma_segment(uv, bd, (ptrdiff_t)cy * puv + 2 * cx + comp, dir ? 2 : puv, dir ? puv : 2, ctc);
                }
            }
        }
    }
}

void orc_hevc_deblock_bd(void *y, void *uv, int py, int puv, int wc, int hc, int qp, int is_intra,
                         const uint8_t *cbf, const int32_t *mv4, int num_slices, int bd)
{
    orc_hevc_deblock_ctu(y, uv, py, puv, wc, hc, qp, is_intra, cbf, mv4, num_slices, 4, bd, NULL);
}

void orc_hevc_deblock(uint8_t *y, uint8_t *uv, int py, int puv, int wc, int hc, int qp, int is_intra,
                      const uint8_t *cbf, const int32_t *mv4, int num_slices)
{
    orc_hevc_deblock_bd(y, uv, py, puv, wc, hc, qp, is_intra, cbf, mv4, num_slices, 8);
}
