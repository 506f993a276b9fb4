/*
 * orc_h264.c - CPU oracle for the H.264 compute-encoder stages
 * (TEST INFRASTRUCTURE: the checker for the HIP path, never the product).
 *
 * Every function restates the semantics of a reference stage; the citation
 * names the reference file:line it follows (paths relative to the
 * reference's approach1-compute-encoder/).  Deviations from the reference
 * compute path are listed where they occur:
 *   - integer SAD is exact (the shader's float-normalised SAD with
 *     truncation, shaders/motion_estimation.comp:320,357, is not kept);
 *   - integer ME is an exhaustive +-R search (the reference runs a +-16
 *     diamond, motion_estimation.comp:340-374);
 *   - sample clamping uses the coded (16-aligned) frame, which is what a
 *     decoder clamps to (H.264 8.4.2.2: PicWidthInSamples);
 *   - Plane prediction also requires the top-left MB to be available
 *     (H.264 8.3.3.4); the shader only tests top and left
 *     (shaders/intra_wavefront.comp:449-458);
 *   - deblocking is the spec process of H.264 8.7 (luma and chroma, bS=4
 *     strong filter, macroblock raster order) instead of the shader's
 *     luma-only subset (shaders/deblock_filter.comp).
 */
#include "orc.h"

#include <stdlib.h>
#include <string.h>

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline int clip255(int v) { return v < 0 ? 0 : (v > 255 ? 255 : v); }
static inline int iabs(int v) { return v < 0 ? -v : v; }

/* ------------------------------------------------------------------ tables */

/* H.264 Table 8-15, QPc from qPI (shaders/quantize.comp:84-89). */
int orc_chroma_qp(int qpy)
{
    static const int t[22] = {29, 30, 31, 32, 32, 33, 34, 34, 35, 35, 36,
                              36, 37, 37, 37, 38, 38, 38, 39, 39, 39, 39};
    qpy = clampi(qpy, 0, 51);
    return qpy < 30 ? qpy : t[qpy - 30];
}

/* Position class of raster position p: 0 both even, 1 both odd, 2 mixed
 * (shaders/quantize.comp:62-70). */
static inline int pos_class(int p)
{
    int r = (p >> 2) & 1, c = p & 1;
    return (r == 0 && c == 0) ? 0 : ((r == 1 && c == 1) ? 1 : 2);
}

/* Forward multiplier MF (shaders/quantize.comp:53-60) and dequant V
 * (shaders/reconstruct.comp:124-131), both indexed [qp%6][class]. */
static const int MF[6][3] = {{13107, 5243, 8066}, {11916, 4660, 7490},
                             {10082, 4194, 6554}, {9362, 3647, 5825},
                             {8192, 3355, 5243},  {7282, 2893, 4559}};
static const int VQ[6][3] = {{10, 16, 13}, {11, 18, 14}, {13, 20, 16},
                             {14, 23, 18}, {16, 25, 20}, {18, 29, 23}};

/* ------------------------------------------------------------- transforms */

/* H.264 forward core transform, rows then columns
 * (shaders/dct_transform.comp:34-73). */
void orc_fwd4x4(const int32_t in[16], int32_t out[16])
{
    int32_t m[16];
    for (int i = 0; i < 4; i++) {
        const int32_t *p = in + 4 * i;
        int a0 = p[0] + p[3], a1 = p[1] + p[2], a2 = p[1] - p[2], a3 = p[0] - p[3];
        m[4 * i + 0] = a0 + a1;
        m[4 * i + 1] = 2 * a3 + a2;
        m[4 * i + 2] = a0 - a1;
        m[4 * i + 3] = a3 - 2 * a2;
    }
    for (int j = 0; j < 4; j++) {
        int p0 = m[j], p1 = m[4 + j], p2 = m[8 + j], p3 = m[12 + j];
        int a0 = p0 + p3, a1 = p1 + p2, a2 = p1 - p2, a3 = p0 - p3;
        out[j] = a0 + a1;
        out[4 + j] = 2 * a3 + a2;
        out[8 + j] = a0 - a1;
        out[12 + j] = a3 - 2 * a2;
    }
}

/* H.264 8.5.12.2 inverse core transform incl. the final (x+32)>>6
 * (shaders/reconstruct.comp:162-177). */
void orc_inv4x4(const int32_t in[16], int32_t out[16])
{
    int32_t t[16];
    for (int i = 0; i < 4; i++) {
        const int32_t *d = in + 4 * i;
        int z0 = d[0] + d[2], z1 = d[0] - d[2];
        int z2 = (d[1] >> 1) - d[3], z3 = d[1] + (d[3] >> 1);
        t[4 * i + 0] = z0 + z3;
        t[4 * i + 1] = z1 + z2;
        t[4 * i + 2] = z1 - z2;
        t[4 * i + 3] = z0 - z3;
    }
    for (int j = 0; j < 4; j++) {
        int p0 = t[j], p1 = t[4 + j], p2 = t[8 + j], p3 = t[12 + j];
        int z0 = p0 + p2, z1 = p0 - p2, z2 = (p1 >> 1) - p3, z3 = p1 + (p3 >> 1);
        out[j] = (z0 + z3 + 32) >> 6;
        out[4 + j] = (z1 + z2 + 32) >> 6;
        out[8 + j] = (z1 - z2 + 32) >> 6;
        out[12 + j] = (z0 - z3 + 32) >> 6;
    }
}

/* AC quantiser, level = sign*((|c|*MF + f) >> (15+qp/6)), f = 2^(15+qp/6)/3
 * intra or /6 inter (shaders/quantize.comp:105-124). */
int orc_quant_ac(int c, int pos, int qp, int is_intra)
{
    int per = qp / 6, rem = qp % 6;
    int f = (1 << (15 + per)) / (is_intra ? 3 : 6);
    int a = iabs(c);
    int l = (int)(((int64_t)a * MF[rem][pos_class(pos)] + f) >> (15 + per));
    l = c < 0 ? -l : l;
    return clampi(l, -32768, 32767);
}

/* H.264 8.5.12.1 scaling with the flat list (shaders/reconstruct.comp:148-154). */
int orc_dequant_ac(int level, int pos, int qp)
{
    int per = qp / 6, rem = qp % 6;
    int s = 16 * VQ[rem][pos_class(pos)];
    if (per >= 4) return (level * s) << (per - 4);
    return (level * s + (1 << (3 - per))) >> (4 - per);
}

/* Encoder-side luma DC Hadamard with the (x+1)>>1 on the second pass and
 * the implicit transpose of its two-pass butterfly
 * (src/encoder_h264.c:711-724; shaders/reconstruct.comp:180-197). */
void orc_luma_dc_hadamard_fwd(const int32_t in[16], int32_t out[16])
{
    int32_t t[16];
    for (int i = 0; i < 4; i++) {
        int s01 = in[4 * i] + in[4 * i + 1], d01 = in[4 * i] - in[4 * i + 1];
        int s23 = in[4 * i + 2] + in[4 * i + 3], d23 = in[4 * i + 2] - in[4 * i + 3];
        t[0 * 4 + i] = s01 + s23;
        t[1 * 4 + i] = s01 - s23;
        t[2 * 4 + i] = d01 - d23;
        t[3 * 4 + i] = d01 + d23;
    }
    for (int i = 0; i < 4; i++) {
        int s01 = t[4 * i] + t[4 * i + 1], d01 = t[4 * i] - t[4 * i + 1];
        int s23 = t[4 * i + 2] + t[4 * i + 3], d23 = t[4 * i + 2] - t[4 * i + 3];
        out[4 * i + 0] = (s01 + s23 + 1) >> 1;
        out[4 * i + 1] = (s01 - s23 + 1) >> 1;
        out[4 * i + 2] = (d01 - d23 + 1) >> 1;
        out[4 * i + 3] = (d01 + d23 + 1) >> 1;
    }
}

/* Same butterfly without normalisation (shaders/reconstruct.comp:204-216). */
void orc_luma_dc_hadamard_inv(const int32_t in[16], int32_t out[16])
{
    int32_t t[16];
    for (int i = 0; i < 4; i++) {
        int s01 = in[4 * i] + in[4 * i + 1], d01 = in[4 * i] - in[4 * i + 1];
        int s23 = in[4 * i + 2] + in[4 * i + 3], d23 = in[4 * i + 2] - in[4 * i + 3];
        t[0 * 4 + i] = s01 + s23;
        t[1 * 4 + i] = s01 - s23;
        t[2 * 4 + i] = d01 - d23;
        t[3 * 4 + i] = d01 + d23;
    }
    for (int i = 0; i < 4; i++) {
        int s01 = t[4 * i] + t[4 * i + 1], d01 = t[4 * i] - t[4 * i + 1];
        int s23 = t[4 * i + 2] + t[4 * i + 3], d23 = t[4 * i + 2] - t[4 * i + 3];
        out[4 * i + 0] = s01 + s23;
        out[4 * i + 1] = s01 - s23;
        out[4 * i + 2] = d01 - d23;
        out[4 * i + 3] = d01 + d23;
    }
}

/* 2x2 chroma DC Hadamard, self-inverse up to 4 (src/encoder_h264.c:791-797). */
void orc_chroma_dc_hadamard(const int32_t c[4], int32_t o[4])
{
    o[0] = c[0] + c[1] + c[2] + c[3];
    o[1] = c[0] - c[1] + c[2] - c[3];
    o[2] = c[0] + c[1] - c[2] - c[3];
    o[3] = c[0] - c[1] - c[2] + c[3];
}

static const int DC_LS0[6] = {160, 176, 208, 224, 256, 288};

/* Round-to-nearest DC quantiser with threshold 5 for both luma and chroma
 * (src/encoder_h264.c:653-670 and :687-704). */
static int quant_dc_k5(int v, int qp)
{
    int per = qp / 6, rem = qp % 6;
    int64_t ls = DC_LS0[rem], av = v < 0 ? -(int64_t)v : v, num, den;
    if (per >= 5) { num = av; den = ls << (per - 5); }
    else          { num = av << (5 - per); den = ls; }
    int64_t l = (num + den / 2) / den;
    return (int)(v < 0 ? -l : l);
}
int orc_quant_dc_luma(int v, int qp) { return quant_dc_k5(v, qp); }
int orc_quant_dc_chroma(int v, int qpc) { return quant_dc_k5(v, qpc); }

/* H.264 8.5.10 (shaders/reconstruct.comp:280-286). */
int orc_dequant_dc_luma(int f, int qp)
{
    int per = qp / 6, ls = DC_LS0[qp % 6];
    if (per >= 6) return (f * ls) << (per - 6);
    return (f * ls + (1 << (5 - per))) >> (6 - per);
}

/* H.264 8.5.11.2 for 4:2:0 (shaders/reconstruct.comp:288-294). */
int orc_dequant_dc_chroma(int f, int qpc)
{
    int per = qpc / 6, ls = DC_LS0[qpc % 6];
    if (per >= 5) return (f * ls) << (per - 5);
    return (f * ls) >> (5 - per);
}

/* Transmitted I16x16 luma DC levels: Hadamard, quantise, transpose
 * (src/encoder_h264.c:775-787). */
void orc_intra16_luma_dc_levels(const int32_t dc_in[16], int qp, int32_t out[16])
{
    int32_t raw[16], nat[16];
    orc_luma_dc_hadamard_fwd(dc_in, raw);
    for (int i = 0; i < 16; i++) nat[i] = orc_quant_dc_luma(raw[i], qp);
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) out[r * 4 + c] = nat[c * 4 + r];
}

/* ---------------------------------------------------------- interpolation */

static inline int G(const uint8_t *p, int py, int wc, int hc, int x, int y)
{
    return p[(size_t)clampi(y, 0, hc - 1) * py + clampi(x, 0, wc - 1)];
}
static inline int tap6(int a, int b, int c, int d, int e, int f)
{
    return a - 5 * b + 20 * c + 20 * d - 5 * e + f;
}
static inline int hraw(const uint8_t *p, int py, int wc, int hc, int x, int y)
{
    return tap6(G(p, py, wc, hc, x - 2, y), G(p, py, wc, hc, x - 1, y),
                G(p, py, wc, hc, x, y), G(p, py, wc, hc, x + 1, y),
                G(p, py, wc, hc, x + 2, y), G(p, py, wc, hc, x + 3, y));
}
static inline int vraw(const uint8_t *p, int py, int wc, int hc, int x, int y)
{
    return tap6(G(p, py, wc, hc, x, y - 2), G(p, py, wc, hc, x, y - 1),
                G(p, py, wc, hc, x, y), G(p, py, wc, hc, x, y + 1),
                G(p, py, wc, hc, x, y + 2), G(p, py, wc, hc, x, y + 3));
}

/* Luma sample at quarter-pel position (xq, yq) per H.264 8.4.2.2.1, with
 * the letter positions of Figure 8-4
 * (shaders/motion_estimation.comp:180-215). */
int orc_luma_qpel(const uint8_t *ref, int py, int wc, int hc, int xq, int yq)
{
    int x = xq >> 2, y = yq >> 2, fx = xq & 3, fy = yq & 3;
    int g = G(ref, py, wc, hc, x, y);
    if (!fx && !fy) return g;
    int H = G(ref, py, wc, hc, x + 1, y), M = G(ref, py, wc, hc, x, y + 1);
    int b = clip255((hraw(ref, py, wc, hc, x, y) + 16) >> 5);
    int h = clip255((vraw(ref, py, wc, hc, x, y) + 16) >> 5);
    int m = clip255((vraw(ref, py, wc, hc, x + 1, y) + 16) >> 5);
    int s = clip255((hraw(ref, py, wc, hc, x, y + 1) + 16) >> 5);
    int j1 = tap6(hraw(ref, py, wc, hc, x, y - 2), hraw(ref, py, wc, hc, x, y - 1),
                  hraw(ref, py, wc, hc, x, y), hraw(ref, py, wc, hc, x, y + 1),
                  hraw(ref, py, wc, hc, x, y + 2), hraw(ref, py, wc, hc, x, y + 3));
    int j = clip255((j1 + 512) >> 10);
#define AVG(a, c) (((a) + (c) + 1) >> 1)
    switch (fy * 4 + fx) {
    case 1: return AVG(g, b);
    case 2: return b;
    case 3: return AVG(b, H);
    case 4: return AVG(g, h);
    case 5: return AVG(b, h);
    case 6: return AVG(b, j);
    case 7: return AVG(b, m);
    case 8: return h;
    case 9: return AVG(h, j);
    case 10: return j;
    case 11: return AVG(j, m);
    case 12: return AVG(h, M);
    case 13: return AVG(h, s);
    case 14: return AVG(j, s);
    default: return AVG(m, s);
    }
#undef AVG
}

/* Chroma sample at eighth-pel (x8, y8), bilinear per H.264 8.4.2.2.2
 * (shaders/residual_predict.comp:606-631).  comp 0 = Cb, 1 = Cr. */
int orc_chroma_8th(const uint8_t *uv, int puv, int wcc, int hcc, int comp, int x8, int y8)
{
    int x = x8 >> 3, y = y8 >> 3, fx = x8 & 7, fy = y8 & 7;
    int x0 = clampi(x, 0, wcc - 1), x1 = clampi(x + 1, 0, wcc - 1);
    int y0 = clampi(y, 0, hcc - 1), y1 = clampi(y + 1, 0, hcc - 1);
    int A = uv[(size_t)y0 * puv + 2 * x0 + comp], B = uv[(size_t)y0 * puv + 2 * x1 + comp];
    int C = uv[(size_t)y1 * puv + 2 * x0 + comp], D = uv[(size_t)y1 * puv + 2 * x1 + comp];
    return ((8 - fx) * (8 - fy) * A + fx * (8 - fy) * B + (8 - fx) * fy * C + fx * fy * D + 32) >> 6;
}

void orc_pred_luma_block(const uint8_t *ref, int py, int wc, int hc, int x0, int y0,
                         int w, int h, int mvx, int mvy, uint8_t *dst, int dpitch)
{
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++)
            dst[y * dpitch + x] = (uint8_t)orc_luma_qpel(ref, py, wc, hc,
                                                         (x0 + x) * 4 + mvx, (y0 + y) * 4 + mvy);
}

/* ----------------------------------------------------- motion estimation */

/* Exhaustive integer search of one MB over [-R,R]^2 against the
 * edge-replicated reference.  Costs:
 *   cost8_q = SAD_q + lambda8*(|dx|+|dy|)        for the four 8x8 quadrants
 *   cost16  = SAD16 + 4*lambda8*(|dx|+|dy|)
 * and the winner is the smallest key cost*2^15 + ((dy+R)*(2R+1) + dx+R),
 * i.e. lowest cost, ties to the first candidate in raster order.  The
 * reference's integer search is a +-16 diamond with lambda*|mv|
 * (shaders/motion_estimation.comp:340-374); full search is the form asked
 * of the MI355X build. */
void orc_me_fullsearch_mb(const uint8_t *cur, const uint8_t *ref, int py, int wc, int hc,
                          int range, int lambda8, int mbx, int mby,
                          orc_imv_t *int16, orc_imv_t *int8)
{
    const int n = 2 * range + 1;
    const int x0 = mbx * 16, y0 = mby * 16;
    uint8_t c[16][16];
    for (int y = 0; y < 16; y++)
        for (int x = 0; x < 16; x++) c[y][x] = cur[(size_t)(y0 + y) * py + x0 + x];
    /* window of the reference with replicated edges */
    const int ww = 16 + 2 * range;
    uint8_t *win = (uint8_t *)malloc((size_t)ww * ww);
    for (int y = 0; y < ww; y++)
        for (int x = 0; x < ww; x++)
            win[(size_t)y * ww + x] = (uint8_t)G(ref, py, wc, hc, x0 - range + x, y0 - range + y);
    uint64_t best16 = UINT64_MAX, best8[4] = {UINT64_MAX, UINT64_MAX, UINT64_MAX, UINT64_MAX};
    uint32_t sad16b = 0, sad8b[4] = {0, 0, 0, 0};
    for (int dy = -range; dy <= range; dy++) {
        for (int dx = -range; dx <= range; dx++) {
            uint32_t q[4] = {0, 0, 0, 0};
            for (int y = 0; y < 16; y++) {
                const uint8_t *r = win + (size_t)(y + dy + range) * ww + dx + range;
                for (int x = 0; x < 16; x++)
                    q[((y >> 3) << 1) | (x >> 3)] += (uint32_t)iabs(c[y][x] - r[x]);
            }
            uint32_t l1 = (uint32_t)(iabs(dx) + iabs(dy));
            uint64_t idx = (uint64_t)((dy + range) * n + (dx + range));
            uint32_t s16 = q[0] + q[1] + q[2] + q[3];
            uint64_t k16 = ((uint64_t)(s16 + 4u * lambda8 * l1) << 15) | idx;
            if (k16 < best16) { best16 = k16; sad16b = s16; }
            for (int b = 0; b < 4; b++) {
                uint64_t k = ((uint64_t)(q[b] + lambda8 * l1) << 15) | idx;
                if (k < best8[b]) { best8[b] = k; sad8b[b] = q[b]; }
            }
        }
    }
    free(win);
    int i16 = (int)(best16 & 0x7fff);
    int16->dx = i16 % n - range;
    int16->dy = i16 / n - range;
    int16->cost = (uint32_t)(best16 >> 15);
    int16->sad = sad16b;
    for (int b = 0; b < 4; b++) {
        int i = (int)(best8[b] & 0x7fff);
        int8[b].dx = i % n - range;
        int8[b].dy = i / n - range;
        int8[b].cost = (uint32_t)(best8[b] >> 15);
        int8[b].sad = sad8b[b];
    }
}

void orc_me_fullsearch(const uint8_t *cur, const uint8_t *ref, int py, int wc, int hc,
                       int range, int lambda8, orc_imv_t *int16, orc_imv_t *int8)
{
    const int mbw = wc / 16, mbh = hc / 16;
    for (int mby = 0; mby < mbh; mby++)
        for (int mbx = 0; mbx < mbw; mbx++) {
            int m = mby * mbw + mbx;
            orc_me_fullsearch_mb(cur, ref, py, wc, hc, range, lambda8, mbx, mby,
                                 &int16[m], &int8[4 * m]);
        }
}

static uint32_t sad_qpel(const uint8_t *cur, const uint8_t *ref, int py, int wc, int hc,
                         int x0, int y0, int bw, int bh, int mvx, int mvy)
{
    uint32_t s = 0;
    for (int y = 0; y < bh; y++)
        for (int x = 0; x < bw; x++)
            s += (uint32_t)iabs(cur[(size_t)(y0 + y) * py + x0 + x] -
                                orc_luma_qpel(ref, py, wc, hc, (x0 + x) * 4 + mvx, (y0 + y) * 4 + mvy));
    return s;
}

/* Half-pel then quarter-pel refinement around an integer winner: the 8
 * candidates of each ring in the reference's order, plain SAD, strict <
 * so the earlier candidate keeps a tie
 * (shaders/motion_estimation.comp:395-459).  The starting value is the
 * integer winner's own SAD. */
static void refine_block(const uint8_t *cur, const uint8_t *ref, int py, int wc, int hc,
                         int x0, int y0, int bw, int bh, int dx, int dy, orc_mv_t *out)
{
    static const int HO[8][2] = {{2, 0}, {-2, 0}, {0, 2}, {0, -2}, {2, 2}, {2, -2}, {-2, 2}, {-2, -2}};
    static const int QO[8][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}, {1, 1}, {1, -1}, {-1, 1}, {-1, -1}};
    int cx = dx * 4, cy = dy * 4;
    uint32_t best = sad_qpel(cur, ref, py, wc, hc, x0, y0, bw, bh, cx, cy);
    int bx = 0, by = 0;
    for (int c = 0; c < 8; c++) {
        uint32_t s = sad_qpel(cur, ref, py, wc, hc, x0, y0, bw, bh, cx + HO[c][0], cy + HO[c][1]);
        if (s < best) { best = s; bx = HO[c][0]; by = HO[c][1]; }
    }
    cx += bx; cy += by;
    bx = by = 0;
    for (int c = 0; c < 8; c++) {
        uint32_t s = sad_qpel(cur, ref, py, wc, hc, x0, y0, bw, bh, cx + QO[c][0], cy + QO[c][1]);
        if (s < best) { best = s; bx = QO[c][0]; by = QO[c][1]; }
    }
    out->mvx = cx + bx;
    out->mvy = cy + by;
    out->sad = best;
    out->pad = 0;
}

void orc_me_refine(const uint8_t *cur, const uint8_t *ref, int py, int wc, int hc,
                   const orc_imv_t *int16, const orc_imv_t *int8, orc_mv_t *mv16, orc_mv_t *mv8)
{
    const int mbw = wc / 16, mbh = hc / 16;
    for (int mby = 0; mby < mbh; mby++)
        for (int mbx = 0; mbx < mbw; mbx++) {
            int m = mby * mbw + mbx;
            refine_block(cur, ref, py, wc, hc, mbx * 16, mby * 16, 16, 16,
                         int16[m].dx, int16[m].dy, &mv16[m]);
            if (mv8)
                for (int b = 0; b < 4; b++)
                    refine_block(cur, ref, py, wc, hc, mbx * 16 + (b & 1) * 8, mby * 16 + (b >> 1) * 8,
                                 8, 8, int8[4 * m + b].dx, int8[4 * m + b].dy, &mv8[4 * m + b]);
        }
}

/* The same refinement for one MB (its 16x16 result and its four 8x8 ones),
 * so a checker can refine a sample of MBs of a large frame. */
void orc_me_refine_mb(const uint8_t *cur, const uint8_t *ref, int py, int wc, int hc, int mbx, int mby,
                      const orc_imv_t *int16, const orc_imv_t *int8, orc_mv_t *mv16, orc_mv_t *mv8)
{
    refine_block(cur, ref, py, wc, hc, mbx * 16, mby * 16, 16, 16, int16->dx, int16->dy, mv16);
    for (int b = 0; b < 4; b++)
        refine_block(cur, ref, py, wc, hc, mbx * 16 + (b & 1) * 8, mby * 16 + (b >> 1) * 8, 8, 8, int8[b].dx,
                     int8[b].dy, &mv8[b]);
}

/* ------------------------------------------------------- P16x16 encode */

/* One P macroblock: MC prediction (residual_predict.comp:417-434 luma,
 * :606-631 chroma), forward transform (dct_transform.comp:34-98), inter AC
 * quantisation of all 16 positions (quantize.comp:105-126), and the
 * reconstruction of reconstruct.comp:346-440: luma position 0 dequantised
 * as AC, chroma DC through Hadamard -> quant -> Hadamard -> dequant. */
static void encode_p_mb(const uint8_t *cy, const uint8_t *cuv, const uint8_t *ry,
                        const uint8_t *ruv, int py, int puv, int wc, int hc, int qp,
                        int mbx, int mby, orc_mv_t mv, int16_t *lv, int32_t *dcb,
                        uint32_t *nzb, uint8_t *oy, uint8_t *ouv)
{
    const int qpc = orc_chroma_qp(qp);
    int32_t pred[24][16], res[24][16], co[24][16];
    for (int y = 0; y < 16; y++)
        for (int x = 0; x < 16; x++) {
            int X = mbx * 16 + x, Y = mby * 16 + y;
            int p = orc_luma_qpel(ry, py, wc, hc, X * 4 + mv.mvx, Y * 4 + mv.mvy);
            int b = (y >> 2) * 4 + (x >> 2), k = (y & 3) * 4 + (x & 3);
            pred[b][k] = p;
            res[b][k] = cy[(size_t)Y * py + X] - p;
        }
    for (int comp = 0; comp < 2; comp++)
        for (int v = 0; v < 8; v++)
            for (int u = 0; u < 8; u++) {
                int X = mbx * 8 + u, Y = mby * 8 + v;
                int p = orc_chroma_8th(ruv, puv, wc / 2, hc / 2, comp, X * 8 + mv.mvx, Y * 8 + mv.mvy);
                int b = 16 + comp * 4 + (v >> 2) * 2 + (u >> 2), k = (v & 3) * 4 + (u & 3);
                pred[b][k] = p;
                res[b][k] = cuv[(size_t)Y * puv + 2 * X + comp] - p;
            }
    for (int b = 0; b < 24; b++) {
        orc_fwd4x4(res[b], co[b]);
        dcb[b] = co[b][0];
        int q = b < 16 ? qp : qpc;
        uint32_t m = 0;
        for (int k = 0; k < 16; k++) {
            int l = orc_quant_ac(co[b][k], k, q, 0);
            lv[b * 16 + k] = (int16_t)l;
            if (l) m |= 1u << k;
        }
        nzb[b] = m;
    }
    /* reconstruction */
    for (int b = 0; b < 16; b++) {
        int32_t d[16], r[16];
        for (int k = 0; k < 16; k++) d[k] = orc_dequant_ac(lv[b * 16 + k], k, qp);
        orc_inv4x4(d, r);
        int bx = b & 3, by = b >> 2;
        for (int k = 0; k < 16; k++) {
            int X = mbx * 16 + bx * 4 + (k & 3), Y = mby * 16 + by * 4 + (k >> 2);
            oy[(size_t)Y * py + X] = (uint8_t)clip255(pred[b][k] + r[k]);
        }
    }
    for (int comp = 0; comp < 2; comp++) {
        int32_t din[4], raw[4], lvl[4], f[4];
        for (int s = 0; s < 4; s++) din[s] = co[16 + comp * 4 + s][0];
        orc_chroma_dc_hadamard(din, raw);
        for (int s = 0; s < 4; s++) lvl[s] = orc_quant_dc_chroma(raw[s], qpc);
        orc_chroma_dc_hadamard(lvl, f);
        for (int s = 0; s < 4; s++) {
            int b = 16 + comp * 4 + s;
            int32_t d[16], r[16];
            d[0] = orc_dequant_dc_chroma(f[s], qpc);
            for (int k = 1; k < 16; k++) d[k] = orc_dequant_ac(lv[b * 16 + k], k, qpc);
            orc_inv4x4(d, r);
            for (int k = 0; k < 16; k++) {
                int X = mbx * 8 + (s & 1) * 4 + (k & 3), Y = mby * 8 + (s >> 1) * 4 + (k >> 2);
                ouv[(size_t)Y * puv + 2 * X + comp] = (uint8_t)clip255(pred[b][k] + r[k]);
            }
        }
    }
}

void orc_h264_encode_p(const uint8_t *cy, const uint8_t *cuv, const uint8_t *ry,
                       const uint8_t *ruv, int py, int puv, int wc, int hc, int qp,
                       const orc_mv_t *mvs, int16_t *levels, int32_t *dc, uint32_t *nz,
                       uint32_t *modes, uint8_t *oy, uint8_t *ouv)
{
    const int mbw = wc / 16, mbh = hc / 16;
    for (int mby = 0; mby < mbh; mby++)
        for (int mbx = 0; mbx < mbw; mbx++) {
            int m = mby * mbw + mbx;
            encode_p_mb(cy, cuv, ry, ruv, py, puv, wc, hc, qp, mbx, mby, mvs[m],
                        levels + (size_t)m * 384, dc + m * 24, nz + m * 24, oy, ouv);
            modes[m] = 2; /* DC: residual_predict.comp:422 */
        }
}

/* ------------------------------------------------------- I16x16 encode */

/* Slice of MB `mb` for the partition start(s) = floor(s*total/num_slices)
 * (shaders/intra_wavefront.comp:196-204). */
int orc_slice_of(int mb, int total, int ns)
{
    if (ns <= 1 || total <= 0) return 0;
    int s = (int)(((int64_t)mb * ns) / total);
    if (s >= ns) s = ns - 1;
    while (s + 1 < ns && ((int64_t)(s + 1) * total) / ns <= mb) s++;
    while (s > 0 && ((int64_t)s * total) / ns > mb) s--;
    return s;
}

/* Intra 16x16 prediction of H.264 8.3.3 (modes 0 V, 1 H, 2 DC, 3 Plane)
 * from neighbour samples; ta / la say whether the top / left neighbours
 * are available (shaders/intra_wavefront.comp:440-475). */
void orc_pred16x16(int mode, const int top[16], const int left[16], int corner, int ta, int la,
                   uint8_t out[256])
{
    int st = 0, sl = 0;
    for (int i = 0; i < 16; i++) { st += top[i]; sl += left[i]; }
    const int dcv = (ta && la) ? (st + sl + 16) >> 5 : ta ? (st + 8) >> 4 : la ? (sl + 8) >> 4 : 128;
    int H = 0, V = 0;
    for (int k = 0; k < 8; k++) {
        H += (k + 1) * (top[8 + k] - (6 - k >= 0 ? top[6 - k] : corner));
        V += (k + 1) * (left[8 + k] - (6 - k >= 0 ? left[6 - k] : corner));
    }
    const int a = 16 * (left[15] + top[15]), b = (5 * H + 32) >> 6, c = (5 * V + 32) >> 6;
    for (int y = 0; y < 16; y++)
        for (int x = 0; x < 16; x++)
            out[y * 16 + x] = (uint8_t)(mode == 0 ? top[x] : mode == 1 ? left[y] : mode == 3
                                        ? clip255((a + b * (x - 7) + c * (y - 7) + 16) >> 5) : dcv);
}

/* Chroma DC prediction of H.264 8.3.4.1-3 for a 4:2:0 8x8 block: each 4x4
 * quadrant takes its own DC from the neighbours the standard assigns it
 * (shaders/intra_wavefront.comp:610-633). */
void orc_pred_chroma_dc(const int t[8], const int l[8], int ta, int la, uint8_t out[64])
{
    const int t0 = t[0] + t[1] + t[2] + t[3], t1 = t[4] + t[5] + t[6] + t[7];
    const int l0 = l[0] + l[1] + l[2] + l[3], l1 = l[4] + l[5] + l[6] + l[7];
    int d[4];
    d[0] = (ta && la) ? (t0 + l0 + 4) >> 3 : ta ? (t0 + 2) >> 2 : la ? (l0 + 2) >> 2 : 128;
    d[1] = ta ? (t1 + 2) >> 2 : la ? (l0 + 2) >> 2 : 128;
    d[2] = la ? (l1 + 2) >> 2 : ta ? (t0 + 2) >> 2 : 128;
    d[3] = (ta && la) ? (t1 + l1 + 4) >> 3 : ta ? (t1 + 2) >> 2 : la ? (l1 + 2) >> 2 : 128;
    for (int y = 0; y < 8; y++)
        for (int x = 0; x < 8; x++) out[y * 8 + x] = (uint8_t)d[(y >> 2) * 2 + (x >> 2)];
}

/* One I16x16 macroblock, predicted from already reconstructed neighbours
 * (shaders/intra_wavefront.comp:398-756): SAD mode choice over DC, V, H,
 * Plane in that order with strict <, intra AC quantisation of positions
 * 1..15 (position 0 written as 0, its nz bit clear), luma DC through the
 * Hadamard path, chroma DC prediction of H.264 8.3.4.1-3 with the same
 * chain. */
static void encode_i_mb(const uint8_t *cy, const uint8_t *cuv, int py, int puv, int wc,
                        int hc, int qp, int ns, int mbx, int mby, int16_t *lv,
                        int32_t *dcb, uint32_t *nzb, uint32_t *mode_out, uint8_t *oy,
                        uint8_t *ouv)
{
    const int mbw = wc / 16, mbh = hc / 16, total = mbw * mbh;
    const int m = mby * mbw + mbx, sl = orc_slice_of(m, total, ns);
    const int qpc = orc_chroma_qp(qp);
    const int te = mby > 0, le = mbx > 0;
    const int ta = te && orc_slice_of(m - mbw, total, ns) == sl;
    const int la = le && orc_slice_of(m - 1, total, ns) == sl;
    const int ca = te && le && orc_slice_of(m - mbw - 1, total, ns) == sl;
    const int X0 = mbx * 16, Y0 = mby * 16;
    int top[16], left[16], corner = 128;
    for (int i = 0; i < 16; i++) {
        top[i] = te ? oy[(size_t)(Y0 - 1) * py + X0 + i] : 128;
        left[i] = le ? oy[(size_t)(Y0 + i) * py + X0 - 1] : 128;
    }
    if (te && le) corner = oy[(size_t)(Y0 - 1) * py + X0 - 1];
    const int plane_ok = ta && la && ca;
    uint8_t pr[4][256];
    uint32_t sad[4] = {0, 0, 0, 0};
    for (int md = 0; md < 4; md++) {
        orc_pred16x16(md, top, left, corner, ta, la, pr[md]);
        for (int y = 0; y < 16; y++)
            for (int x = 0; x < 16; x++)
                sad[md] += (uint32_t)iabs(cy[(size_t)(Y0 + y) * py + X0 + x] - pr[md][y * 16 + x]);
    }
    /* DC first, then V, H, Plane, strict < */
    uint32_t best = sad[2], mode = 2;
    if (ta && sad[0] < best) { best = sad[0]; mode = 0; }
    if (la && sad[1] < best) { best = sad[1]; mode = 1; }
    if (plane_ok && sad[3] < best) { best = sad[3]; mode = 3; }
    *mode_out = mode;

    int32_t pred[16][16], co[16][16], acl[16][16];
    int32_t dcin[16];
    for (int y = 0; y < 16; y++)
        for (int x = 0; x < 16; x++)
            pred[(y >> 2) * 4 + (x >> 2)][(y & 3) * 4 + (x & 3)] = pr[mode][y * 16 + x];
    for (int b = 0; b < 16; b++) {
        int32_t r[16];
        int bx = b & 3, by = b >> 2;
        for (int k = 0; k < 16; k++)
            r[k] = cy[(size_t)(Y0 + by * 4 + (k >> 2)) * py + X0 + bx * 4 + (k & 3)] - pred[b][k];
        orc_fwd4x4(r, co[b]);
        dcb[b] = co[b][0];
        dcin[b] = co[b][0];
        lv[b * 16] = 0;
        acl[b][0] = 0;
        uint32_t msk = 0;
        for (int k = 1; k < 16; k++) {
            int l = orc_quant_ac(co[b][k], k, qp, 1);
            acl[b][k] = l;
            lv[b * 16 + k] = (int16_t)l;
            if (l) msk |= 1u << k;
        }
        nzb[b] = msk;
    }
    int32_t raw[16], dl[16], f[16];
    orc_luma_dc_hadamard_fwd(dcin, raw);
    for (int b = 0; b < 16; b++) dl[b] = orc_quant_dc_luma(raw[b], qp);
    orc_luma_dc_hadamard_inv(dl, f);
    for (int b = 0; b < 16; b++) {
        int32_t d[16], r[16];
        d[0] = orc_dequant_dc_luma(f[b], qp);
        for (int k = 1; k < 16; k++) d[k] = orc_dequant_ac(acl[b][k], k, qp);
        orc_inv4x4(d, r);
        int bx = b & 3, by = b >> 2;
        for (int k = 0; k < 16; k++)
            oy[(size_t)(Y0 + by * 4 + (k >> 2)) * py + X0 + bx * 4 + (k & 3)] =
                (uint8_t)clip255(pred[b][k] + r[k]);
    }

    /* chroma: DC prediction per 4x4 quadrant, H.264 8.3.4.1-3
     * (shaders/intra_wavefront.comp:600-633) */
    const int U0 = mbx * 8, V0 = mby * 8;
    for (int comp = 0; comp < 2; comp++) {
        int t[8], l[8];
        for (int i = 0; i < 8; i++) {
            t[i] = ta ? ouv[(size_t)(V0 - 1) * puv + 2 * (U0 + i) + comp] : 128;
            l[i] = la ? ouv[(size_t)(V0 + i) * puv + 2 * (U0 - 1) + comp] : 128;
        }
        uint8_t cp[64];
        orc_pred_chroma_dc(t, l, ta, la, cp);
        int dcq[4];
        for (int q = 0; q < 4; q++) dcq[q] = cp[(q >> 1) * 32 + (q & 1) * 4];
        int32_t cco[4][16], cac[4][16], cdin[4];
        for (int s = 0; s < 4; s++) {
            int b = 16 + comp * 4 + s;
            int32_t r[16];
            for (int k = 0; k < 16; k++) {
                int X = U0 + (s & 1) * 4 + (k & 3), Y = V0 + (s >> 1) * 4 + (k >> 2);
                r[k] = cuv[(size_t)Y * puv + 2 * X + comp] - dcq[s];
            }
            orc_fwd4x4(r, cco[s]);
            dcb[b] = cco[s][0];
            cdin[s] = cco[s][0];
            lv[b * 16] = 0;
            cac[s][0] = 0;
            uint32_t msk = 0;
            for (int k = 1; k < 16; k++) {
                int lq = orc_quant_ac(cco[s][k], k, qpc, 1);
                cac[s][k] = lq;
                lv[b * 16 + k] = (int16_t)lq;
                if (lq) msk |= 1u << k;
            }
            nzb[b] = msk;
        }
        int32_t craw[4], clv[4], cf[4];
        orc_chroma_dc_hadamard(cdin, craw);
        for (int s = 0; s < 4; s++) clv[s] = orc_quant_dc_chroma(craw[s], qpc);
        orc_chroma_dc_hadamard(clv, cf);
        for (int s = 0; s < 4; s++) {
            int32_t d[16], r[16];
            d[0] = orc_dequant_dc_chroma(cf[s], qpc);
            for (int k = 1; k < 16; k++) d[k] = orc_dequant_ac(cac[s][k], k, qpc);
            orc_inv4x4(d, r);
            for (int k = 0; k < 16; k++) {
                int X = U0 + (s & 1) * 4 + (k & 3), Y = V0 + (s >> 1) * 4 + (k >> 2);
                ouv[(size_t)Y * puv + 2 * X + comp] = (uint8_t)clip255(dcq[s] + r[k]);
            }
        }
    }
}

/* Raster order satisfies the left / top / top-left dependencies of every
 * MB, so it gives the same result as the anti-diagonal wavefront of
 * src/gpu_compute.c:3217-3251. */
void orc_h264_encode_i(const uint8_t *cy, const uint8_t *cuv, int py, int puv, int wc, int hc,
                       int qp, int num_slices, int16_t *levels, int32_t *dc, uint32_t *nz,
                       uint32_t *modes, uint8_t *oy, uint8_t *ouv)
{
    const int mbw = wc / 16, mbh = hc / 16;
    for (int mby = 0; mby < mbh; mby++)
        for (int mbx = 0; mbx < mbw; mbx++) {
            int m = mby * mbw + mbx;
            encode_i_mb(cy, cuv, py, puv, wc, hc, qp, num_slices, mbx, mby,
                        levels + (size_t)m * 384, dc + m * 24, nz + m * 24, &modes[m], oy, ouv);
        }
}

/* ------------------------------------------------------------ deblocking */

/* H.264 Tables 8-16 / 8-17 (alpha', beta', tC0' by indexA / indexB). */
static const uint8_t ALPHA[52] = {
    0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 4, 4, 5, 6, 7, 8, 9, 10, 12, 13,
    15, 17, 20, 22, 25, 28, 32, 36, 40, 45, 50, 56, 63, 71, 80, 90, 101, 113, 127, 144, 162, 182,
    203, 226, 255, 255};
static const uint8_t BETA[52] = {
    0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4,
    6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13, 14, 14, 15, 15, 16, 16, 17, 17, 18, 18};
static const uint8_t TC0[52][3] = {
    {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0},
    {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0},
    {0, 0, 0}, {0, 0, 1}, {0, 0, 1}, {0, 0, 1}, {0, 0, 1}, {0, 1, 1}, {0, 1, 1}, {1, 1, 1},
    {1, 1, 1}, {1, 1, 1}, {1, 1, 1}, {1, 1, 2}, {1, 1, 2}, {1, 1, 2}, {1, 1, 2}, {1, 2, 3},
    {1, 2, 3}, {2, 2, 3}, {2, 2, 4}, {2, 3, 4}, {2, 3, 4}, {3, 3, 5}, {3, 4, 6}, {3, 4, 6},
    {4, 5, 7}, {4, 5, 8}, {4, 6, 9}, {5, 7, 10}, {6, 8, 11}, {6, 8, 13}, {7, 10, 14}, {8, 11, 16},
    {9, 12, 18}, {10, 13, 20}, {11, 15, 23}, {13, 17, 25}};

int orc_h264_alpha(int i) { return ALPHA[clampi(i, 0, 51)]; }
int orc_h264_beta(int i) { return BETA[clampi(i, 0, 51)]; }
int orc_h264_tc0(int i, int bs) { return TC0[clampi(i, 0, 51)][bs - 1]; }

/* One line of samples across a luma edge (H.264 8.7.2.3 for bS<4, 8.7.2.4
 * for bS=4), FilterOffsetA = FilterOffsetB = 0 so indexA = indexB = qPav.
 * q0 points at the first q sample, step is +1 across a vertical edge and
 * +pitch across a horizontal one. */
void orc_h264_filter_line_luma(uint8_t *q0p, int d, int bs, int qp)
{
    const int alpha = ALPHA[qp], beta = BETA[qp];
    int p0 = q0p[-d], p1 = q0p[-2 * d], p2 = q0p[-3 * d], p3 = q0p[-4 * d];
    int q0 = q0p[0], q1 = q0p[d], q2 = q0p[2 * d], q3 = q0p[3 * d];
    if (!(iabs(p0 - q0) < alpha && iabs(p1 - p0) < beta && iabs(q1 - q0) < beta)) return;
    int ap = iabs(p2 - p0), aq = iabs(q2 - q0);
    if (bs < 4) {
        int tc0 = TC0[qp][bs - 1];
        int tc = tc0 + (ap < beta) + (aq < beta);
        int delta = clampi((((q0 - p0) << 2) + (p1 - q1) + 4) >> 3, -tc, tc);
        q0p[-d] = (uint8_t)clip255(p0 + delta);
        q0p[0] = (uint8_t)clip255(q0 - delta);
        if (ap < beta) q0p[-2 * d] = (uint8_t)(p1 + clampi((p2 + ((p0 + q0 + 1) >> 1) - (p1 << 1)) >> 1, -tc0, tc0));
        if (aq < beta) q0p[d] = (uint8_t)(q1 + clampi((q2 + ((p0 + q0 + 1) >> 1) - (q1 << 1)) >> 1, -tc0, tc0));
    } else {
        int narrow = iabs(p0 - q0) < ((alpha >> 2) + 2);
        if (ap < beta && narrow) {
            q0p[-d] = (uint8_t)((p2 + 2 * p1 + 2 * p0 + 2 * q0 + q1 + 4) >> 3);
            q0p[-2 * d] = (uint8_t)((p2 + p1 + p0 + q0 + 2) >> 2);
            q0p[-3 * d] = (uint8_t)((2 * p3 + 3 * p2 + p1 + p0 + q0 + 4) >> 3);
        } else {
            q0p[-d] = (uint8_t)((2 * p1 + p0 + q1 + 2) >> 2);
        }
        if (aq < beta && narrow) {
            q0p[0] = (uint8_t)((p1 + 2 * p0 + 2 * q0 + 2 * q1 + q2 + 4) >> 3);
            q0p[d] = (uint8_t)((p0 + q0 + q1 + q2 + 2) >> 2);
            q0p[2 * d] = (uint8_t)((2 * q3 + 3 * q2 + q1 + q0 + p0 + 4) >> 3);
        } else {
            q0p[0] = (uint8_t)((2 * q1 + q0 + p1 + 2) >> 2);
        }
    }
}

/* Chroma line (chromaEdgeFlag = 1): only p0/q0 change, tc = tc0 + 1. */
void orc_h264_filter_line_chroma(uint8_t *q0p, int d, int bs, int qpc)
{
    const int alpha = ALPHA[qpc], beta = BETA[qpc];
    int p0 = q0p[-d], p1 = q0p[-2 * d], q0 = q0p[0], q1 = q0p[d];
    if (!(iabs(p0 - q0) < alpha && iabs(p1 - p0) < beta && iabs(q1 - q0) < beta)) return;
    if (bs < 4) {
        int tc = TC0[qpc][bs - 1] + 1;
        int delta = clampi((((q0 - p0) << 2) + (p1 - q1) + 4) >> 3, -tc, tc);
        q0p[-d] = (uint8_t)clip255(p0 + delta);
        q0p[0] = (uint8_t)clip255(q0 - delta);
    } else {
        q0p[-d] = (uint8_t)((2 * p1 + p0 + q1 + 2) >> 2);
        q0p[0] = (uint8_t)((2 * q1 + q0 + p1 + 2) >> 2);
    }
}

/* Boundary strength of H.264 8.7.2.1 for the encoder's pictures: every MB
 * is I16x16 (intra pictures) or P16x16 with reference index 0 (inter
 * pictures), so the mixed-reference cases reduce to the MV test. */
static int bs_of(int intra, int mb_edge, uint32_t nzp, uint32_t nzq, orc_mv_t mp, orc_mv_t mq)
{
    if (intra) return mb_edge ? 4 : 3;
    if (nzp || nzq) return 2;
    if (iabs(mp.mvx - mq.mvx) >= 4 || iabs(mp.mvy - mq.mvy) >= 4) return 1;
    return 0;
}

/* Macroblocks in raster order; in each, luma vertical edges left to right,
 * then horizontal top to bottom, chroma edges 0 and 2 with the strength of
 * the luma edge they sit on (H.264 8.7; the order of
 * src/h264_deblock.c:263-355). */
void orc_h264_deblock(uint8_t *y, uint8_t *uv, int py, int puv, int wc, int hc, int qp,
                      int is_intra, const uint32_t *nz, const orc_mv_t *mvs, int ns, int idc)
{
    if (idc == 1) return;
    const int mbw = wc / 16, mbh = hc / 16, total = mbw * mbh;
    const int qpc = orc_chroma_qp(qp);
    const orc_mv_t zero = {0, 0, 0, 0};
    for (int my = 0; my < mbh; my++)
        for (int mx = 0; mx < mbw; mx++) {
            const int m = my * mbw + mx;
            const int sl = orc_slice_of(m, total, ns);
            const int skip_left = mx == 0 || (idc == 2 && orc_slice_of(m - 1, total, ns) != sl);
            const int skip_top = my == 0 || (idc == 2 && orc_slice_of(m - mbw, total, ns) != sl);
            uint8_t *Y = y + (size_t)my * 16 * py + mx * 16;
            uint8_t *C = uv + (size_t)my * 8 * puv + mx * 16;
            orc_mv_t mq = is_intra ? zero : mvs[m];
            for (int dir = 0; dir < 2; dir++) {
                for (int e = 0; e < 4; e++) {
                    if (e == 0 && (dir == 0 ? skip_left : skip_top)) continue;
                    const int pm = e ? m : (dir == 0 ? m - 1 : m - mbw);
                    orc_mv_t mp = is_intra ? zero : mvs[pm];
                    for (int i = 0; i < 4; i++) {
                        /* block q at (e, i) for vertical, (i, e) for horizontal */
                        int bq = dir == 0 ? i * 4 + e : e * 4 + i;
                        int bp = dir == 0 ? (e ? i * 4 + e - 1 : i * 4 + 3) : (e ? (e - 1) * 4 + i : 12 + i);
                        int bs = bs_of(is_intra, e == 0, nz[pm * 24 + bp], nz[m * 24 + bq], mp, mq);
                        if (!bs) continue;
                        for (int k = 0; k < 4; k++) {
                            if (dir == 0) orc_h264_filter_line_luma(Y + (size_t)(i * 4 + k) * py + e * 4, 1, bs, qp);
                            else orc_h264_filter_line_luma(Y + (size_t)(e * 4) * py + i * 4 + k, py, bs, qp);
                        }
                        if (e == 0 || e == 2) {
                            for (int comp = 0; comp < 2; comp++)
                                for (int k = 0; k < 2; k++) {
                                    if (dir == 0)
                                        orc_h264_filter_line_chroma(C + (size_t)(i * 2 + k) * puv + 2 * (e * 2) + comp, 2, bs, qpc);
                                    else
                                        orc_h264_filter_line_chroma(C + (size_t)(e * 2) * puv + 2 * (i * 2 + k) + comp, puv, bs, qpc);
                                }
                        }
                    }
                }
            }
        }
}

/* ------------------------------------------------------------ whole frame */

void orc_h264_encode_frame(const uint8_t *cy, const uint8_t *cuv, const uint8_t *ry,
                           const uint8_t *ruv, int py, int puv, int wc, int hc, int qp,
                           int is_intra, int ns, int range, int lambda8, int idc,
                           orc_mv_t *mvs, orc_mv_t *mv8, int16_t *levels, int32_t *dc,
                           uint32_t *nz, uint32_t *modes, uint8_t *oy, uint8_t *ouv)
{
    const int nmb = (wc / 16) * (hc / 16);
    if (is_intra) {
        memset(mvs, 0, sizeof(orc_mv_t) * nmb);
        if (mv8) memset(mv8, 0, sizeof(orc_mv_t) * nmb * 4);
        orc_h264_encode_i(cy, cuv, py, puv, wc, hc, qp, ns, levels, dc, nz, modes, oy, ouv);
    } else {
        orc_imv_t *i16 = (orc_imv_t *)malloc(sizeof(orc_imv_t) * nmb);
        orc_imv_t *i8 = (orc_imv_t *)malloc(sizeof(orc_imv_t) * nmb * 4);
        orc_me_fullsearch(cy, ry, py, wc, hc, range, lambda8, i16, i8);
        orc_me_refine(cy, ry, py, wc, hc, i16, i8, mvs, mv8);
        free(i16);
        free(i8);
        orc_h264_encode_p(cy, cuv, ry, ruv, py, puv, wc, hc, qp, mvs, levels, dc, nz, modes, oy, ouv);
    }
    orc_h264_deblock(oy, ouv, py, puv, wc, hc, qp, is_intra, nz, mvs, ns, idc);
}

/* ------------------------------------------------------------------------
 * Intra 4x4 and 8x8 luma prediction (H.264 8.3.1.2 and 8.3.2.2), restated
 * from the standard's equations; the reference decoder's own forms are
 * enc/src/h264_pred.c:45-150 and :321-470.  Neighbours in natural order:
 * left[y] = p[-1,y], corner = p[-1,-1], top[x] = p[x,-1] (x < 8 for 4x4
 * including the top-right run, x < 16 for 8x8).  Modes 0..8 = V, H, DC,
 * diagonal down-left, diagonal down-right, vertical-right, horizontal-down,
 * vertical-left, horizontal-up.
 * ---------------------------------------------------------------------- */
static int f3(int a, int b, int c) { return (a + 2 * b + c + 2) >> 2; }
static int f2(int a, int b) { return (a + b + 1) >> 1; }

/* One sample of the nine directional / DC modes for an n x n block
 * (n = 4 or 8) from (already filtered, for 8x8) references; T(i) = p[i,-1]
 * for i >= -1 (T(-1) = corner), L(i) = p[-1,i] for i >= -1. */
static int dir_sample(int n, int mode, int x, int y, const int *top, const int *left, int corner, int dc)
{
#define T(i) ((i) < 0 ? corner : top[i])
#define L(i) ((i) < 0 ? corner : left[i])
    const int last = 2 * n - 1;   /* index of the last top sample */
    switch (mode) {
    case 0: return top[x];
    case 1: return left[y];
    case 2: return dc;
    case 3:
        if (x == n - 1 && y == n - 1) return (top[last - 1] + 3 * top[last] + 2) >> 2;
        return f3(top[x + y], top[x + y + 1], top[x + y + 2]);
    case 4:
        if (x > y) return f3(T(x - y - 2), T(x - y - 1), top[x - y]);
        if (x < y) return f3(L(y - x - 2), L(y - x - 1), left[y - x]);
        return f3(top[0], corner, left[0]);
    case 5: {
        const int z = 2 * x - y;
        if (z >= 0 && !(z & 1)) return f2(T(x - (y >> 1) - 1), top[x - (y >> 1)]);
        if (z > 0) return f3(T(x - (y >> 1) - 2), T(x - (y >> 1) - 1), top[x - (y >> 1)]);
        if (z == -1) return f3(left[0], corner, top[0]);
        return f3(L(y - 2 * x - 1), L(y - 2 * x - 2), L(y - 2 * x - 3));
    }
    case 6: {
        const int z = 2 * y - x;
        if (z >= 0 && !(z & 1)) return f2(L(y - (x >> 1) - 1), left[y - (x >> 1)]);
        if (z > 0) return f3(L(y - (x >> 1) - 2), L(y - (x >> 1) - 1), left[y - (x >> 1)]);
        if (z == -1) return f3(left[0], corner, top[0]);
        return f3(T(x - 2 * y - 1), T(x - 2 * y - 2), T(x - 2 * y - 3));
    }
    case 7:
        if (!(y & 1)) return f2(top[x + (y >> 1)], top[x + (y >> 1) + 1]);
        return f3(top[x + (y >> 1)], top[x + (y >> 1) + 1], top[x + (y >> 1) + 2]);
    case 8: {
        const int z = x + 2 * y, lim = 2 * n - 3;
        if (z < lim && !(z & 1)) return f2(left[y + (x >> 1)], left[y + (x >> 1) + 1]);
        if (z < lim) return f3(left[y + (x >> 1)], left[y + (x >> 1) + 1], left[y + (x >> 1) + 2]);
        if (z == lim) return (left[n - 2] + 3 * left[n - 1] + 2) >> 2;
        return left[n - 1];
    }
    default: return 128;
    }
#undef T
#undef L
}

void orc_h264_pred4x4(const uint8_t top8[8], const uint8_t left4[4], int corner, int mode, int avail_top,
                      int avail_left, uint8_t out[16])
{
    int t[8], l[4];
    for (int i = 0; i < 8; i++) t[i] = top8[i];
    for (int i = 0; i < 4; i++) l[i] = left4[i];
    int dc;
    if (avail_top && avail_left) dc = (t[0] + t[1] + t[2] + t[3] + l[0] + l[1] + l[2] + l[3] + 4) >> 3;
    else if (avail_left) dc = (l[0] + l[1] + l[2] + l[3] + 2) >> 2;
    else if (avail_top) dc = (t[0] + t[1] + t[2] + t[3] + 2) >> 2;
    else dc = 128;
    for (int y = 0; y < 4; y++)
        for (int x = 0; x < 4; x++) out[y * 4 + x] = (uint8_t)dir_sample(4, mode, x, y, t, l, corner, dc);
}

/* 8.3.2.2.1 reference filtering, then the modes on the filtered samples. */
void orc_h264_pred8x8(const uint8_t top16[16], const uint8_t left8[8], int corner, int mode, int avail_top,
                      int avail_left, int avail_corner, int avail_top_right, uint8_t out[64])
{
    int t[16], l[8], ft[16], fl[8], fc = corner;
    for (int i = 0; i < 16; i++) t[i] = (i >= 8 && avail_top && !avail_top_right) ? top16[7] : top16[i];
    for (int i = 0; i < 8; i++) l[i] = left8[i];
    for (int i = 0; i < 16; i++) ft[i] = t[i];
    for (int i = 0; i < 8; i++) fl[i] = l[i];
    if (avail_top) {
        ft[0] = avail_corner ? f3(corner, t[0], t[1]) : (3 * t[0] + t[1] + 2) >> 2;
        for (int i = 1; i < 15; i++) ft[i] = f3(t[i - 1], t[i], t[i + 1]);
        ft[15] = (t[14] + 3 * t[15] + 2) >> 2;
    }
    if (avail_left) {
        fl[0] = avail_corner ? f3(corner, l[0], l[1]) : (3 * l[0] + l[1] + 2) >> 2;
        for (int i = 1; i < 7; i++) fl[i] = f3(l[i - 1], l[i], l[i + 1]);
        fl[7] = (l[6] + 3 * l[7] + 2) >> 2;
    }
    if (avail_corner && avail_top && avail_left) fc = f3(t[0], corner, l[0]);
    int dc;
    int st = 0, sl = 0;
    for (int i = 0; i < 8; i++) { st += ft[i]; sl += fl[i]; }
    if (avail_top && avail_left) dc = (st + sl + 8) >> 4;
    else if (avail_top) dc = (st + 4) >> 3;
    else if (avail_left) dc = (sl + 4) >> 3;
    else dc = 128;
    for (int y = 0; y < 8; y++)
        for (int x = 0; x < 8; x++) out[y * 8 + x] = (uint8_t)dir_sample(8, mode, x, y, ft, fl, fc, dc);
}

/* ------------------------------------------------------------------------
 * 8x8 transform (High profile).  The reference encoder codes only 4x4
 * blocks; this is the decoder's side of 8.5.13 (normative) with an encoder
 * forward transform and quantiser chosen to invert it:
 *   forward   Y = C X C^T, C the 8x8 integer basis below (exact, no rounding)
 *   quant     level = floor((|Y| * 2^14 + off) / step), step(i,j) =
 *             n_i n_j v8[qp%6][class] 2^(qp/6) with n = diag(C C^T) and off
 *             = round_q12/12 of a step: the algebraic inverse of the scaling
 *             below followed by the inverse transform
 *   dequant   8.5.13.1 with a flat weight scale (16): d = (c * 16 * v8 <<
 *             qp/6) >> 6, rounding below qp 36
 *   inverse   8.5.13.2 butterflies, rows then columns, (x + 32) >> 6
 * Reference decoder: enc/src/h264_recon.c:87-93 (scale8), :150-210.
 * ---------------------------------------------------------------------- */
static const int kC8[8][8] = {
    {8, 8, 8, 8, 8, 8, 8, 8},         {12, 10, 6, 3, -3, -6, -10, -12}, {8, 4, -4, -8, -8, -4, 4, 8},
    {10, -3, -12, -6, 6, 12, 3, -10}, {8, -8, -8, 8, 8, -8, -8, 8},     {6, -12, 3, 10, -10, -3, 12, -6},
    {4, -8, 8, -4, -4, 8, -8, 4},     {3, -6, 10, -12, 12, -10, 6, -3}};
static const int kN8[8] = {512, 578, 320, 578, 512, 578, 320, 578};
/* normAdjust8x8 (8.5.9) v[qp%6][class] */
static const int kV8[6][6] = {{20, 18, 32, 19, 25, 24}, {22, 19, 35, 21, 28, 26}, {26, 23, 42, 24, 33, 31},
                              {28, 25, 45, 26, 35, 33}, {32, 28, 51, 30, 40, 38}, {36, 32, 58, 34, 46, 43}};

/* 8.5.9 position class of raster position (i row, j column) of an 8x8 block */
int orc_h264_class8(int i, int j)
{
    if (i % 4 == 0 && j % 4 == 0) return 0;
    if (i % 2 == 1 && j % 2 == 1) return 1;
    if (i % 4 == 2 && j % 4 == 2) return 2;
    if ((i % 4 == 0 && j % 2 == 1) || (i % 2 == 1 && j % 4 == 0)) return 3;
    if ((i % 4 == 0 && j % 4 == 2) || (i % 4 == 2 && j % 4 == 0)) return 4;
    return 5;
}

int orc_h264_v8(int qp_rem, int cls) { return kV8[qp_rem][cls]; }

void orc_h264_fwd8x8(const int32_t in[64], int64_t out[64])
{
    int64_t t[64];
    for (int i = 0; i < 8; i++)
        for (int c = 0; c < 8; c++) {
            int64_t s = 0;
            for (int r = 0; r < 8; r++) s += (int64_t)kC8[i][r] * in[r * 8 + c];
            t[i * 8 + c] = s;
        }
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 8; j++) {
            int64_t s = 0;
            for (int c = 0; c < 8; c++) s += (int64_t)kC8[j][c] * t[i * 8 + c];
            out[i * 8 + j] = s;
        }
}

void orc_h264_quant8x8(const int64_t y[64], int qp, int round_q12, int16_t out[64])
{
    for (int p = 0; p < 64; p++) {
        const int i = p >> 3, j = p & 7;
        const uint64_t step = (uint64_t)kN8[i] * kN8[j] * kV8[qp % 6][orc_h264_class8(i, j)] << (qp / 6);
        const uint64_t off = step * (uint64_t)round_q12 / 12;
        const uint64_t mag = (uint64_t)(y[p] < 0 ? -y[p] : y[p]);
        uint64_t l = ((mag << 14) + off) / step;
        if (l > 32767) l = 32767;
        out[p] = (int16_t)(y[p] < 0 ? -(int64_t)l : (int64_t)l);
    }
}

void orc_h264_dequant8x8(const int16_t c[64], int qp, int32_t d[64])
{
    const int per = qp / 6;
    for (int p = 0; p < 64; p++) {
        const int32_t ls = 16 * kV8[qp % 6][orc_h264_class8(p >> 3, p & 7)];
        d[p] = per >= 6 ? (c[p] * ls) * (1 << (per - 6)) : (c[p] * ls + (1 << (5 - per))) >> (6 - per);
    }
}

/* 8.5.13.2 one-dimensional inverse over v[0..7] at stride s, in place */
static void inv8_1d(int32_t *v, int s)
{
    const int32_t d0 = v[0], d1 = v[s], d2 = v[2 * s], d3 = v[3 * s], d4 = v[4 * s], d5 = v[5 * s],
                  d6 = v[6 * s], d7 = v[7 * s];
    const int32_t e0 = d0 + d4, e1 = -d3 + d5 - d7 - (d7 >> 1), e2 = d0 - d4, e3 = d1 + d7 - d3 - (d3 >> 1);
    const int32_t e4 = (d2 >> 1) - d6, e5 = -d1 + d7 + d5 + (d5 >> 1), e6 = d2 + (d6 >> 1),
                  e7 = d3 + d5 + d1 + (d1 >> 1);
    const int32_t f0 = e0 + e6, f1 = e1 + (e7 >> 2), f2 = e2 + e4, f3 = e3 + (e5 >> 2);
    const int32_t f4 = e2 - e4, f5 = (e3 >> 2) - e5, f6 = e0 - e6, f7 = e7 - (e1 >> 2);
    v[0] = f0 + f7;
    v[s] = f2 + f5;
    v[2 * s] = f4 + f3;
    v[3 * s] = f6 + f1;
    v[4 * s] = f6 - f1;
    v[5 * s] = f4 - f3;
    v[6 * s] = f2 - f5;
    v[7 * s] = f0 - f7;
}

void orc_h264_inv8x8(const int32_t d[64], int32_t out[64])
{
    int32_t t[64];
    for (int p = 0; p < 64; p++) t[p] = d[p];
    for (int i = 0; i < 8; i++) inv8_1d(t + i * 8, 1);
    for (int i = 0; i < 8; i++) inv8_1d(t + i, 8);
    for (int p = 0; p < 64; p++) out[p] = (t[p] + 32) >> 6;
}

/* residual -> levels, and levels -> reconstructed residual */
void orc_h264_tq8x8(const int16_t res[64], int qp, int round_q12, int16_t out[64])
{
    int32_t x[64];
    int64_t y[64];
    for (int p = 0; p < 64; p++) x[p] = res[p];
    orc_h264_fwd8x8(x, y);
    orc_h264_quant8x8(y, qp, round_q12, out);
}

void orc_h264_dqit8x8(const int16_t lv[64], int qp, int16_t out[64])
{
    int32_t d[64], r[64];
    orc_h264_dequant8x8(lv, qp, d);
    orc_h264_inv8x8(d, r);
    for (int p = 0; p < 64; p++) out[p] = (int16_t)r[p];
}

/* ---------------------------------------------------------------------------
 * The reference's Tier-2 motion search, cpu_simd_me_search_frame
 * (src/cpu_simd_me.c:275-497; the one the H.264 encoder runs on the CPU and
 * hands to dispatch_encode_ext as cpu_mvs, encoder_h264.c:2360-2382).  Per
 * 16x16 MB, in raster order (the left neighbour's result feeds the next MB):
 *   1. SAD at (0,0); <= 512 -> (0,0).
 *   2. Left neighbour's vector / 4 (C division, towards zero), if non-zero,
 *      within +-R and inside the MB-padded frame: cost SAD + 5*|mv|; if it
 *      beats the zero SAD and costs < 768, take it and stop.
 *   3. Square search around the best, steps R/2, R/4, .. 1, the 8 offsets
 *      (0,1) (0,-1) (1,0) (-1,0) (1,1) (-1,-1) (1,-1) (-1,1), cost SAD +
 *      5*|mv|, strict '<'; stop early at cost 0.
 *   4. If 128 < cost < 3000 and the block has a one-sample margin, the four
 *      half-sample candidates (+-2, 0) (0, +-2) (quarter units) predicted by
 *      the rounded-up average of two neighbouring integer samples, compared
 *      by plain SAD against the step-3 cost; strict '<'.
 * R = search_radius clamped to 2..16.  out: {mvx, mvy (quarter samples),
 * cost, 0} per MB.  Planes hold at least the MB-padded frame.
 * ------------------------------------------------------------------------- */
static uint32_t sad16x16(const uint8_t *a, int ap, const uint8_t *b, int bp)
{
    uint32_t s = 0;
    for (int r = 0; r < 16; r++)
        for (int c = 0; c < 16; c++) s += (uint32_t)abs((int)a[r * ap + c] - (int)b[r * bp + c]);
    return s;
}

void orc_square_me(const uint8_t *src, int sp, const uint8_t *ref, int rp, int width, int height, int search_radius,
                   int32_t *out)
{
    static const int pat[8][2] = {{0, 1}, {0, -1}, {1, 0}, {-1, 0}, {1, 1}, {-1, -1}, {1, -1}, {-1, 1}};
    const int wmb = (width + 15) / 16, hmb = (height + 15) / 16;
    const int R = search_radius < 2 ? 2 : (search_radius > 16 ? 16 : search_radius);
    const int max_x = wmb * 16, max_y = hmb * 16;
    const uint32_t lambda = 5;
    for (int mby = 0; mby < hmb; mby++)
        for (int mbx = 0; mbx < wmb; mbx++) {
            const int mb = mby * wmb + mbx, px = mbx * 16, py = mby * 16;
            int32_t *o = out + 4 * mb;
            const uint8_t *cur = src + (size_t)py * sp + px;
            const uint32_t zero = sad16x16(cur, sp, ref + (size_t)py * rp + px, rp);
            if (zero <= 512) {
                o[0] = o[1] = 0;
                o[2] = (int32_t)zero;
                o[3] = 0;
                continue;
            }
            int bx = 0, by = 0;
            uint32_t best = zero;
            if (mbx > 0) {
                const int lx = o[-4] / 4, ly = o[-3] / 4;
                if ((lx || ly) && abs(lx) <= R && abs(ly) <= R) {
                    const int tx = px + lx, ty = py + ly;
                    if (tx >= 0 && tx + 16 <= max_x && ty >= 0 && ty + 16 <= max_y) {
                        const uint32_t c = sad16x16(cur, sp, ref + (size_t)ty * rp + tx, rp) +
                                           lambda * (uint32_t)(abs(lx) + abs(ly));
                        if (c < best) {
                            best = c;
                            bx = lx;
                            by = ly;
                        }
                    }
                }
            }
            if ((bx || by) && best < 768) {
                o[0] = bx * 4;
                o[1] = by * 4;
                o[2] = (int32_t)best;
                o[3] = 0;
                continue;
            }
            for (int step = R / 2; step >= 1; step /= 2) {
                const int cx = bx, cy = by;
                for (int k = 0; k < 8; k++) {
                    const int x = cx + pat[k][0] * step, y = cy + pat[k][1] * step;
                    if (abs(x) > R || abs(y) > R) continue;
                    const int tx = px + x, ty = py + y;
                    if (tx < 0 || tx + 16 > max_x || ty < 0 || ty + 16 > max_y) continue;
                    const uint32_t c = sad16x16(cur, sp, ref + (size_t)ty * rp + tx, rp) + lambda * (uint32_t)(abs(x) + abs(y));
                    if (c < best) {
                        best = c;
                        bx = x;
                        by = y;
                    }
                }
                if (best == 0) break;
            }
            int qx = bx * 4, qy = by * 4;
            if (best > 128 && best < 3000) {
                const int ox = px + bx, oy = py + by;
                if (ox > 0 && ox + 17 < max_x && oy > 0 && oy + 17 < max_y) {
                    static const int hp[4][2] = {{2, 0}, {-2, 0}, {0, 2}, {0, -2}};
                    int sx_best = 0, sy_best = 0;
                    uint32_t low = best;
                    for (int h = 0; h < 4; h++) {
                        const int dx = hp[h][0] > 0 ? 1 : (hp[h][0] < 0 ? -1 : 0);
                        const int dy = hp[h][1] > 0 ? 1 : (hp[h][1] < 0 ? -1 : 0);
                        uint32_t s = 0;
                        for (int r = 0; r < 16; r++)
                            for (int c = 0; c < 16; c++) {
                                const int a = ref[(size_t)(oy + r) * rp + ox + c];
                                const int b = ref[(size_t)(oy + r + dy) * rp + ox + c + dx];
                                s += (uint32_t)abs((int)cur[r * sp + c] - ((a + b + 1) >> 1));
                            }
                        if (s < low) {
                            low = s;
                            sx_best = hp[h][0];
                            sy_best = hp[h][1];
                        }
                    }
                    qx += sx_best;
                    qy += sy_best;
                    best = low;
                }
            }
            o[0] = qx;
            o[1] = qy;
            o[2] = (int32_t)best;
            o[3] = 0;
        }
}

/* ---------------------------------------------------------------------------
 * The reference's GPU motion search, motion_estimation.comp:278-469, with
 * exact integer SADs in place of its float-normalised, truncated ones.  Per
 * MB (independent of every other MB):
 *   1. full SAD at (0,0) (clamped reference); <= 512 -> (0,0) with that SAD.
 *   2. diamond (0,1) (0,-1) (1,0) (-1,0) at steps 8, 4, 2, 1 around the step's
 *      centre, within +-16; cost = 2 x (SAD over the pixels with even x + y)
 *      + lambda * |mv|, compared with the best cost so far (the zero SAD at
 *      first), strict '<'; stop after a step whose best cost is < 768.
 *   3. 8 half-sample candidates (+-2,0) (0,+-2) (2,2) (2,-2) (-2,2) (-2,-2)
 *      around the integer winner, then 8 quarter-sample ones (+-1 ...) around
 *      the half winner, plain SAD of the 8.4.2.2.1 samples, strict '<'
 *      against the running best (the step-2 cost at first).
 * lambda 5, or 8 in the reference's fast mode.  out: {mvx, mvy (quarter
 * samples), cost, 0} per MB.  Planes are coded-size (edges replicated).
 * ------------------------------------------------------------------------- */
void orc_diamond_me(const uint8_t *cur, const uint8_t *ref, int py, int wc, int hc, int lambda, int32_t *out)
{
    static const int dia[4][2] = {{0, 1}, {0, -1}, {1, 0}, {-1, 0}};
    static const int half[8][2] = {{2, 0}, {-2, 0}, {0, 2}, {0, -2}, {2, 2}, {2, -2}, {-2, 2}, {-2, -2}};
    const int wmb = wc / 16, hmb = hc / 16;
    for (int mby = 0; mby < hmb; mby++)
        for (int mbx = 0; mbx < wmb; mbx++) {
            const int x0 = mbx * 16, y0 = mby * 16;
            int32_t *o = out + 4 * (mby * wmb + mbx);
#define REFC(x, y) ref[(size_t)clampi((y), 0, hc - 1) * py + clampi((x), 0, wc - 1)]
            uint32_t zero = 0;
            for (int r = 0; r < 16; r++)
                for (int c = 0; c < 16; c++) zero += (uint32_t)abs(cur[(size_t)(y0 + r) * py + x0 + c] - REFC(x0 + c, y0 + r));
            if (zero <= 512) {
                o[0] = o[1] = 0;
                o[2] = (int32_t)zero;
                o[3] = 0;
                continue;
            }
            int bx = 0, by = 0;
            uint32_t best = zero;
            for (int step = 8; step >= 1; step /= 2) {
                const int cx = bx, cy = by;
                for (int k = 0; k < 4; k++) {
                    const int x = cx + dia[k][0] * step, y = cy + dia[k][1] * step;
                    if (abs(x) > 16 || abs(y) > 16) continue;
                    uint32_t s = 0;
                    for (int r = 0; r < 16; r++)
                        for (int c = 0; c < 16; c++)
                            if (((r + c) & 1) == 0)
                                s += 2u * (uint32_t)abs(cur[(size_t)(y0 + r) * py + x0 + c] - REFC(x0 + c + x, y0 + r + y));
                    const uint32_t cost = s + (uint32_t)lambda * (uint32_t)(abs(x) + abs(y));
                    if (cost < best) {
                        best = cost;
                        bx = x;
                        by = y;
                    }
                }
                if (best < 768) break;
            }
#undef REFC
            int qx = bx * 4, qy = by * 4;
            for (int stage = 0; stage < 2; stage++) {
                const int ax = qx, ay = qy;
                int sx = 0, sy = 0;
                for (int k = 0; k < 8; k++) {
                    const int ox = stage ? half[k][0] / 2 : half[k][0], oy = stage ? half[k][1] / 2 : half[k][1];
                    uint32_t s = 0;
                    for (int r = 0; r < 16; r++)
                        for (int c = 0; c < 16; c++)
                            s += (uint32_t)abs(cur[(size_t)(y0 + r) * py + x0 + c] -
                                               orc_luma_qpel(ref, py, wc, hc, (x0 + c) * 4 + ax + ox, (y0 + r) * 4 + ay + oy));
                    if (s < best) {
                        best = s;
                        sx = ox;
                        sy = oy;
                    }
                }
                qx = ax + sx;
                qy = ay + sy;
            }
            o[0] = qx;
            o[1] = qy;
            o[2] = (int32_t)best;
            o[3] = 0;
        }
}
