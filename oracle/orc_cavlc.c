/* CPU restatement of the H.264 CAVLC slice writer of the reference's
 * compute-encoder path -- TEST INFRASTRUCTURE.  Only tests/, smoke() and
 * bench.py's cpu_baseline leg load it, as the checker.
 *
 * Follows src/encoder_h264.c:3031-3240 (slice header, I16x16 / P16x16 /
 * P_Skip macroblock loop, trailing bits), encode_mb_i16x16 / encode_mb_p16x16
 * (src/encoder_h264.c:1121-1329), the nC neighbour rule (luma_nc / chroma_nc,
 * :800-842), the motion vector predictors (mv_predictor / skip_mv_predictor,
 * :940-1058) and the residual writers of src/cavlc.c:75-796 (H.264 9.2,
 * Tables 9-5, 9-7, 9-8, 9-9, 9-10).  The code tables are held here as the
 * standard's codeword strings.
 *
 * Pinned by: the reference's own src/cavlc.c + src/bitstream.c compiled
 * into oracle/_ref (block writers compared bit for bit on random and
 * exhaustive-pattern blocks), the known answers of the reference's
 * tests/test_cavlc.c, and a slice parser written from H.264 7.3.4 / 7.3.5 /
 * 9.2 / 8.4.1.3 that decodes every written slice back to its inputs.
 */
#include <stdlib.h>
#include <string.h>

#include "orc.h"

/* ------------------------------------------------------------ bit writer */
typedef struct {
    uint8_t *buf;
    int64_t cap;   /* bytes */
    int64_t bits;  /* bits written */
    int overflow;
} obits;

static void put(obits *b, int n, uint32_t v)
{
    for (int i = n - 1; i >= 0; i--) {
        const int64_t byte = b->bits >> 3;
        if (byte >= b->cap) { b->overflow = 1; return; }
        const int bit = (v >> i) & 1;
        if ((b->bits & 7) == 0) b->buf[byte] = 0;
        b->buf[byte] |= (uint8_t)(bit << (7 - (b->bits & 7)));
        b->bits++;
    }
}

static void put_str(obits *b, const char *s)
{
    for (; *s; s++) put(b, 1, *s == '1');
}

static void put_ue(obits *b, uint32_t v)
{
    const uint64_t cn = (uint64_t)v + 1;
    int k = 0;
    while ((cn >> (k + 1)) != 0) k++;
    put(b, k, 0);
    /* k + 1 bits of cn: at most 33, written as two pieces */
    if (k + 1 > 32) { put(b, 1, 1); put(b, 32, (uint32_t)cn); }
    else put(b, k + 1, (uint32_t)cn);
}

static void put_se(obits *b, int32_t v)
{
    put_ue(b, v <= 0 ? (uint32_t)(-(int64_t)v * 2) : (uint32_t)v * 2u - 1u);
}

/* ------------------------------------------------------------------ tables */
/* coeff_token (Table 9-5) for 0 <= nC < 2, 2 <= nC < 4 and 4 <= nC < 8,
 * [TotalCoeff][TrailingOnes]; "" where the pair cannot occur.  nC >= 8 is
 * the 6-bit fixed-length form and nC == -1 (chroma DC) has its own table. */
static const char *const CT[3][17][4] = {
{
    {"1", "", "", ""},
    {"000101", "01", "", ""},
    {"00000111", "000100", "001", ""},
    {"000000111", "00000110", "0000101", "00011"},
    {"0000000111", "000000110", "00000101", "000011"},
    {"00000000111", "0000000110", "000000101", "0000100"},
    {"0000000001111", "00000000110", "0000000101", "00000100"},
    {"0000000001011", "0000000001110", "00000000101", "000000100"},
    {"0000000001000", "0000000001010", "0000000001101", "0000000100"},
    {"00000000001111", "00000000001110", "0000000001001", "00000000100"},
    {"00000000001011", "00000000001010", "00000000001101", "0000000001100"},
    {"000000000001111", "000000000001110", "00000000001001", "00000000001100"},
    {"000000000001011", "000000000001010", "000000000001101", "00000000001000"},
    {"0000000000001111", "000000000000001", "000000000001001", "000000000001100"},
    {"0000000000001011", "0000000000001110", "0000000000001101", "000000000001000"},
    {"0000000000000111", "0000000000001010", "0000000000001001", "0000000000001100"},
    {"0000000000000100", "0000000000000110", "0000000000000101", "0000000000001000"},
},
{
    {"11", "", "", ""},
    {"001011", "10", "", ""},
    {"000111", "00111", "011", ""},
    {"0000111", "001010", "001001", "0101"},
    {"00000111", "000110", "000101", "0100"},
    {"00000100", "0000110", "0000101", "00110"},
    {"000000111", "00000110", "00000101", "001000"},
    {"00000001111", "000000110", "000000101", "000100"},
    {"00000001011", "00000001110", "00000001101", "0000100"},
    {"000000001111", "00000001010", "00000001001", "000000100"},
    {"000000001011", "000000001110", "000000001101", "00000001100"},
    {"000000001000", "000000001010", "000000001001", "00000001000"},
    {"0000000001111", "0000000001110", "0000000001101", "000000001100"},
    {"0000000001011", "0000000001010", "0000000001001", "0000000001100"},
    {"0000000000111", "00000000001011", "0000000000110", "0000000001000"},
    {"00000000001001", "00000000001000", "00000000001010", "0000000000001"},
    {"00000000000111", "00000000000110", "00000000000101", "00000000000100"},
},
{
    {"1111", "", "", ""},
    {"001111", "1110", "", ""},
    {"001011", "01111", "1101", ""},
    {"001000", "01100", "01110", "1100"},
    {"0001111", "01010", "01011", "1011"},
    {"0001011", "01000", "01001", "1010"},
    {"0001001", "001110", "001101", "1001"},
    {"0001000", "001010", "001001", "1000"},
    {"00001111", "0001110", "0001101", "01101"},
    {"00001011", "00001110", "0001010", "001100"},
    {"000001111", "00001010", "00001101", "0001100"},
    {"000001011", "000001110", "00001001", "00001100"},
    {"000001000", "000001010", "000001101", "00001000"},
    {"0000001101", "000000111", "000001001", "000001100"},
    {"0000001001", "0000001100", "0000001011", "0000001010"},
    {"0000000101", "0000001000", "0000000111", "0000000110"},
    {"0000000001", "0000000100", "0000000011", "0000000010"},
},
};
static const char *const CT_CDC[5][4] = {
    {"01", "", "", ""},
    {"000111", "1", "", ""},
    {"000100", "000110", "001", ""},
    {"000011", "0000011", "0000010", "000101"},
    {"000010", "00000011", "00000010", "0000000"},
};
/* total_zeros (Tables 9-7, 9-8), [TotalCoeff - 1][total_zeros] */
static const char *const TZ[15][16] = {
    {"1", "011", "010", "0011", "0010", "00011", "00010", "000011", "000010", "0000011", "0000010", "00000011", "00000010", "000000011", "000000010", "000000001"},
    {"111", "110", "101", "100", "011", "0101", "0100", "0011", "0010", "00011", "00010", "000011", "000010", "000001", "000000"},
    {"0101", "111", "110", "101", "0100", "0011", "100", "011", "0010", "00011", "00010", "000001", "00001", "000000"},
    {"00011", "111", "0101", "0100", "110", "101", "100", "0011", "011", "0010", "00010", "00001", "00000"},
    {"0101", "0100", "0011", "111", "110", "101", "100", "011", "0010", "00001", "0001", "00000"},
    {"000001", "00001", "111", "110", "101", "100", "011", "010", "0001", "001", "000000"},
    {"000001", "00001", "101", "100", "011", "11", "010", "0001", "001", "000000"},
    {"000001", "0001", "00001", "011", "11", "10", "010", "001", "000000"},
    {"000001", "000000", "0001", "11", "10", "001", "01", "00001"},
    {"00001", "00000", "001", "11", "10", "01", "0001"},
    {"0000", "0001", "001", "010", "1", "011"},
    {"0000", "0001", "01", "1", "001"},
    {"000", "001", "1", "01"},
    {"00", "01", "1"},
    {"0", "1"},
};
/* chroma DC total_zeros (Table 9-9a), [TotalCoeff - 1][total_zeros] */
static const char *const TZ_CDC[3][4] = {
    {"1", "01", "001", "000"},
    {"1", "01", "00"},
    {"1", "0"},
};
/* run_before (Table 9-10), [min(zerosLeft, 7) - 1][run_before] */
static const char *const RB[7][15] = {
    {"1", "0"},
    {"1", "01", "00"},
    {"11", "10", "01", "00"},
    {"11", "10", "01", "001", "000"},
    {"11", "10", "011", "010", "001", "000"},
    {"11", "000", "001", "011", "010", "101", "100"},
    {"111", "110", "101", "100", "011", "010", "001", "0001", "00001", "000001", "0000001", "00000001", "000000001", "0000000001", "00000000001"},
};

static const int ZZ[16] = {0, 1, 4, 8, 5, 2, 3, 6, 9, 12, 13, 10, 7, 11, 14, 15};

/* One level (9.2.2.1 inverted), returning the next suffixLength
 * (src/cavlc.c:429-480). */
static int put_level(obits *b, int level, int first_after_lt3_t1, int sl)
{
    int code = level > 0 ? 2 * level - 2 : -2 * level - 1;
    if (first_after_lt3_t1) code -= 2;
    if (sl == 0 && code < 14) {
        put(b, code, 0);
        put(b, 1, 1);
    } else if (sl == 0 && code < 30) {
        put(b, 14, 0);
        put(b, 1, 1);
        put(b, 4, (uint32_t)(code - 14));
    } else if (sl > 0 && (code >> sl) < 15) {
        put(b, code >> sl, 0);
        put(b, 1, 1);
        put(b, sl, (uint32_t)(code & ((1 << sl) - 1)));
    } else {
        /* level_prefix >= 15: the escape carries levelCode minus what the
         * shorter prefixes cover, in a suffix of level_prefix - 3 bits */
        const uint32_t rest = (uint32_t)(code - (sl == 0 ? 30 : (15 << sl))) + 4096u;
        int k = 12;
        while ((rest >> (k + 1)) != 0) k++;
        put(b, k + 3, 0);
        put(b, 1, 1);
        put(b, k, rest - (1u << k));
    }
    const int a = level < 0 ? -level : level;
    if (sl == 0) sl = 1;
    if (a > (3 << (sl - 1)) && sl < 6) sl++;
    return sl;
}

/* residual_block_cavlc (7.3.5.3.2) of `n` coefficients in scan order; nc
 * -1 is the chroma DC form.  Returns TotalCoeff (src/cavlc.c:603-796). */
static int put_residual(obits *b, const int *scan, int n, int nc)
{
    int tc = 0, t1 = 0, signs = 0, nlev = 0, levels[16], runs[16], hi = -1, prev = -1;
    for (int i = n - 1; i >= 0; i--) {
        if (scan[i] == 0) continue;
        if (hi < 0) hi = i;
        if (prev >= 0) runs[tc - 1] = prev - i - 1;
        if (tc < 3 && t1 == tc && (scan[i] == 1 || scan[i] == -1)) {
            t1++;
            signs = signs << 1 | (scan[i] < 0);
        } else {
            levels[nlev++] = scan[i];
        }
        tc++;
        prev = i;
    }
    const int tz = hi < 0 ? 0 : hi + 1 - tc;
    if (nc < 0) put_str(b, CT_CDC[tc][t1]);
    else if (nc >= 8) put(b, 6, tc == 0 ? 3u : (uint32_t)((tc - 1) << 2 | t1));
    else put_str(b, CT[nc < 2 ? 0 : nc < 4 ? 1 : 2][tc][t1]);
    if (tc == 0) return 0;
    put(b, t1, (uint32_t)signs);
    int sl = tc > 10 && t1 < 3 ? 1 : 0;
    for (int i = 0; i < nlev; i++) sl = put_level(b, levels[i], i == 0 && t1 < 3, sl);
    if (tc < n) put_str(b, nc < 0 ? TZ_CDC[tc - 1][tz] : TZ[tc - 1][tz]);
    int left = tz;
    for (int i = 0; i + 1 < tc && left > 0; i++) {
        put_str(b, RB[(left < 7 ? left : 7) - 1][runs[i]]);
        left -= runs[i];
    }
    return tc;
}

int orc_cavlc_block(const int16_t coef[16], int ac, int nc, uint8_t *buf, int cap, int *nbits)
{
    int scan[16];
    for (int i = ac; i < 16; i++) scan[i - ac] = coef[ZZ[i]];
    obits b = {buf, cap, 0, 0};
    const int tc = put_residual(&b, scan, 16 - ac, nc);
    *nbits = (int)b.bits;
    return b.overflow ? -1 : tc;
}

int orc_cavlc_chroma_dc(const int32_t c[4], uint8_t *buf, int cap, int *nbits)
{
    int scan[4] = {c[0], c[1], c[2], c[3]};
    obits b = {buf, cap, 0, 0};
    const int tc = put_residual(&b, scan, 4, -1);
    *nbits = (int)b.bits;
    return b.overflow ? -1 : tc;
}

int orc_exp_golomb(int value, int is_signed, uint8_t *buf, int cap, int *nbits)
{
    obits b = {buf, cap, 0, 0};
    if (is_signed) put_se(&b, value);
    else put_ue(&b, (uint32_t)value);
    *nbits = (int)b.bits;
    return b.overflow ? -1 : 0;
}

/* ------------------------------------------------------------ macroblocks */
/* 4x4 luma block blkIdx (6.4.3) -> its (x, y) in 4-sample units */
static int bx4(int k) { return (k & 1) | (k & 4) >> 1; }
static int by4(int k) { return (k & 2) >> 1 | (k & 8) >> 2; }
static int blk_at(int x, int y) { return (x & 1) | (y & 1) << 1 | (x & 2) << 1 | (y & 2) << 2; }

typedef struct {
    int wmb, start;
    uint8_t (*nzl)[16];   /* TotalCoeff per luma blkIdx */
    uint8_t (*nzc)[2][4]; /* per chroma component and 4x4 block */
} nctx;

static int nc_avg(int na, int nb)
{
    if (na >= 0 && nb >= 0) return (na + nb + 1) >> 1;
    return na >= 0 ? na : nb >= 0 ? nb : 0;
}

/* nC of luma block k (9.2.1 with the slice as the availability bound;
 * src/encoder_h264.c:800-821) */
static int luma_nc(const nctx *c, int mb, int k)
{
    const int x = bx4(k), y = by4(k), mbx = mb % c->wmb;
    int na = -1, nb = -1;
    if (x > 0) na = c->nzl[mb][blk_at(x - 1, y)];
    else if (mbx > 0 && mb - 1 >= c->start) na = c->nzl[mb - 1][blk_at(3, y)];
    if (y > 0) nb = c->nzl[mb][blk_at(x, y - 1)];
    else if (mb - c->wmb >= c->start) nb = c->nzl[mb - c->wmb][blk_at(x, 3)];
    return nc_avg(na, nb);
}

static int chroma_nc(const nctx *c, int mb, int comp, int k)
{
    const int x = k & 1, y = k >> 1, mbx = mb % c->wmb;
    int na = -1, nb = -1;
    if (x > 0) na = c->nzc[mb][comp][k - 1];
    else if (mbx > 0 && mb - 1 >= c->start) na = c->nzc[mb - 1][comp][k + 1];
    if (y > 0) nb = c->nzc[mb][comp][k - 2];
    else if (mb - c->wmb >= c->start) nb = c->nzc[mb - c->wmb][comp][k + 2];
    return nc_avg(na, nb);
}

/* neighbour (dx, dy) of mb inside the picture and the slice */
static int nb_mv(const orc_mv_t *mv, int wmb, int start, int mb, int dx, int dy, int *x, int *y)
{
    const int nx = mb % wmb + dx, ny = mb / wmb + dy;
    *x = *y = 0;
    if (nx < 0 || ny < 0 || nx >= wmb) return 0;
    const int n = ny * wmb + nx;
    if (n < start) return 0;
    *x = mv[n].mvx;
    *y = mv[n].mvy;
    return 1;
}

static int med3(int a, int b, int c)
{
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    return c < lo ? lo : c > hi ? hi : c;
}

/* The reference's 16x16 predictor: C, or D when C is missing; A alone when
 * B and C(D) are missing; otherwise the median with missing ones as zero
 * (src/encoder_h264.c:974-1003).  Unlike 8.4.1.3.1 it does not take a lone
 * available B or C by itself. */
void orc_h264_mv_pred(const orc_mv_t *mv, int wmb, int start, int mb, int *px, int *py)
{
    int ax, ay, bx, by, cx, cy;
    const int a = nb_mv(mv, wmb, start, mb, -1, 0, &ax, &ay);
    const int b = nb_mv(mv, wmb, start, mb, 0, -1, &bx, &by);
    int c = nb_mv(mv, wmb, start, mb, 1, -1, &cx, &cy);
    if (!c) c = nb_mv(mv, wmb, start, mb, -1, -1, &cx, &cy);
    if (!b && !c && a) { *px = ax; *py = ay; return; }
    *px = med3(ax, bx, cx);
    *py = med3(ay, by, cy);
}

/* P_Skip vector (8.4.1.1; src/encoder_h264.c:1036-1058) */
void orc_h264_skip_mv(const orc_mv_t *mv, int wmb, int start, int mb, int *px, int *py)
{
    int ax, ay, bx, by;
    const int a = nb_mv(mv, wmb, start, mb, -1, 0, &ax, &ay);
    const int b = nb_mv(mv, wmb, start, mb, 0, -1, &bx, &by);
    if (!a || !b || (ax == 0 && ay == 0) || (bx == 0 && by == 0)) { *px = *py = 0; return; }
    orc_h264_mv_pred(mv, wmb, start, mb, px, py);
}

static const int16_t *blk_of(const int16_t *levels, int mb, int raster) { return levels + ((size_t)mb * 24 + raster) * 16; }

static int any_nz(const int16_t *p, int from)
{
    for (int i = from; i < 16; i++)
        if (p[i]) return 1;
    return 0;
}

/* chroma DC levels of one component: 2x2 Hadamard of the raw DC terms and
 * the DC quantiser at the chroma QP; *raw_nz says whether the transform
 * (before quantisation) has a nonzero term */
static void chroma_dc(const int32_t *dc, int mb, int comp, int qp, int32_t lv[4], int *raw_nz)
{
    int32_t h[4];
    orc_chroma_dc_hadamard(dc + (size_t)mb * 24 + 16 + comp * 4, h);
    const int qpc = orc_chroma_qp(qp);
    for (int i = 0; i < 4; i++) {
        lv[i] = orc_quant_dc_chroma(h[i], qpc);
        if (h[i]) *raw_nz = 1;
    }
}

/* mode sanitising of src/encoder_h264.c:858-887 (16x16: V 0, H 1, DC 2,
 * Plane 3; chroma: DC 0, H 1, V 2, Plane 3) */
static int fix_mode(int m, int top, int left, int dcm, int hm, int vm, int pm)
{
    if (!top && !left) return dcm;
    if (!top && (m == vm || m == pm)) return left ? hm : dcm;
    if (!left && (m == hm || m == pm)) return top ? vm : dcm;
    return m;
}

/* coded chroma AC blocks and their counts, shared by both MB kinds */
static void put_chroma(obits *b, nctx *c, const int16_t *levels, int mb, int cbp_c, const int32_t lv[2][4])
{
    if (cbp_c >= 1)
        for (int comp = 0; comp < 2; comp++) {
            int s[4] = {lv[comp][0], lv[comp][1], lv[comp][2], lv[comp][3]};
            put_residual(b, s, 4, -1);
        }
    for (int comp = 0; comp < 2; comp++)
        for (int k = 0; k < 4; k++) {
            int tc = 0;
            if (cbp_c == 2) {
                int s[15];
                const int16_t *p = blk_of(levels, mb, 16 + comp * 4 + k);
                for (int i = 1; i < 16; i++) s[i - 1] = p[ZZ[i]];
                tc = put_residual(b, s, 15, chroma_nc(c, mb, comp, k));
            }
            c->nzc[mb][comp][k] = (uint8_t)tc;
        }
}

/* inter coded_block_pattern -> codeNum (Table 9-4, Inter column) */
static int cbp_code_inter(int cbp)
{
    static const uint8_t tab[48] = {0,  16, 1,  2,  4,  8,  32, 3,  5,  10, 12, 15, 47, 7,  11, 13,
                                    14, 6,  9,  31, 35, 37, 42, 44, 33, 34, 36, 40, 39, 43, 45, 46,
                                    17, 18, 20, 24, 19, 21, 26, 28, 23, 27, 29, 30, 22, 25, 38, 41};
    for (int i = 0; i < 48; i++)
        if (tab[i] == cbp) return i;
    return cbp;
}

/* mb_type I_16x16_<pm>_<cbp_c>_<ac>, intra_chroma_pred_mode, mb_qp_delta
 * (src/cavlc.c:75-128) */
static void put_i16_header(obits *b, int pm, int cm, int cbp_c, int ac, int dqp)
{
    put_ue(b, (uint32_t)(1 + pm + 4 * cbp_c + 12 * (ac != 0)));
    put_ue(b, (uint32_t)cm);
    put_se(b, dqp);
}

/* mb_type P_L0_16x16, mvd, coded_block_pattern, mb_qp_delta when coded
 * (src/cavlc.c:135-160) */
static void put_p_header(obits *b, int mvdx, int mvdy, int cbp, int dqp)
{
    put_ue(b, 0);
    put_se(b, mvdx);
    put_se(b, mvdy);
    put_ue(b, (uint32_t)cbp_code_inter(cbp));
    if (cbp > 0) put_se(b, dqp);
}

/* kind 0: I16x16 header (a pm, b cm, c cbp_chroma, d cbp_luma), 1: P16x16
 * header (a mvd_x, b mvd_y, c cbp, d qp_delta), 2: mb_skip_run a */
int orc_h264_mb_header(int kind, int a, int b, int c, int d, uint8_t *buf, int cap, int *nbits)
{
    obits o = {buf, cap, 0, 0};
    if (kind == 0) put_i16_header(&o, a, b, c, d, 0);
    else if (kind == 1) put_p_header(&o, a, b, c, d);
    else put_ue(&o, (uint32_t)a);
    *nbits = (int)o.bits;
    return o.overflow ? -1 : 0;
}

static void mb_i16(obits *b, nctx *c, const int16_t *levels, const int32_t *dc, const uint32_t *modes, int mb,
                   int qp)
{
    const int mbx = mb % c->wmb;
    const int left = mbx > 0 && mb - 1 >= c->start, top = mb - c->wmb >= c->start;
    const int pm = fix_mode((int)(modes[mb] & 3), top, left, 2, 1, 0, 3);
    const int cm = fix_mode((int)(modes[mb] >> 2 & 3), top, left, 0, 1, 2, 3);
    int32_t din[16], dl[16], lv[2][4];
    for (int i = 0; i < 16; i++) din[i] = dc[(size_t)mb * 24 + i];
    orc_intra16_luma_dc_levels(din, qp, dl);
    int ac = 0, cac = 0, cdc = 0, raw = 0;
    for (int r = 0; r < 16; r++) ac |= any_nz(blk_of(levels, mb, r), 1);
    chroma_dc(dc, mb, 0, qp, lv[0], &raw);
    chroma_dc(dc, mb, 1, qp, lv[1], &raw);
    for (int i = 0; i < 8; i++) cdc |= lv[i >> 2][i & 3] != 0;
    for (int r = 16; r < 24; r++) cac |= any_nz(blk_of(levels, mb, r), 1);
    const int cbp_c = cac ? 2 : cdc ? 1 : 0;
    put_i16_header(b, pm, cm, cbp_c, ac, 0);
    int s[16];
    for (int i = 0; i < 16; i++) s[i] = dl[ZZ[i]];
    put_residual(b, s, 16, luma_nc(c, mb, 0));
    for (int k = 0; k < 16; k++) {
        int tc = 0;
        if (ac) {
            const int16_t *p = blk_of(levels, mb, by4(k) * 4 + bx4(k));
            for (int i = 1; i < 16; i++) s[i - 1] = p[ZZ[i]];
            tc = put_residual(b, s, 15, luma_nc(c, mb, k));
        }
        c->nzl[mb][k] = (uint8_t)tc;
    }
    put_chroma(b, c, levels, mb, cbp_c, lv);
}

/* P slice MB: returns 1 when coded, 0 when skipped (src/encoder_h264.c:3200-3233) */
static int mb_p(obits *b, nctx *c, const int16_t *levels, const int32_t *dc, const orc_mv_t *mv, int mb, int qp,
                int relax, uint32_t *run)
{
    int zl = 1, zc = 1, raw = 0;
    for (int r = 0; r < 16; r++)
        if (any_nz(blk_of(levels, mb, r), 0)) zl = 0;
    int32_t lv[2][4];
    chroma_dc(dc, mb, 0, qp, lv[0], &raw);
    chroma_dc(dc, mb, 1, qp, lv[1], &raw);
    int cac = 0;
    for (int r = 16; r < 24; r++) cac |= any_nz(blk_of(levels, mb, r), 1);
    if (raw || cac) zc = 0;
    if (relax && zl && !zc && abs(dc[(size_t)mb * 2]) <= 1 && abs(dc[(size_t)mb * 2 + 1]) <= 1) zc = 1;
    int sx, sy;
    orc_h264_skip_mv(mv, c->wmb, c->start, mb, &sx, &sy);
    if (zl && zc && mv[mb].mvx == sx && mv[mb].mvy == sy) {
        memset(c->nzl[mb], 0, 16);
        memset(c->nzc[mb], 0, 8);
        (*run)++;
        return 0;
    }
    put_ue(b, *run);
    *run = 0;
    int px, py;
    orc_h264_mv_pred(mv, c->wmb, c->start, mb, &px, &py);
    int cbp_l = 0, cdc = 0;
    for (int k = 0; k < 16; k++)
        if (any_nz(blk_of(levels, mb, by4(k) * 4 + bx4(k)), 0)) cbp_l |= 1 << (k >> 2);
    for (int i = 0; i < 8; i++) cdc |= lv[i >> 2][i & 3] != 0;
    const int cbp_c = cac ? 2 : cdc ? 1 : 0, cbp = cbp_l | cbp_c << 4;
    put_p_header(b, mv[mb].mvx - px, mv[mb].mvy - py, cbp, 0);
    for (int k = 0; k < 16; k++) {
        int tc = 0;
        if (cbp_l >> (k >> 2) & 1) {
            int s[16];
            const int16_t *p = blk_of(levels, mb, by4(k) * 4 + bx4(k));
            for (int i = 0; i < 16; i++) s[i] = p[ZZ[i]];
            tc = put_residual(b, s, 16, luma_nc(c, mb, k));
        }
        c->nzl[mb][k] = (uint8_t)tc;
    }
    put_chroma(b, c, levels, mb, cbp_c, lv);
    return 1;
}

/* ----------------------------------------------------------------- slices */
/* cfg: width_in_mbs, height_in_mbs, qp, is_idr, num_slices, pps_id,
 * pic_init_qp_minus26, frame_num, log2_max_frame_num_minus4, idr_pic_id,
 * poc, log2_max_poc_lsb_minus4, deblock_idc, relax_chroma_skip.
 * Writes the slices' RBSPs back to back into out; slice s occupies
 * [off[s], off[s] + len[s]).  Returns the total bytes, or -1 when cap is
 * too small. */
int64_t orc_h264_cavlc_slices(const int16_t *levels, const int32_t *dc, const uint32_t *modes,
                              const orc_mv_t *mv, const int32_t *cfg, uint8_t *out, int64_t cap, int64_t *off,
                              int64_t *len)
{
    const int wmb = cfg[0], total = cfg[0] * cfg[1], qp = cfg[2], idr = cfg[3], ns = cfg[4];
    nctx c;
    c.wmb = wmb;
    c.nzl = calloc((size_t)total, 16);
    c.nzc = calloc((size_t)total, 8);
    int64_t pos = 0, rc = 0;
    for (int s = 0; s < ns && rc == 0; s++) {
        const int start = (int)((int64_t)s * total / ns), end = (int)((int64_t)(s + 1) * total / ns);
        c.start = start;
        obits b = {out + pos, cap - pos, 0, 0};
        put_ue(&b, (uint32_t)start);
        put_ue(&b, idr ? 2u : 0u);
        put_ue(&b, (uint32_t)cfg[5]);
        put(&b, cfg[8] + 4, (uint32_t)cfg[7]);
        if (idr) put_ue(&b, (uint32_t)cfg[9]);
        const int pb = cfg[11] + 4;
        put(&b, pb, (uint32_t)cfg[10] & ((1u << pb) - 1));
        /* P: num_ref_idx_active_override, ref_pic_list_modification_l0,
         * adaptive_ref_pic_marking; IDR: no_output_of_prior_pics,
         * long_term_reference -- all 0 */
        put(&b, idr ? 2 : 3, 0);
        put_se(&b, qp - 26 - cfg[6]);
        put_ue(&b, (uint32_t)cfg[12]);
        if (cfg[12] != 1) {
            put_se(&b, 0);
            put_se(&b, 0);
        }
        if (idr) {
            for (int mb = start; mb < end; mb++) mb_i16(&b, &c, levels, dc, modes, mb, qp);
        } else {
            uint32_t run = 0;
            for (int mb = start; mb < end; mb++) mb_p(&b, &c, levels, dc, mv, mb, qp, cfg[13], &run);
            if (run > 0) put_ue(&b, run);
        }
        put(&b, 1, 1);
        if (b.bits & 7) put(&b, 8 - (int)(b.bits & 7), 0);
        if (b.overflow) rc = -1;
        off[s] = pos;
        len[s] = b.bits >> 3;
        pos += b.bits >> 3;
    }
    free(c.nzl);
    free(c.nzc);
    return rc < 0 ? -1 : pos;
}
