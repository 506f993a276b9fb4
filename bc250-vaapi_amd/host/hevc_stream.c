/*
 * hevc_stream.c - the HEVC bitstream of the pipeline's pictures (the
 * interface and the syntax it writes: include/bc250_hevc_stream.h).
 *
 * Parameter sets and slice headers are plain RBSP bit strings; the slice
 * data is CABAC (Rec. ITU-T H.265 9.3): the arithmetic coder below is the
 * encoder the standard family describes for this engine (H.264 9.3.4.2:
 * low / range registers, PutBit with outstanding bits, flush at the end of
 * the slice), the context tables are H.265's (Tables 9-5 to 9-37, 9-46,
 * 9-47) and the binarisations and context selections follow 9.3.3 / 9.3.4.2.
 * The reference writes this part with its own coder on the host after a CPU
 * encode (src/hevc_cabac.c, src/hevc_enc_template.c:596-662,1063-1170,
 * 1409-1497, src/encoder_h265.c:160-345,1419-1486); here the decisions are
 * the GPU's, read from the staging.
 */
#include "bc250_hevc_stream.h"

#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- bits */

typedef struct {
    uint8_t *buf;
    size_t cap, n;
    uint64_t acc;
    int nacc;  /* pending bits in acc (0..7 between calls) */
    int ovf;
} bw_t;

/* n <= 24 bits, v < 2^n: the coder's path, kept small enough to inline. */
static inline void bw_put(bw_t *w, uint32_t v, int n)
{
    w->acc = (w->acc << n) | v;
    w->nacc += n;
    while (w->nacc >= 8) {
        w->nacc -= 8;
        if (w->n < w->cap) w->buf[w->n] = (uint8_t)(w->acc >> w->nacc);
        else w->ovf = 1;
        w->n++;
    }
}

static void bw_bits(bw_t *w, uint32_t v, int n)
{
    if (n <= 0) return;
    if (n > 24) {  /* keep acc within 64 bits */
        bw_bits(w, v >> 16, n - 16);
        bw_bits(w, v & 0xffffu, 16);
        return;
    }
    w->acc = (w->acc << n) | (v & ((1u << n) - 1));
    w->nacc += n;
    while (w->nacc >= 8) {
        w->nacc -= 8;
        if (w->n < w->cap) w->buf[w->n] = (uint8_t)(w->acc >> w->nacc);
        else w->ovf = 1;
        w->n++;
    }
}

static void bw_ue(bw_t *w, uint32_t v)
{
    const uint64_t x = (uint64_t)v + 1;
    int len = 0;
    while ((x >> (len + 1)) != 0) len++;
    bw_bits(w, 0, len);
    if (len + 1 > 32) {
        bw_bits(w, (uint32_t)(x >> 32), len + 1 - 32);
        bw_bits(w, (uint32_t)x, 32);
    } else {
        bw_bits(w, (uint32_t)x, len + 1);
    }
}

static void bw_se(bw_t *w, int32_t v)
{
    bw_ue(w, v > 0 ? 2u * (uint32_t)v - 1u : 2u * (uint32_t)(-(int64_t)v));
}

static void bw_align_zero(bw_t *w)
{
    if (w->nacc) bw_bits(w, 0, 8 - w->nacc);
}

/* rbsp_trailing_bits / byte_alignment(): a one, then zeros to the byte. */
static void bw_trailing(bw_t *w)
{
    bw_bits(w, 1, 1);
    bw_align_zero(w);
}

/* One NAL unit in Annex B form: start code, two-byte header (layer 0,
 * temporal id 0), the RBSP with emulation prevention bytes (7.4.2). */
static size_t nal_out(uint8_t *out, size_t cap, size_t at, int type, const uint8_t *rbsp, size_t n)
{
    size_t o = at;
    const size_t worst = 6 + n + n / 2 + 1;
    if (o + worst > cap) return 0;
    out[o++] = 0; out[o++] = 0; out[o++] = 0; out[o++] = 1;
    out[o++] = (uint8_t)(type << 1);
    out[o++] = 1;
    int zeros = 0;
    for (size_t i = 0; i < n; i++) {
        if (zeros >= 2 && rbsp[i] <= 3) {
            out[o++] = 3;
            zeros = 0;
        }
        out[o++] = rbsp[i];
        zeros = rbsp[i] == 0 ? zeros + 1 : 0;
    }
    return o - at;
}

/* --------------------------------------------------------------- CABAC */

/* Table 9-46: rangeTabLps[pStateIdx][qRangeIdx]. */
static const uint8_t kLps[64][4] = {
    {128, 176, 208, 240}, {128, 167, 197, 227}, {128, 158, 187, 216}, {123, 150, 178, 205}, {116, 142, 169, 195},
    {111, 135, 160, 185}, {105, 128, 152, 175}, {100, 122, 144, 166}, {95, 116, 137, 158},  {90, 110, 130, 150},
    {85, 104, 123, 142},  {81, 99, 117, 135},   {77, 94, 111, 128},   {73, 89, 105, 122},   {69, 85, 100, 116},
    {66, 80, 95, 110},    {62, 76, 90, 104},    {59, 72, 86, 99},     {56, 69, 81, 94},     {53, 65, 77, 89},
    {51, 62, 73, 85},     {48, 59, 69, 80},     {46, 56, 66, 76},     {43, 53, 63, 72},     {41, 50, 59, 69},
    {39, 48, 56, 65},     {37, 45, 54, 62},     {35, 43, 51, 59},     {33, 41, 48, 56},     {32, 39, 46, 53},
    {30, 37, 43, 50},     {29, 35, 41, 48},     {27, 33, 39, 45},     {26, 31, 37, 43},     {24, 30, 35, 41},
    {23, 28, 33, 39},     {22, 27, 32, 37},     {21, 26, 30, 35},     {20, 24, 29, 33},     {19, 23, 27, 31},
    {18, 22, 26, 30},     {17, 21, 25, 28},     {16, 20, 23, 27},     {15, 19, 22, 25},     {14, 18, 21, 24},
    {14, 17, 20, 23},     {13, 16, 19, 22},     {12, 15, 18, 21},     {12, 14, 17, 20},     {11, 14, 16, 19},
    {11, 13, 15, 18},     {10, 12, 15, 17},     {10, 12, 14, 16},     {9, 11, 13, 15},      {9, 11, 12, 14},
    {8, 10, 12, 14},      {8, 9, 11, 13},       {7, 9, 11, 12},       {7, 9, 10, 12},       {7, 8, 10, 11},
    {6, 8, 9, 11},        {6, 7, 9, 10},        {6, 7, 8, 9},         {2, 2, 2, 2}};

/* Table 9-47: transIdxLps (transIdxMps is min(pStateIdx + 1, 62)). */
static const uint8_t kNextLps[64] = {0,  0,  1,  2,  2,  4,  4,  5,  6,  7,  8,  9,  9,  11, 11, 12,
                                     13, 13, 15, 15, 16, 16, 18, 18, 19, 19, 21, 21, 22, 22, 23, 24,
                                     24, 25, 26, 26, 27, 27, 28, 29, 29, 30, 30, 30, 31, 32, 32, 33,
                                     33, 33, 34, 34, 35, 35, 35, 36, 36, 36, 37, 37, 37, 38, 38, 63};

/* The contexts this writer uses, by syntax element (offsets into st[]). */
enum {
    C_SPLIT_CU = 0,   /* split_cu_flag, 3 */
    C_SKIP = 3,       /* cu_skip_flag, 3 */
    C_PRED_MODE = 6,  /* pred_mode_flag */
    C_PART_MODE = 7,  /* part_mode bin 0 */
    C_PREV_INTRA = 8, /* prev_intra_luma_pred_flag */
    C_CHROMA = 9,     /* intra_chroma_pred_mode bin 0 */
    C_MERGE_FLAG = 10,
    C_MERGE_IDX = 11,
    C_MVD_G0 = 12,    /* abs_mvd_greater0_flag */
    C_MVD_G1 = 13,    /* abs_mvd_greater1_flag */
    C_MVP = 14,       /* mvp_l0_flag */
    C_ROOT_CBF = 15,  /* rqt_root_cbf */
    C_SPLIT_TF = 16,  /* split_transform_flag, 3 */
    C_CBF_LUMA = 19,  /* 2 */
    C_CBF_CHROMA = 21, /* 4 */
    C_LAST_X = 25,    /* last_sig_coeff_x_prefix, 18 */
    C_LAST_Y = 43,    /* last_sig_coeff_y_prefix, 18 */
    C_CSBF = 61,      /* coded_sub_block_flag, 4 */
    C_SIG = 65,       /* sig_coeff_flag, 42 (27 luma, 15 chroma) */
    C_G1 = 107,       /* coeff_abs_level_greater1_flag, 24 */
    C_G2 = 131,       /* coeff_abs_level_greater2_flag, 6 */
    NCTX = 137
};

/* initValue per context, [0] initType 0 (I slices), [1] initType 1 (P
 * slices, cabac_init_flag 0).  Elements an I slice never codes hold 154. */
static const uint8_t kInit[2][NCTX] = {
    {
        139, 141, 157,                /* split_cu_flag (Table 9-7) */
        154, 154, 154,                /* cu_skip_flag (P only) */
        154,                          /* pred_mode_flag (P only) */
        184,                          /* part_mode (Table 9-11) */
        184,                          /* prev_intra_luma_pred_flag (Table 9-12) */
        63,                           /* intra_chroma_pred_mode (Table 9-13) */
        154, 154, 154, 154, 154, 154, /* merge_flag, merge_idx, mvd x2, mvp_l0_flag, rqt_root_cbf (P only) */
        153, 138, 138,                /* split_transform_flag (Table 9-20) */
        111, 141,                     /* cbf_luma (Table 9-21) */
        94, 138, 182, 154,            /* cbf_cb / cbf_cr (Table 9-22) */
        110, 110, 124, 125, 140, 153, 125, 127, 140, 109, 111, 143, 127, 111, 79, 108, 123, 63, /* last x */
        110, 110, 124, 125, 140, 153, 125, 127, 140, 109, 111, 143, 127, 111, 79, 108, 123, 63, /* last y */
        91, 171, 134, 141,            /* coded_sub_block_flag (Table 9-28) */
        111, 111, 125, 110, 110, 94, 124, 108, 124, 107, 125, 141, 179, 153, 125, 107, 125, 141, 179, 153, 125,
        107, 125, 141, 179, 153, 125, 140, 139, 182, 182, 152, 136, 152, 136, 153, 136, 139, 111, 136, 139, 111,
        140, 92, 137, 138, 140, 152, 138, 139, 153, 74, 149, 92, 139, 107, 122, 152, 140, 179, 166, 182, 140, 227,
        122, 197,                     /* greater1 (Table 9-30) */
        138, 153, 136, 167, 152, 152, /* greater2 (Table 9-31) */
    },
    {
        107, 139, 126,                /* split_cu_flag */
        197, 185, 201,                /* cu_skip_flag (Table 9-8) */
        149,                          /* pred_mode_flag (Table 9-10) */
        154,                          /* part_mode */
        154,                          /* prev_intra_luma_pred_flag */
        152,                          /* intra_chroma_pred_mode */
        110,                          /* merge_flag (Table 9-15) */
        122,                          /* merge_idx (Table 9-16) */
        140,                          /* abs_mvd_greater0_flag (Table 9-24) */
        198,                          /* abs_mvd_greater1_flag (Table 9-25) */
        168,                          /* mvp_l0_flag (Table 9-19) */
        79,                           /* rqt_root_cbf (Table 9-14) */
        124, 138, 94,                 /* split_transform_flag */
        153, 111,                     /* cbf_luma */
        149, 107, 167, 154,           /* cbf_cb / cbf_cr */
        125, 110, 94, 110, 95, 79, 125, 111, 110, 78, 110, 111, 111, 95, 94, 108, 123, 108, /* last x */
        125, 110, 94, 110, 95, 79, 125, 111, 110, 78, 110, 111, 111, 95, 94, 108, 123, 108, /* last y */
        121, 140, 61, 154,            /* coded_sub_block_flag */
        155, 154, 139, 153, 139, 123, 123, 63, 153, 166, 183, 140, 136, 153, 154, 166, 183, 140, 136, 153, 154,
        166, 183, 140, 136, 153, 154, 170, 153, 123, 123, 107, 121, 107, 121, 167, 151, 183, 140, 151, 183, 140,
        154, 196, 196, 167, 154, 152, 167, 182, 182, 134, 149, 136, 153, 121, 136, 137, 169, 194, 166, 167, 154,
        167, 137, 182,                /* greater1 */
        107, 167, 91, 122, 107, 167,  /* greater2 */
    }};

/* The arithmetic coder writes whole bytes straight into the RBSP buffer
 * (the slice data starts byte aligned).  low keeps the 10-bit interval
 * base and, above it, the queue + 8 bits not yet written; the bit above
 * those is a carry into the last byte written.  A byte of ones is held
 * back (ff) until a later carry decides whether it stays 0xff or becomes
 * 0x00.  This writes the bits of the PutBit / outstanding-bit encoder of
 * H.264 9.3.4.2 (whose first bit, always zero, is not written), byte by
 * byte. */
typedef struct {
    bw_t *w;
    uint8_t *p, *start, *end;
    uint32_t low, range;
    int queue, ff;
    uint8_t st[NCTX]; /* (pStateIdx << 1) | valMps */
} cab_t;

static int clampi(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }

/* 9.3.2.2: the contexts' initial states for a slice QP, and the coder's
 * registers at the start of the slice data (which is byte aligned). */
static void cab_start(cab_t *c, bw_t *w, int init_type, int slice_qp)
{
    const int qp = clampi(slice_qp, 0, 51);
    for (int i = 0; i < NCTX; i++) {
        const int v = kInit[init_type][i];
        const int m = (v >> 4) * 5 - 45, n = ((v & 15) << 3) - 16;
        const int pre = clampi(((m * qp) >> 4) + n, 1, 126);
        const int mps = pre > 63;
        c->st[i] = (uint8_t)(((mps ? pre - 64 : 63 - pre) << 1) | mps);
    }
    c->w = w;
    c->start = c->p = w->buf + (w->n < w->cap ? w->n : w->cap);
    c->end = w->buf + w->cap;
    c->low = 0;
    c->range = 510;
    c->queue = -9;
    c->ff = 0;
}

static inline void cab_byte(cab_t *c)
{
    if (c->queue < 0) return;
    const uint32_t out = c->low >> (c->queue + 10);
    c->low &= (0x400u << c->queue) - 1;
    c->queue -= 8;
    if ((out & 0xff) == 0xff) {
        c->ff++;
        return;
    }
    if (c->p + c->ff + 1 >= c->end) {
        c->w->ovf = 1;
        return;
    }
    const uint32_t carry = out >> 8;
    if (c->p > c->start) c->p[-1] = (uint8_t)(c->p[-1] + carry);
    for (; c->ff > 0; c->ff--) *c->p++ = (uint8_t)(0xff + carry);
    *c->p++ = (uint8_t)out;
}

static inline void cab_bin(cab_t *c, int ctx, int bin)
{
    const int s = c->st[ctx];
    int p = s >> 1, mps = s & 1;
    const uint32_t lps = kLps[p][(c->range >> 6) & 3];
    c->range -= lps;
    if ((bin != 0) != mps) {
        c->low += c->range;
        c->range = lps;
        if (p == 0) mps ^= 1;
        p = kNextLps[p];
    } else if (p < 62) {
        p++;
    }
    c->st[ctx] = (uint8_t)((p << 1) | mps);
    if (c->range < 256) {
        const int sh = __builtin_clz(c->range) - 23;
        c->range <<= sh;
        c->low <<= sh;
        c->queue += sh;
        cab_byte(c);
    }
}

static inline void cab_bypass(cab_t *c, int bin)
{
    c->low <<= 1;
    if (bin) c->low += c->range;
    c->queue++;
    cab_byte(c);
}

/* n bypass bins, the bits of v from the most significant: up to eight at
 * a time, each group one multiply-add into low (the queue then holds at
 * most 7 + 8 bits below the carry, within 32). */
static void cab_bypass_bits(cab_t *c, uint32_t v, int n)
{
    while (n > 0) {
        const int k = n > 8 ? 8 : n;
        n -= k;
        c->low = (c->low << k) + c->range * ((v >> n) & ((1u << k) - 1));
        c->queue += k;
        cab_byte(c);
    }
}

/* end_of_slice_segment_flag.  With bin 1 the coder is flushed (H.264
 * 9.3.4.5: seven renormalisation shifts, then bits 9 and 8 of low and a
 * one, which is the slice data's rbsp_stop_one_bit) and the RBSP writer
 * takes over at the first bit after it. */
static void cab_terminate(cab_t *c, int bin)
{
    c->range -= 2;
    if (!bin) {
        if (c->range < 256) {
            c->range <<= 1;
            c->low <<= 1;
            c->queue++;
            cab_byte(c);
        }
        return;
    }
    c->low += c->range;
    c->low <<= 7;
    c->queue += 7;
    cab_byte(c);
    c->low = (c->low & ~0x7fu) | 0x80u;
    c->low <<= 3;
    c->queue += 3;
    cab_byte(c);
    const int nb = c->queue + 8; /* bits still held, 0..7 */
    const uint32_t carry = (c->low >> (10 + nb)) & 1;
    if (c->p + c->ff + 1 >= c->end) {
        c->w->ovf = 1;
        return;
    }
    if (c->p > c->start) c->p[-1] = (uint8_t)(c->p[-1] + carry);
    for (; c->ff > 0; c->ff--) *c->p++ = (uint8_t)(0xff + carry);
    bw_t *w = c->w;
    w->n = (size_t)(c->p - w->buf);
    w->acc = 0;
    w->nacc = 0;
    if (nb) bw_put(w, (c->low >> 10) & ((1u << nb) - 1), nb);
}

/* ------------------------------------------------------- residual coding */

/* Scan orders (6.5.3 up-right diagonal, 6.5.4 horizontal, 6.5.5 vertical)
 * of blocks of side 1, 2, 4 and 8: [scanIdx][log2 side][position] = x | y << 4. */
static uint8_t g_scan[3][4][64];
static int g_scan_ready;

static void make_scans(void)
{
    if (g_scan_ready) return;
    for (int l = 0; l < 4; l++) {
        const int n = 1 << l;
        int i = 0, x = 0, y = 0;
        while (i < n * n) {
            while (y >= 0) {
                if (x < n && y < n) g_scan[0][l][i++] = (uint8_t)(x | y << 4);
                y--;
                x++;
            }
            y = x;
            x = 0;
        }
        i = 0;
        for (y = 0; y < n; y++)
            for (x = 0; x < n; x++) g_scan[1][l][i++] = (uint8_t)(x | y << 4);
        i = 0;
        for (x = 0; x < n; x++)
            for (y = 0; y < n; y++) g_scan[2][l][i++] = (uint8_t)(x | y << 4);
    }
    g_scan_ready = 1;
}

/* last_sig_coeff_{x,y}_prefix of one coordinate: its bins (TR, cMax = 2
 * log2 - 1) and, returned through *sfx / *nsfx, the suffix (7.4.9.11). */
static void last_prefix(cab_t *c, int v, int log2, int cidx, int base, int *sfx, int *nsfx)
{
    int p;
    if (v < 4) {
        p = v;
        *nsfx = 0;
        *sfx = 0;
    } else {
        int g = 0;
        while ((v >> (g + 1)) != 0) g++;
        p = 2 * g + ((v >> (g - 1)) & 1);
        *nsfx = (p >> 1) - 1;
        *sfx = v - ((1 << ((p >> 1) - 1)) * (2 + (p & 1)));
    }
    const int cmax = (log2 << 1) - 1;
    const int off = cidx ? 15 : 3 * (log2 - 2) + ((log2 - 1) >> 2);
    const int shift = cidx ? log2 - 2 : (log2 + 1) >> 2;
    for (int b = 0; b < p; b++) cab_bin(c, base + off + (b >> shift), 1);
    if (p < cmax) cab_bin(c, base + off + (p >> shift), 0);
}

/* coeff_abs_level_remaining with Rice parameter k (9.3.3.11). */
static void cab_remaining(cab_t *c, uint32_t v, int k)
{
    if (v < (3u << k)) {
        const uint32_t q = v >> k; /* q ones, a zero, k bits */
        cab_bypass_bits(c, ((((1u << q) - 1) << 1) << k) | (v & ((1u << k) - 1)), (int)q + 1 + k);
        return;
    }
    uint32_t r = v - (3u << k);
    int len = k;
    while (r >= (1u << len)) {
        r -= 1u << len;
        len++;
    }
    const int ones = 3 + len - k;
    cab_bypass_bits(c, ((1u << ones) - 1) << 1, ones + 1);
    cab_bypass_bits(c, r, len);
}

/* 9.3.4.2.5: sig_coeff_flag's context. */
static int sig_ctx(int x, int y, int log2, int cidx, int scan, int prev_csbf)
{
    static const uint8_t map4[16] = {0, 1, 4, 5, 2, 3, 4, 5, 6, 6, 8, 8, 7, 7, 8, 8};
    int s;
    if (log2 == 2) {
        s = map4[(y << 2) + x];
    } else if (x + y == 0) {
        s = 0;
    } else {
        const int xp = x & 3, yp = y & 3;
        if (prev_csbf == 0) s = xp + yp == 0 ? 2 : xp + yp < 3 ? 1 : 0;
        else if (prev_csbf == 1) s = yp == 0 ? 2 : yp == 1 ? 1 : 0;
        else if (prev_csbf == 2) s = xp == 0 ? 2 : xp == 1 ? 1 : 0;
        else s = 2;
        if (cidx == 0) {
            if ((x >> 2) + (y >> 2) > 0) s += 3;
            s += log2 == 3 ? (scan == 0 ? 9 : 15) : 21;
        } else {
            s += log2 == 3 ? 9 : 12;
        }
    }
    return cidx ? 27 + s : s;
}

/* residual_coding() (7.3.8.11) of one transform block of side 1 << log2
 * with at least one non-zero level: co[y * (1 << log2) + x]. */
static void residual(cab_t *c, const int16_t *co, int log2, int cidx, int scan)
{
    const int n = 1 << log2, lsb = log2 - 2, sb = 1 << lsb, nsb = sb * sb;
    const uint8_t *S = g_scan[scan][lsb], *P = g_scan[scan][2];
    int last_i = -1, last_k = -1;
    for (int i = nsb - 1; i >= 0 && last_i < 0; i--) {
        const int xs = S[i] & 15, ys = S[i] >> 4;
        for (int k = 15; k >= 0; k--) {
            const int x = xs * 4 + (P[k] & 15), y = ys * 4 + (P[k] >> 4);
            if (co[y * n + x]) {
                last_i = i;
                last_k = k;
                break;
            }
        }
    }
    if (last_i < 0) return;
    int lx = (S[last_i] & 15) * 4 + (P[last_k] & 15), ly = (S[last_i] >> 4) * 4 + (P[last_k] >> 4);
    if (scan == 2) {  /* the vertical scan sends the coordinates swapped */
        const int t = lx;
        lx = ly;
        ly = t;
    }
    int sx, nsx, sy, nsy;
    last_prefix(c, lx, log2, cidx, C_LAST_X, &sx, &nsx);
    last_prefix(c, ly, log2, cidx, C_LAST_Y, &sy, &nsy);
    cab_bypass_bits(c, (uint32_t)sx, nsx);
    cab_bypass_bits(c, (uint32_t)sy, nsy);

    uint8_t csbf[64];
    memset(csbf, 0, sizeof csbf);
    int c1_carry = 1; /* greater1Ctx state carried between sub-blocks */
    for (int i = last_i; i >= 0; i--) {
        const int xs = S[i] & 15, ys = S[i] >> 4;
        int lv[16], any = 0;
        for (int k = 0; k < 16; k++) {
            lv[k] = co[(ys * 4 + (P[k] >> 4)) * n + xs * 4 + (P[k] & 15)];
            any |= lv[k] != 0;
        }
        const int right = xs < sb - 1 ? csbf[ys * sb + xs + 1] : 0;
        const int below = ys < sb - 1 ? csbf[(ys + 1) * sb + xs] : 0;
        int infer_dc = 0;
        if (i < last_i && i > 0) {
            cab_bin(c, C_CSBF + (right | below) + (cidx ? 2 : 0), any);
            csbf[ys * sb + xs] = (uint8_t)any;
            infer_dc = 1;
            if (!any) continue;
        } else {
            csbf[ys * sb + xs] = 1;
        }
        const int prev_csbf = right | (below << 1);
        int sig[16];
        memset(sig, 0, sizeof sig);
        const int start = i == last_i ? last_k - 1 : 15;
        if (i == last_i) sig[last_k] = 1;
        for (int k = start; k >= 0; k--) {
            if (k > 0 || !infer_dc) {
                const int x = xs * 4 + (P[k] & 15), y = ys * 4 + (P[k] >> 4);
                sig[k] = lv[k] != 0;
                cab_bin(c, C_SIG + sig_ctx(x, y, log2, cidx, scan, prev_csbf), sig[k]);
                if (sig[k]) infer_dc = 0;
            } else {
                sig[k] = 1; /* the sub-block's only non-zero level, inferred */
            }
        }
        int nsig = 0;
        for (int k = 0; k < 16; k++) nsig += sig[k];
        if (!nsig) continue;

        /* greater1 flags of the first eight, the greater2 flag of the first above one */
        int ctx_set = (i == 0 || cidx > 0) ? 0 : 2;
        if (c1_carry == 0) ctx_set++;
        int g1ctx = 1, ng1 = 0, first_g1 = -1;
        int g1[16], g2[16];
        memset(g1, 0, sizeof g1);
        memset(g2, 0, sizeof g2);
        for (int k = 15; k >= 0; k--) {
            if (!sig[k] || ng1 >= 8) continue;
            const int a = lv[k] < 0 ? -lv[k] : lv[k];
            g1[k] = a > 1;
            cab_bin(c, C_G1 + ctx_set * 4 + (g1ctx > 3 ? 3 : g1ctx) + (cidx ? 16 : 0), g1[k]);
            ng1++;
            if (g1ctx > 0) g1ctx = g1[k] ? 0 : g1ctx + 1;
            if (g1[k] && first_g1 < 0) first_g1 = k;
        }
        c1_carry = g1ctx;
        if (first_g1 >= 0) {
            const int a = lv[first_g1] < 0 ? -lv[first_g1] : lv[first_g1];
            g2[first_g1] = a > 2;
            cab_bin(c, C_G2 + ctx_set + (cidx ? 4 : 0), g2[first_g1]);
        }
        uint32_t signs = 0;
        for (int k = 15; k >= 0; k--)
            if (sig[k]) signs = signs << 1 | (lv[k] < 0);
        cab_bypass_bits(c, signs, nsig);
        int nseen = 0, rice = 0;
        for (int k = 15; k >= 0; k--) {
            if (!sig[k]) continue;
            const int a = lv[k] < 0 ? -lv[k] : lv[k];
            const int base = 1 + g1[k] + g2[k];
            const int thr = nseen < 8 ? (k == first_g1 ? 3 : 2) : 1;
            if (base == thr) {
                cab_remaining(c, (uint32_t)(a - base), rice);
                if (a > 3 * (1 << rice)) rice = rice + 1 > 4 ? 4 : rice + 1;
            }
            nseen++;
        }
    }
}

/* ------------------------------------------------------------- pictures */

typedef struct {
    const bc250_hevc_stream_cfg_t *cfg;
    int w, h, L, wctb, hctb, a16w;
    int *slice_of_row;   /* per CTU row */
    uint8_t *depth8;     /* CtDepth per 8x8 */
    uint8_t *mode4;      /* IntraPredModeY per 4x4 */
    int32_t *mv;         /* per 8x8: x, y (quarter samples) */
    uint8_t *skip8;      /* cu_skip_flag per 8x8 */
    uint8_t zs4[8][8];   /* MinTbAddrZs inside a CTB by (y, x) at 4x4 granularity */
    const int16_t *coef;
    const uint8_t *modes;
    const uint32_t *mv8;
    cab_t c;
} pic_t;

/* MinTbAddrZs order inside a CTB at 4x4 granularity (6.5.2): the bits of
 * the 4x4 column and row interleaved. */
static void make_zs(pic_t *p)
{
    for (int yi = 0; yi < 8; yi++)
        for (int xi = 0; xi < 8; xi++) {
            int r = 0;
            for (int b = 0; b < 3; b++) r |= ((xi >> b) & 1) << (2 * b) | ((yi >> b) & 1) << (2 * b + 1);
            p->zs4[yi][xi] = (uint8_t)r;
        }
}

/* 6.4.1: whether the block at (xn, yn) is available to the one at (xc, yc):
 * inside the picture, in the same slice, and earlier in decoding order. */
static inline int avail(const pic_t *p, int xc, int yc, int xn, int yn)
{
    if (xn < 0 || yn < 0 || xn >= p->w || yn >= p->h) return 0;
    const int L = p->L, rn = yn >> L, rc = yc >> L;
    const int cn = xn >> L, cc = xc >> L;
    if (rn == rc && cn == cc) {
        const int m = (1 << L) - 1;
        return p->zs4[(yn & m) >> 2][(xn & m) >> 2] < p->zs4[(yc & m) >> 2][(xc & m) >> 2];
    }
    if (p->slice_of_row[rn] != p->slice_of_row[rc]) return 0;
    return rn < rc || (rn == rc && cn < cc);
}

/* The staging offset of the 16x16 area holding (x, y). */
static size_t area_of(const pic_t *p, int x, int y) { return (size_t)(y >> 4) * p->a16w + (x >> 4); }

static void set_depth(pic_t *p, int x0, int y0, int log2, int depth)
{
    const int s = 1 << (log2 - 3), w8 = p->w >> 3;
    for (int j = 0; j < s && (y0 >> 3) + j < (p->h >> 3); j++)
        for (int i = 0; i < s && (x0 >> 3) + i < w8; i++) p->depth8[((y0 >> 3) + j) * w8 + (x0 >> 3) + i] = (uint8_t)depth;
}

/* split_cu_flag's context (9.3.4.2.2). */
static int split_ctx(const pic_t *p, int x0, int y0, int depth)
{
    const int w8 = p->w >> 3;
    const int l = avail(p, x0, y0, x0 - 1, y0) && p->depth8[(y0 >> 3) * w8 + ((x0 - 1) >> 3)] > depth;
    const int a = avail(p, x0, y0, x0, y0 - 1) && p->depth8[((y0 - 1) >> 3) * w8 + (x0 >> 3)] > depth;
    return l + a;
}

/* n a multiple of 16 */
static int any_nz(const int16_t *v, int n)
{
    uint64_t acc = 0;
    for (int i = 0; i < n; i += 16) {
        uint64_t w[4];
        memcpy(w, v + i, sizeof w);
        acc |= w[0] | w[1] | w[2] | w[3];
        if (acc) return 1;
    }
    return 0;
}

/* 8.4.2: the three most probable modes of the PU at (x, y). */
static void mpm_list(const pic_t *p, int x, int y, int cand[3])
{
    const int w4 = p->w >> 2;
    const int ca = avail(p, x, y, x - 1, y) ? p->mode4[(y >> 2) * w4 + ((x - 1) >> 2)] : 1;
    const int above_in_ctb = y - 1 >= ((y >> p->L) << p->L);
    const int cb = (above_in_ctb && avail(p, x, y, x, y - 1)) ? p->mode4[((y - 1) >> 2) * w4 + (x >> 2)] : 1;
    if (ca == cb) {
        if (ca < 2) {
            cand[0] = 0;
            cand[1] = 1;
            cand[2] = 26;
        } else {
            cand[0] = ca;
            cand[1] = 2 + ((ca + 29) % 32);
            cand[2] = 2 + ((ca - 2 + 1) % 32);
        }
    } else {
        cand[0] = ca;
        cand[1] = cb;
        cand[2] = (ca != 0 && cb != 0) ? 0 : (ca != 1 && cb != 1) ? 1 : 26;
    }
}

/* 7.4.9.11: the scan of an intra 4x4 luma block by its mode. */
static int scan_of_mode(int mode) { return (mode >= 6 && mode <= 14) ? 2 : (mode >= 22 && mode <= 30) ? 1 : 0; }

/* An I picture's 8x8 CU: PART_NxN, four 4x4 PUs, chroma DC. */
static void intra_cu(pic_t *p, int x0, int y0, int is_p_slice)
{
    cab_t *c = &p->c;
    const size_t t = area_of(p, x0, y0);
    const int z = ((y0 >> 3) & 1) * 2 + ((x0 >> 3) & 1);
    const int16_t *co = p->coef + 384 * t;
    const uint8_t *md = p->modes + 16 * t + 4 * z;
    const int w4 = p->w >> 2;
    if (is_p_slice) {
        cab_bin(c, C_SKIP, 0);
        cab_bin(c, C_PRED_MODE, 1);
    }
    cab_bin(c, C_PART_MODE, 0); /* PART_NxN */
    int cand[4][3], idx[4];
    for (int u = 0; u < 4; u++) {
        const int x = x0 + 4 * (u & 1), y = y0 + 4 * (u >> 1);
        mpm_list(p, x, y, cand[u]);
        idx[u] = -1;
        for (int i = 0; i < 3; i++)
            if (md[u] == cand[u][i]) {
                idx[u] = i;
                break;
            }
        cab_bin(c, C_PREV_INTRA, idx[u] >= 0);
        p->mode4[(y >> 2) * w4 + (x >> 2)] = md[u]; /* the next PU's neighbour */
    }
    for (int u = 0; u < 4; u++) {
        if (idx[u] >= 0) {
            cab_bypass(c, idx[u] > 0);
            if (idx[u] > 0) cab_bypass(c, idx[u] > 1);
        } else {
            int s[3] = {cand[u][0], cand[u][1], cand[u][2]};
            for (int a = 0; a < 2; a++)
                for (int b = a + 1; b < 3; b++)
                    if (s[b] < s[a]) {
                        const int tmp = s[a];
                        s[a] = s[b];
                        s[b] = tmp;
                    }
            int rem = md[u];
            for (int i = 2; i >= 0; i--)
                if (rem > s[i]) rem--;
            cab_bypass_bits(c, (uint32_t)rem, 5);
        }
    }
    /* chroma DC: 4 (the luma mode) when PU 0 is DC, else 3 */
    if (md[0] == 1) {
        cab_bin(c, C_CHROMA, 0);
    } else {
        cab_bin(c, C_CHROMA, 1);
        cab_bypass_bits(c, 3, 2);
    }
    const int16_t *cbk = co + 256 + 16 * z, *crk = co + 320 + 16 * z;
    const int cbf_cb = any_nz(cbk, 16), cbf_cr = any_nz(crk, 16);
    cab_bin(c, C_CBF_CHROMA + 0, cbf_cb);
    cab_bin(c, C_CBF_CHROMA + 0, cbf_cr);
    for (int u = 0; u < 4; u++) {
        const int16_t *lb = co + 16 * (4 * z + u);
        const int cbf = any_nz(lb, 16);
        cab_bin(c, C_CBF_LUMA + 0, cbf); /* trafoDepth 1 */
        if (cbf) residual(c, lb, 2, 0, scan_of_mode(md[u]));
        if (u == 3) {
            if (cbf_cb) residual(c, cbk, 2, 1, 0);
            if (cbf_cr) residual(c, crk, 2, 2, 0);
        }
    }
    set_depth(p, x0, y0, 3, p->L - 3);
}

/* mvd_coding() (7.3.8.9). */
static void mvd_code(cab_t *c, int dx, int dy)
{
    const int ax = dx < 0 ? -dx : dx, ay = dy < 0 ? -dy : dy;
    cab_bin(c, C_MVD_G0, ax > 0);
    cab_bin(c, C_MVD_G0, ay > 0);
    if (ax > 0) cab_bin(c, C_MVD_G1, ax > 1);
    if (ay > 0) cab_bin(c, C_MVD_G1, ay > 1);
    for (int comp = 0; comp < 2; comp++) {
        const int a = comp ? ay : ax, neg = comp ? dy < 0 : dx < 0;
        if (!a) continue;
        if (a > 1) { /* abs_mvd_minus2: EG1 */
            uint32_t v = (uint32_t)(a - 2);
            int k = 1;
            while (v >= (1u << k)) {
                cab_bypass(c, 1);
                v -= 1u << k;
                k++;
            }
            cab_bypass(c, 0);
            cab_bypass_bits(c, v, k);
        }
        cab_bypass(c, neg);
    }
}

static int mvd_cost(int d)
{
    const int a = d < 0 ? -d : d;
    if (!a) return 1;
    int bits = 3;
    if (a > 1) {
        uint32_t v = (uint32_t)(a - 2);
        int k = 1;
        while (v >= (1u << k)) {
            v -= 1u << k;
            k++;
            bits++;
        }
        bits += 1 + k;
    }
    return bits;
}

/* 8.5.3.2.6 / 8.5.3.2.7: the two AMVP candidates of the 2Nx2N PU at
 * (x, y), side s, from the spatial neighbours (one reference picture, every
 * CU inter; no temporal candidate). */
static void amvp(const pic_t *p, int x, int y, int s, int out[2][2])
{
    const int w8 = p->w >> 3;
#define MVAT(X, Y) (p->mv + 2 * ((size_t)((Y) >> 3) * w8 + ((X) >> 3)))
    const int a0 = avail(p, x, y, x - 1, y + s), a1 = avail(p, x, y, x - 1, y + s - 1);
    int fa = 0, fb = 0, ma[2] = {0, 0}, mb[2] = {0, 0};
    if (a0) {
        fa = 1;
        memcpy(ma, MVAT(x - 1, y + s), sizeof ma);
    } else if (a1) {
        fa = 1;
        memcpy(ma, MVAT(x - 1, y + s - 1), sizeof ma);
    }
    const int bx[3] = {x + s, x + s - 1, x - 1};
    for (int k = 0; k < 3 && !fb; k++)
        if (avail(p, x, y, bx[k], y - 1)) {
            fb = 1;
            memcpy(mb, MVAT(bx[k], y - 1), sizeof mb);
        }
#undef MVAT
    if (!(a0 || a1) && fb) { /* isScaledFlag 0: A takes B (8.5.3.2.7 step 7) */
        fa = 1;
        ma[0] = mb[0];
        ma[1] = mb[1];
    }
    int n = 0;
    if (fa) {
        out[n][0] = ma[0];
        out[n][1] = ma[1];
        n++;
    }
    if (fb && !(fa && ma[0] == mb[0] && ma[1] == mb[1])) {
        out[n][0] = mb[0];
        out[n][1] = mb[1];
        n++;
    }
    while (n < 2) {
        out[n][0] = out[n][1] = 0;
        n++;
    }
}

/* 8.5.3.2.2 - 8.5.3.2.5: the five merge candidates of the 2Nx2N PU at
 * (x, y), side s, in a P slice with one reference picture, every CU inter
 * and no temporal candidate: the spatial ones A1, B1, B0, A0, B2 with the
 * standard's pruning, then zero vectors. */
static void merge_list(const pic_t *p, int x, int y, int s, int out[5][2])
{
    const int w8 = p->w >> 3;
    const int nx[5] = {x - 1, x + s - 1, x + s, x - 1, x - 1}; /* A1 B1 B0 A0 B2 */
    const int ny[5] = {y + s - 1, y - 1, y - 1, y + s, y - 1};
    int av[5], m[5][2];
    for (int k = 0; k < 5; k++) {
        av[k] = avail(p, x, y, nx[k], ny[k]);
        if (av[k]) {
            const int32_t *v = p->mv + 2 * ((size_t)(ny[k] >> 3) * w8 + (nx[k] >> 3));
            m[k][0] = v[0];
            m[k][1] = v[1];
        }
    }
#define SAME(a, b) (av[a] && m[a][0] == m[b][0] && m[a][1] == m[b][1])
    int fl[5];
    fl[0] = av[0];
    fl[1] = av[1] && !SAME(0, 1);
    fl[2] = av[2] && !SAME(1, 2);
    fl[3] = av[3] && !SAME(0, 3);
    fl[4] = av[4] && !SAME(0, 4) && !SAME(1, 4) && fl[0] + fl[1] + fl[2] + fl[3] != 4;
#undef SAME
    int n = 0;
    for (int k = 0; k < 5; k++)
        if (fl[k]) {
            out[n][0] = m[k][0];
            out[n][1] = m[k][1];
            n++;
        }
    for (; n < 5; n++) out[n][0] = out[n][1] = 0;
}

/* A P picture's inter CU of side 1 << log2 at (x0, y0), 2Nx2N with one
 * luma TU of the CU's size and one chroma TU per component of half it.
 * Its vector is sent as a merge index when it is one of the merge
 * candidates (a skip when it has no levels either), else by AMVP: the
 * same prediction and residual either way, so the same picture. */
static void inter_cu(pic_t *p, int x0, int y0, int log2)
{
    cab_t *c = &p->c;
    const int s = 1 << log2, w8 = p->w >> 3;
    const int32_t *mv = p->mv + 2 * ((size_t)(y0 >> 3) * w8 + (x0 >> 3));

    /* the levels: a CU8 in its area's z-slot; a CU16 the area; a CU32 the four areas' chunks */
    int16_t lbuf[1024], cbbuf[256], crbuf[256];
    const int nl = s * s, nc = nl / 4;
    if (log2 == 3) {
        const int16_t *co = p->coef + 384 * area_of(p, x0, y0);
        const int z = ((y0 >> 3) & 1) * 2 + ((x0 >> 3) & 1);
        memcpy(lbuf, co + 64 * z, 64 * sizeof(int16_t));
        memcpy(cbbuf, co + 256 + 16 * z, 16 * sizeof(int16_t));
        memcpy(crbuf, co + 320 + 16 * z, 16 * sizeof(int16_t));
    } else {
        const int nq = log2 == 4 ? 1 : 4;
        for (int q = 0; q < nq; q++) {
            const int16_t *co = p->coef + 384 * area_of(p, x0 + 16 * (q & 1), y0 + 16 * (q >> 1));
            memcpy(lbuf + 256 * q, co, 256 * sizeof(int16_t));
            memcpy(cbbuf + 64 * q, co + 256, 64 * sizeof(int16_t));
            memcpy(crbuf + 64 * q, co + 320, 64 * sizeof(int16_t));
        }
    }
    const int cbf_y = any_nz(lbuf, nl), cbf_cb = any_nz(cbbuf, nc), cbf_cr = any_nz(crbuf, nc);
    const int root = cbf_y | cbf_cb | cbf_cr;

    int mc[5][2], midx = -1;
    merge_list(p, x0, y0, s, mc);
    for (int k = 0; k < 5 && midx < 0; k++)
        if (mc[k][0] == mv[0] && mc[k][1] == mv[1]) midx = k;
    /* cu_skip_flag's context: the left and above CUs' skip flags (9.3.4.2.2) */
    const int sl = avail(p, x0, y0, x0 - 1, y0) && p->skip8[(y0 >> 3) * w8 + ((x0 - 1) >> 3)];
    const int sa = avail(p, x0, y0, x0, y0 - 1) && p->skip8[((y0 - 1) >> 3) * w8 + (x0 >> 3)];
    const int skip = midx >= 0 && !root;
    cab_bin(c, C_SKIP + sl + sa, skip);
    for (int j = 0; j < (s >> 3); j++)
        for (int i = 0; i < (s >> 3); i++) p->skip8[((y0 >> 3) + j) * w8 + (x0 >> 3) + i] = (uint8_t)skip;
    if (!skip) {
        cab_bin(c, C_PRED_MODE, 0); /* MODE_INTER */
        cab_bin(c, C_PART_MODE, 1); /* PART_2Nx2N */
        cab_bin(c, C_MERGE_FLAG, midx >= 0);
    }
    if (midx >= 0) { /* merge_idx: TR, cMax 4, first bin in context */
        cab_bin(c, C_MERGE_IDX, midx > 0);
        for (int k = 1; k < 4 && k <= midx; k++) cab_bypass(c, midx > k);
    } else {
        int cand[2][2];
        amvp(p, x0, y0, s, cand);
        const int c0 = mvd_cost(mv[0] - cand[0][0]) + mvd_cost(mv[1] - cand[0][1]);
        const int c1 = mvd_cost(mv[0] - cand[1][0]) + mvd_cost(mv[1] - cand[1][1]);
        const int idx = c1 < c0;
        mvd_code(c, mv[0] - cand[idx][0], mv[1] - cand[idx][1]);
        cab_bin(c, C_MVP, idx);
        cab_bin(c, C_ROOT_CBF, root); /* not sent for a 2Nx2N merge: there it is 1 */
    }
    if (root) {
        cab_bin(c, C_CBF_CHROMA + 0, cbf_cb);
        cab_bin(c, C_CBF_CHROMA + 0, cbf_cr);
        if (cbf_cb || cbf_cr) cab_bin(c, C_CBF_LUMA + 1, cbf_y); /* trafoDepth 0 */
        if (cbf_y) residual(c, lbuf, log2, 0, 0);
        if (cbf_cb) residual(c, cbbuf, log2 - 1, 1, 0);
        if (cbf_cr) residual(c, crbuf, log2 - 1, 2, 0);
    }
    set_depth(p, x0, y0, log2, p->L - log2);
}

/* The P picture's CU size at (x, y): CTU16 codes 8x8 CUs; at CTU32 the
 * staging's modes hold log2 of the CU covering each 8x8. */
static int cu_log2_at(const pic_t *p, int x, int y)
{
    if (p->L == 4) return 3;
    const int z = ((y >> 3) & 1) * 2 + ((x >> 3) & 1);
    const int v = p->modes[16 * area_of(p, x, y) + z];
    return v >= 3 && v <= 5 ? v : 3;
}

/* coding_quadtree() (7.3.8.4). */
static void quadtree(pic_t *p, int x0, int y0, int log2, int depth, int is_idr)
{
    const int s = 1 << log2;
    int split;
    if (log2 > 3) {
        if (is_idr) split = 1;
        else split = cu_log2_at(p, x0, y0) < log2;
        if (x0 + s <= p->w && y0 + s <= p->h) cab_bin(&p->c, C_SPLIT_CU + split_ctx(p, x0, y0, depth), split);
        else split = 1;
    } else {
        split = 0;
    }
    if (split) {
        const int h = s >> 1;
        for (int q = 0; q < 4; q++) {
            const int x1 = x0 + h * (q & 1), y1 = y0 + h * (q >> 1);
            if (x1 < p->w && y1 < p->h) quadtree(p, x1, y1, log2 - 1, depth + 1, is_idr);
        }
        return;
    }
    if (is_idr) intra_cu(p, x0, y0, 0);
    else inter_cu(p, x0, y0, log2);
}

static int cfg_ok(const bc250_hevc_stream_cfg_t *cfg)
{
    if (!cfg || cfg->width <= 0 || cfg->height <= 0 || (cfg->width & 15) || (cfg->height & 15)) return 0;
    if (cfg->bit_depth != 8 && cfg->bit_depth != 10) return 0;
    if (cfg->ctu_log2 != 4 && cfg->ctu_log2 != 5) return 0;
    const int rows = (cfg->height + (1 << cfg->ctu_log2) - 1) >> cfg->ctu_log2;
    if (cfg->num_slices < 1 || cfg->num_slices > rows) return 0;
    if (cfg->init_qp < 0 || cfg->init_qp > 51) return 0;
    if (cfg->crop_width < 0 || cfg->crop_height < 0 || cfg->crop_width > cfg->width ||
        cfg->crop_height > cfg->height || (cfg->crop_width & 1) || (cfg->crop_height & 1))
        return 0;
    return 1;
}

/* profile_tier_level(1, 0) (7.3.3): Main or Main 10, Main tier. */
static void write_ptl(bw_t *w, int bd, int level_idc)
{
    const int profile = bd > 8 ? 2 : 1;
    bw_bits(w, 0, 2);        /* general_profile_space */
    bw_bits(w, 0, 1);        /* general_tier_flag */
    bw_bits(w, (uint32_t)profile, 5);
    for (int j = 0; j < 32; j++) bw_bits(w, j == profile || (profile == 1 && j == 2), 1);
    bw_bits(w, 1, 1);        /* progressive_source */
    bw_bits(w, 0, 1);        /* interlaced_source */
    bw_bits(w, 0, 1);        /* non_packed_constraint */
    bw_bits(w, 1, 1);        /* frame_only_constraint */
    bw_bits(w, 0, 32);       /* general_reserved_zero_43bits */
    bw_bits(w, 0, 11);
    bw_bits(w, 0, 1);        /* general_inbld_flag / reserved */
    bw_bits(w, (uint32_t)level_idc, 8);
}

/* The lowest level whose MaxLumaPs holds the picture (Table A.8). */
static int level_of(int w, int h)
{
    static const struct { uint32_t ps; int idc; } t[] = {
        {36864, 30}, {122880, 60}, {245760, 63}, {552960, 90}, {983040, 93},
        {2228224, 120}, {8912896, 150}, {35651584, 180}};
    const uint32_t ps = (uint32_t)w * (uint32_t)h;
    for (size_t i = 0; i < sizeof t / sizeof t[0]; i++)
        if (ps <= t[i].ps) return t[i].idc;
    return 186;
}

size_t bc250_hevc_write_parameter_sets(const bc250_hevc_stream_cfg_t *cfg, uint8_t *out, size_t cap)
{
    if (!cfg_ok(cfg) || !out) return 0;
    uint8_t rb[256];
    size_t at = 0, k;
    const int bd = cfg->bit_depth;
    const int cw = cfg->crop_width ? cfg->crop_width : cfg->width, ch = cfg->crop_height ? cfg->crop_height : cfg->height;

    /* VPS (7.3.2.1) */
    bw_t w = {rb, sizeof rb, 0, 0, 0, 0};
    bw_bits(&w, 0, 4);       /* vps_video_parameter_set_id */
    bw_bits(&w, 3, 2);       /* vps_base_layer_internal_flag, vps_base_layer_available_flag */
    bw_bits(&w, 0, 6);       /* vps_max_layers_minus1 */
    bw_bits(&w, 0, 3);       /* vps_max_sub_layers_minus1 */
    bw_bits(&w, 1, 1);       /* vps_temporal_id_nesting_flag */
    bw_bits(&w, 0xffff, 16); /* vps_reserved_0xffff_16bits */
    write_ptl(&w, bd, level_of(cfg->width, cfg->height));
    bw_bits(&w, 1, 1);       /* vps_sub_layer_ordering_info_present_flag */
    bw_ue(&w, 1);            /* vps_max_dec_pic_buffering_minus1 */
    bw_ue(&w, 0);            /* vps_max_num_reorder_pics */
    bw_ue(&w, 0);            /* vps_max_latency_increase_plus1 */
    bw_bits(&w, 0, 6);       /* vps_max_layer_id */
    bw_ue(&w, 0);            /* vps_num_layer_sets_minus1 */
    bw_bits(&w, 0, 1);       /* vps_timing_info_present_flag */
    bw_bits(&w, 0, 1);       /* vps_extension_flag */
    bw_trailing(&w);
    if (w.ovf || !(k = nal_out(out, cap, at, 32, rb, w.n))) return 0;
    at += k;

    /* SPS (7.3.2.2) */
    w = (bw_t){rb, sizeof rb, 0, 0, 0, 0};
    bw_bits(&w, 0, 4);       /* sps_video_parameter_set_id */
    bw_bits(&w, 0, 3);       /* sps_max_sub_layers_minus1 */
    bw_bits(&w, 1, 1);       /* sps_temporal_id_nesting_flag */
    write_ptl(&w, bd, level_of(cfg->width, cfg->height));
    bw_ue(&w, 0);            /* sps_seq_parameter_set_id */
    bw_ue(&w, 1);            /* chroma_format_idc: 4:2:0 */
    bw_ue(&w, (uint32_t)cfg->width);
    bw_ue(&w, (uint32_t)cfg->height);
    const int crop = cw != cfg->width || ch != cfg->height;
    bw_bits(&w, (uint32_t)crop, 1); /* conformance_window_flag */
    if (crop) {              /* offsets in chroma samples (SubWidthC = SubHeightC = 2) */
        bw_ue(&w, 0);
        bw_ue(&w, (uint32_t)(cfg->width - cw) / 2);
        bw_ue(&w, 0);
        bw_ue(&w, (uint32_t)(cfg->height - ch) / 2);
    }
    bw_ue(&w, (uint32_t)(bd - 8)); /* bit_depth_luma_minus8 */
    bw_ue(&w, (uint32_t)(bd - 8)); /* bit_depth_chroma_minus8 */
    bw_ue(&w, 4);            /* log2_max_pic_order_cnt_lsb_minus4: 8-bit POC LSB */
    bw_bits(&w, 1, 1);       /* sps_sub_layer_ordering_info_present_flag */
    bw_ue(&w, 1);            /* sps_max_dec_pic_buffering_minus1 */
    bw_ue(&w, 0);            /* sps_max_num_reorder_pics */
    bw_ue(&w, 0);            /* sps_max_latency_increase_plus1 */
    bw_ue(&w, 0);            /* log2_min_luma_coding_block_size_minus3: 8x8 CUs */
    bw_ue(&w, (uint32_t)(cfg->ctu_log2 - 3)); /* log2_diff_max_min_luma_coding_block_size */
    bw_ue(&w, 0);            /* log2_min_luma_transform_block_size_minus2: 4x4 */
    bw_ue(&w, (uint32_t)(cfg->ctu_log2 == 4 ? 1 : 3)); /* largest TU: 8x8 (CTU16) or 32x32 */
    bw_ue(&w, 0);            /* max_transform_hierarchy_depth_inter: one TU per inter CU */
    bw_ue(&w, 0);            /* max_transform_hierarchy_depth_intra (NxN splits once by itself) */
    bw_bits(&w, 0, 1);       /* scaling_list_enabled_flag */
    bw_bits(&w, 0, 1);       /* amp_enabled_flag */
    bw_bits(&w, 0, 1);       /* sample_adaptive_offset_enabled_flag */
    bw_bits(&w, 0, 1);       /* pcm_enabled_flag */
    bw_ue(&w, 1);            /* num_short_term_ref_pic_sets */
    bw_ue(&w, 1);            /* st_ref_pic_set(0): num_negative_pics */
    bw_ue(&w, 0);            /*   num_positive_pics */
    bw_ue(&w, 0);            /*   delta_poc_s0_minus1: the picture before */
    bw_bits(&w, 1, 1);       /*   used_by_curr_pic_s0_flag */
    bw_bits(&w, 0, 1);       /* long_term_ref_pics_present_flag */
    bw_bits(&w, 0, 1);       /* sps_temporal_mvp_enabled_flag */
    bw_bits(&w, 0, 1);       /* strong_intra_smoothing_enabled_flag */
    bw_bits(&w, 0, 1);       /* vui_parameters_present_flag */
    bw_bits(&w, 0, 1);       /* sps_extension_present_flag */
    bw_trailing(&w);
    if (w.ovf || !(k = nal_out(out, cap, at, 33, rb, w.n))) return 0;
    at += k;

    /* PPS (7.3.2.3) */
    w = (bw_t){rb, sizeof rb, 0, 0, 0, 0};
    bw_ue(&w, 0);            /* pps_pic_parameter_set_id */
    bw_ue(&w, 0);            /* pps_seq_parameter_set_id */
    bw_bits(&w, 0, 1);       /* dependent_slice_segments_enabled_flag */
    bw_bits(&w, 0, 1);       /* output_flag_present_flag */
    bw_bits(&w, 0, 3);       /* num_extra_slice_header_bits */
    bw_bits(&w, 0, 1);       /* sign_data_hiding_enabled_flag */
    bw_bits(&w, 0, 1);       /* cabac_init_present_flag */
    bw_ue(&w, 0);            /* num_ref_idx_l0_default_active_minus1 */
    bw_ue(&w, 0);            /* num_ref_idx_l1_default_active_minus1 */
    bw_se(&w, cfg->init_qp - 26);
    bw_bits(&w, 0, 1);       /* constrained_intra_pred_flag */
    bw_bits(&w, 0, 1);       /* transform_skip_enabled_flag */
    bw_bits(&w, 0, 1);       /* cu_qp_delta_enabled_flag */
    bw_se(&w, 0);            /* pps_cb_qp_offset */
    bw_se(&w, 0);            /* pps_cr_qp_offset */
    bw_bits(&w, 0, 1);       /* pps_slice_chroma_qp_offsets_present_flag */
    bw_bits(&w, 0, 1);       /* weighted_pred_flag */
    bw_bits(&w, 0, 1);       /* weighted_bipred_flag */
    bw_bits(&w, 0, 1);       /* transquant_bypass_enabled_flag */
    bw_bits(&w, 0, 1);       /* tiles_enabled_flag */
    bw_bits(&w, 0, 1);       /* entropy_coding_sync_enabled_flag */
    bw_bits(&w, 0, 1);       /* pps_loop_filter_across_slices_enabled_flag */
    bw_bits(&w, 1, 1);       /* deblocking_filter_control_present_flag */
    bw_bits(&w, 0, 1);       /*   deblocking_filter_override_enabled_flag */
    bw_bits(&w, (uint32_t)!cfg->deblock, 1); /* pps_deblocking_filter_disabled_flag */
    if (cfg->deblock) {
        bw_se(&w, 0);        /*   pps_beta_offset_div2 */
        bw_se(&w, 0);        /*   pps_tc_offset_div2 */
    }
    bw_bits(&w, 0, 1);       /* pps_scaling_list_data_present_flag */
    bw_bits(&w, 0, 1);       /* lists_modification_present_flag */
    bw_ue(&w, 0);            /* log2_parallel_merge_level_minus2 */
    bw_bits(&w, 0, 1);       /* slice_segment_header_extension_present_flag */
    bw_bits(&w, 0, 1);       /* pps_extension_present_flag */
    bw_trailing(&w);
    if (w.ovf || !(k = nal_out(out, cap, at, 34, rb, w.n))) return 0;
    at += k;
    return at;
}

size_t bc250_hevc_write_picture(const bc250_hevc_stream_cfg_t *cfg, int is_idr, int poc, int qp,
                                const int16_t *coef, const uint8_t *modes, const uint32_t *mv8, uint8_t *out,
                                size_t cap)
{
    if (!cfg_ok(cfg) || !out || !coef || !modes || (!is_idr && !mv8) || qp < 0 || qp > 51 || poc < 0) return 0;
    make_scans();
    pic_t p;
    memset(&p, 0, sizeof p);
    p.cfg = cfg;
    p.w = cfg->width;
    p.h = cfg->height;
    p.L = cfg->ctu_log2;
    p.wctb = (p.w + (1 << p.L) - 1) >> p.L;
    p.hctb = (p.h + (1 << p.L) - 1) >> p.L;
    p.a16w = p.w >> 4;
    make_zs(&p);
    p.coef = coef;
    p.modes = modes;
    p.mv8 = mv8;
    const int w8 = p.w >> 3, h8 = p.h >> 3;
    p.slice_of_row = (int *)malloc(sizeof(int) * (size_t)p.hctb);
    p.depth8 = (uint8_t *)calloc((size_t)w8 * h8, 1);
    p.mode4 = (uint8_t *)malloc((size_t)(p.w >> 2) * (p.h >> 2));
    p.mv = (int32_t *)calloc((size_t)w8 * h8 * 2, sizeof(int32_t));
    p.skip8 = (uint8_t *)calloc((size_t)w8 * h8, 1);
    /* the slice data of one slice is at most a few bits per sample */
    const size_t rcap = (size_t)p.w * p.h * 4 + 4096;
    uint8_t *rb = (uint8_t *)malloc(rcap);
    size_t at = 0;
    int ok = p.slice_of_row && p.depth8 && p.mode4 && p.mv && p.skip8 && rb;
    if (ok) {
        memset(p.mode4, 1, (size_t)(p.w >> 2) * (p.h >> 2));
        const int ns = cfg->num_slices;
        for (int s = 0; s < ns; s++) {
            const int r0 = (int)((int64_t)p.hctb * s / ns), r1 = (int)((int64_t)p.hctb * (s + 1) / ns);
            for (int r = r0; r < r1; r++) p.slice_of_row[r] = s;
        }
        if (!is_idr) { /* the vector of every 8x8 CU, in picture raster order */
            for (int y = 0; y < h8; y++)
                for (int x = 0; x < w8; x++) {
                    const size_t cu = 4 * ((size_t)(y >> 1) * p.a16w + (x >> 1)) + (size_t)((y & 1) * 2 + (x & 1));
                    p.mv[2 * ((size_t)y * w8 + x)] = (int32_t)mv8[4 * cu];
                    p.mv[2 * ((size_t)y * w8 + x) + 1] = (int32_t)mv8[4 * cu + 1];
                }
        }
        int addr_bits = 0;
        while ((1 << addr_bits) < p.wctb * p.hctb) addr_bits++;
        for (int s = 0; s < ns && ok; s++) {
            const int r0 = (int)((int64_t)p.hctb * s / ns), r1 = (int)((int64_t)p.hctb * (s + 1) / ns);
            bw_t w = {rb, rcap, 0, 0, 0, 0};
            /* slice_segment_header() (7.3.6.1) */
            bw_bits(&w, s == 0, 1);              /* first_slice_segment_in_pic_flag */
            if (is_idr) bw_bits(&w, 0, 1);       /* no_output_of_prior_pics_flag */
            bw_ue(&w, 0);                        /* slice_pic_parameter_set_id */
            if (s) bw_bits(&w, (uint32_t)(r0 * p.wctb), addr_bits); /* slice_segment_address */
            bw_ue(&w, is_idr ? 2 : 1);           /* slice_type: I or P */
            if (!is_idr) {
                bw_bits(&w, (uint32_t)poc & 255, 8); /* slice_pic_order_cnt_lsb */
                bw_bits(&w, 1, 1);               /* short_term_ref_pic_set_sps_flag */
                bw_bits(&w, 0, 1);               /* num_ref_idx_active_override_flag */
                bw_ue(&w, 0);                    /* five_minus_max_num_merge_cand */
            }
            bw_se(&w, qp - cfg->init_qp);        /* slice_qp_delta */
            bw_trailing(&w);                     /* byte_alignment() */
            /* slice_segment_data() */
            cab_start(&p.c, &w, is_idr ? 0 : 1, qp);
            for (int r = r0; r < r1; r++)
                for (int cx = 0; cx < p.wctb; cx++) {
                    quadtree(&p, cx << p.L, r << p.L, p.L, 0, is_idr);
                    cab_terminate(&p.c, r == r1 - 1 && cx == p.wctb - 1);
                }
            bw_align_zero(&w);                   /* rbsp_slice_segment_trailing_bits */
            size_t k = 0;
            if (w.ovf || !(k = nal_out(out, cap, at, is_idr ? 19 : 1, rb, w.n))) ok = 0;
            at += k;
        }
    }
    free(p.slice_of_row);
    free(p.depth8);
    free(p.mode4);
    free(p.mv);
    free(p.skip8);
    free(rb);
    return ok ? at : 0;
}
