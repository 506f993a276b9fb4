// h264_cavlc.hip - H.264 CAVLC slice data of a whole frame on gfx950.
//
// Replaces the host loop of src/encoder_h264.c:3031-3240 (slice header,
// I16x16 / P16x16 / P_Skip macroblocks, rbsp trailing bits) with the residual
// writers of src/cavlc.c.  Everything a macroblock writes is a function of
// staging data that is complete once the frame is coded (levels, DC terms,
// modes, vectors) and of its neighbours' TotalCoeff, so the frame is written
// macroblock-parallel in five launches:
//   1. classify : per 4x4 block TotalCoeff, per MB cbp / skip / mvd / modes
//   2. size     : bits of every MB (nC from the neighbours' counts)
//   3. scan     : one workgroup; skip runs, MB bit offsets, slice headers'
//                 lengths and the slices' byte extents, back to back
//   4. fill     : zeroes exactly the bytes the slices occupy
//   5. write    : every lane ORs its own syntax elements into place
// A wave holds two macroblocks, 32 lanes each, in bitstream order: lane 0 the
// MB header (with mb_skip_run), 1 the Intra16x16 DC block, 2..17 the luma
// blocks in blkIdx order, 18/19 chroma DC Cb/Cr, 20..27 chroma AC.  Each
// residual block is coded by one lane with fully unrolled loops, so the
// coefficients stay in registers.
#include "bc250_dev.h"
#include "bc250_bitstream.h"

#include <algorithm>
#include <mutex>
#include <stdio.h>

namespace {

constexpr int kLanesPerMb = 32;
constexpr int kTcStride = 32;  // TotalCoeff bytes per MB, indexed by lane

// ------------------------------------------------------------- code tables
// H.264 Tables 9-5 (coeff_token), 9-7 / 9-8 (total_zeros), 9-9a (chroma DC
// total_zeros) and 9-10 (run_before), each entry (length << 8 | code).
// coeff_token rows: [class][TotalCoeff][TrailingOnes], class 0..2 for
// 0 <= nC < 2, 2 <= nC < 4, 4 <= nC < 8, class 3 for chroma DC (nC -1).
__constant__ uint16_t kCoeffToken[4][17][4] = {
    {{0x101, 0, 0, 0},           {0x605, 0x201, 0, 0},       {0x807, 0x604, 0x301, 0},
     {0x907, 0x806, 0x705, 0x503}, {0xa07, 0x906, 0x805, 0x603}, {0xb07, 0xa06, 0x905, 0x704},
     {0xd0f, 0xb06, 0xa05, 0x804}, {0xd0b, 0xd0e, 0xb05, 0x904}, {0xd08, 0xd0a, 0xd0d, 0xa04},
     {0xe0f, 0xe0e, 0xd09, 0xb04}, {0xe0b, 0xe0a, 0xe0d, 0xd0c}, {0xf0f, 0xf0e, 0xe09, 0xe0c},
     {0xf0b, 0xf0a, 0xf0d, 0xe08}, {0x100f, 0xf01, 0xf09, 0xf0c}, {0x100b, 0x100e, 0x100d, 0xf08},
     {0x1007, 0x100a, 0x1009, 0x100c}, {0x1004, 0x1006, 0x1005, 0x1008}},
    {{0x203, 0, 0, 0},           {0x60b, 0x202, 0, 0},       {0x607, 0x507, 0x303, 0},
     {0x707, 0x60a, 0x609, 0x405}, {0x807, 0x606, 0x605, 0x404}, {0x804, 0x706, 0x705, 0x506},
     {0x907, 0x806, 0x805, 0x608}, {0xb0f, 0x906, 0x905, 0x604}, {0xb0b, 0xb0e, 0xb0d, 0x704},
     {0xc0f, 0xb0a, 0xb09, 0x904}, {0xc0b, 0xc0e, 0xc0d, 0xb0c}, {0xc08, 0xc0a, 0xc09, 0xb08},
     {0xd0f, 0xd0e, 0xd0d, 0xc0c}, {0xd0b, 0xd0a, 0xd09, 0xd0c}, {0xd07, 0xe0b, 0xd06, 0xd08},
     {0xe09, 0xe08, 0xe0a, 0xd01}, {0xe07, 0xe06, 0xe05, 0xe04}},
    {{0x40f, 0, 0, 0},           {0x60f, 0x40e, 0, 0},       {0x60b, 0x50f, 0x40d, 0},
     {0x608, 0x50c, 0x50e, 0x40c}, {0x70f, 0x50a, 0x50b, 0x40b}, {0x70b, 0x508, 0x509, 0x40a},
     {0x709, 0x60e, 0x60d, 0x409}, {0x708, 0x60a, 0x609, 0x408}, {0x80f, 0x70e, 0x70d, 0x50d},
     {0x80b, 0x80e, 0x70a, 0x60c}, {0x90f, 0x80a, 0x80d, 0x70c}, {0x90b, 0x90e, 0x809, 0x80c},
     {0x908, 0x90a, 0x90d, 0x808}, {0xa0d, 0x907, 0x909, 0x90c}, {0xa09, 0xa0c, 0xa0b, 0xa0a},
     {0xa05, 0xa08, 0xa07, 0xa06}, {0xa01, 0xa04, 0xa03, 0xa02}},
    {{0x201, 0, 0, 0}, {0x607, 0x101, 0, 0}, {0x604, 0x606, 0x301, 0}, {0x603, 0x703, 0x702, 0x605},
     {0x602, 0x803, 0x802, 0x700}},
};
__constant__ uint16_t kTotalZeros[15][16] = {
    {0x101, 0x303, 0x302, 0x403, 0x402, 0x503, 0x502, 0x603, 0x602, 0x703, 0x702, 0x803, 0x802, 0x903, 0x902,
     0x901},
    {0x307, 0x306, 0x305, 0x304, 0x303, 0x405, 0x404, 0x403, 0x402, 0x503, 0x502, 0x603, 0x602, 0x601, 0x600},
    {0x405, 0x307, 0x306, 0x305, 0x404, 0x403, 0x304, 0x303, 0x402, 0x503, 0x502, 0x601, 0x501, 0x600},
    {0x503, 0x307, 0x405, 0x404, 0x306, 0x305, 0x304, 0x403, 0x303, 0x402, 0x502, 0x501, 0x500},
    {0x405, 0x404, 0x403, 0x307, 0x306, 0x305, 0x304, 0x303, 0x402, 0x501, 0x401, 0x500},
    {0x601, 0x501, 0x307, 0x306, 0x305, 0x304, 0x303, 0x302, 0x401, 0x301, 0x600},
    {0x601, 0x501, 0x305, 0x304, 0x303, 0x203, 0x302, 0x401, 0x301, 0x600},
    {0x601, 0x401, 0x501, 0x303, 0x203, 0x202, 0x302, 0x301, 0x600},
    {0x601, 0x600, 0x401, 0x203, 0x202, 0x301, 0x201, 0x501},
    {0x501, 0x500, 0x301, 0x203, 0x202, 0x201, 0x401},
    {0x400, 0x401, 0x301, 0x302, 0x101, 0x303},
    {0x400, 0x401, 0x201, 0x101, 0x301},
    {0x300, 0x301, 0x101, 0x201},
    {0x200, 0x201, 0x101},
    {0x100, 0x101},
};
__constant__ uint16_t kTotalZerosCdc[3][4] = {
    {0x101, 0x201, 0x301, 0x300}, {0x101, 0x201, 0x200, 0}, {0x101, 0x100, 0, 0}};
__constant__ uint16_t kRunBefore[7][15] = {
    {0x101, 0x100},
    {0x101, 0x201, 0x200},
    {0x203, 0x202, 0x201, 0x200},
    {0x203, 0x202, 0x201, 0x301, 0x300},
    {0x203, 0x202, 0x303, 0x302, 0x301, 0x300},
    {0x203, 0x300, 0x301, 0x303, 0x302, 0x305, 0x304},
    {0x307, 0x306, 0x305, 0x304, 0x303, 0x302, 0x301, 0x401, 0x501, 0x601, 0x701, 0x801, 0x901, 0xa01, 0xb01},
};
// Table 9-4, Inter column, as coded_block_pattern -> codeNum
__constant__ uint8_t kInterCbpCode[48] = {
    0,  2,  3,  7,  4,  8,  17, 13, 5,  18, 9,  14, 10, 15, 16, 11, 1,  32, 33, 36, 34, 37, 44, 40,
    35, 45, 38, 41, 39, 42, 43, 19, 6,  24, 25, 20, 26, 21, 46, 28, 27, 47, 22, 29, 23, 30, 31, 12};

constexpr int kZigzag[16] = {0, 1, 4, 8, 5, 2, 3, 6, 9, 12, 13, 10, 7, 11, 14, 15};

// ------------------------------------------------------------- bit sinks
struct CountSink {
    uint32_t bits = 0;
    __device__ void put(int len, uint32_t) { bits += (uint32_t)len; }
};

// MSB-first writer at an arbitrary bit position of a zeroed word buffer;
// words shared with a neighbour's bits are combined with atomicOr.
struct WriteSink {
    uint32_t *out;
    uint32_t cap_words, word;
    uint64_t acc;  // n pending bits, right-aligned
    int n;
    uint32_t *overflow;
    __device__ WriteSink(uint32_t *o, uint32_t cap, uint32_t bitpos, uint32_t *ovf)
        : out(o), cap_words(cap), word(bitpos >> 5), acc(0), n((int)(bitpos & 31)), overflow(ovf) {}
    __device__ __forceinline__ void emit(uint32_t w)
    {
        if (w) {
            if (word < cap_words) atomicOr(out + word, __builtin_bswap32(w));
            else atomicOr(overflow, 1u);
        }
        word++;
    }
    __device__ __forceinline__ void put(int len, uint32_t v)
    {
        if (len <= 0) return;
        acc = acc << len | (uint64_t)(len == 32 ? v : v & ((1u << len) - 1));
        n += len;
        if (n >= 32) {
            n -= 32;
            emit((uint32_t)(acc >> n));
            acc &= (1ull << n) - 1;
        }
    }
    __device__ void finish()
    {
        if (n > 0) emit((uint32_t)(acc << (32 - n)));
    }
};

template <class S> __device__ __forceinline__ void put_code(S &s, uint16_t e) { s.put(e >> 8, e & 0xff); }

template <class S> __device__ __forceinline__ void put_ue(S &s, uint32_t v)
{
    const uint32_t cn = v + 1;
    const int k = 31 - __builtin_clz(cn);
    s.put(k, 0);
    s.put(k + 1, cn);
}
template <class S> __device__ __forceinline__ void put_se(S &s, int v)
{
    put_ue(s, v <= 0 ? (uint32_t)(-2 * v) : (uint32_t)(2 * v - 1));
}

__device__ __forceinline__ uint32_t ue_bits(uint32_t v) { return 2u * (31u - __builtin_clz(v + 1)) + 1u; }

// One level of 9.2.2.1 in the encoding direction: level_prefix zeros, then
// the 1 and level_suffix as one piece; returns the next suffixLength.
template <class S> __device__ __forceinline__ int put_level(S &s, int level, bool first_lt3, int sl)
{
    int code = level > 0 ? 2 * level - 2 : -2 * level - 1;
    if (first_lt3) code -= 2;
    int z, sz;
    uint32_t suf;
    if (sl == 0 && code < 14) {
        z = code, sz = 0, suf = 0;
    } else if (sl == 0 && code < 30) {
        z = 14, sz = 4, suf = (uint32_t)(code - 14);
    } else if (sl > 0 && (code >> sl) < 15) {
        z = code >> sl, sz = sl, suf = (uint32_t)code & ((1u << sl) - 1);
    } else {
        // level_prefix >= 15: a suffix of level_prefix - 3 bits carries the rest
        const uint32_t rest = (uint32_t)(code - (sl == 0 ? 30 : (15 << sl))) + 4096u;
        sz = 31 - __builtin_clz(rest);
        z = sz + 3;
        suf = rest - (1u << sz);
    }
    s.put(z, 0);
    s.put(sz + 1, 1u << sz | suf);
    const int a = level < 0 ? -level : level;
    if (sl == 0) sl = 1;
    if (a > (3 << (sl - 1)) && sl < 6) sl++;
    return sl;
}

// residual_block_cavlc (7.3.5.3.2) of n coefficients c[0..n) in scan order
// (c[n..16) zero); nc -1 selects the chroma DC tables.
template <class S> __device__ __forceinline__ void code_block(S &s, const int (&c)[16], int n, int nc)
{
    uint32_t mask = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) mask |= (uint32_t)(c[i] != 0) << i;
    const int tc = __builtin_popcount(mask);
    int t1 = 0, seen = 0;
    uint32_t signs = 0;
    bool run1 = true;
#pragma unroll
    for (int i = 15; i >= 0; i--)
        if (c[i] != 0) {
            if (run1 && seen < 3 && (c[i] == 1 || c[i] == -1)) {
                t1++;
                signs = signs << 1 | (uint32_t)(c[i] < 0);
            } else {
                run1 = false;
            }
            seen++;
        }
    if (nc < 0) put_code(s, kCoeffToken[3][tc][t1]);
    else if (nc >= 8) s.put(6, tc == 0 ? 3u : (uint32_t)((tc - 1) << 2 | t1));
    else put_code(s, kCoeffToken[nc < 2 ? 0 : nc < 4 ? 1 : 2][tc][t1]);
    if (tc == 0) return;
    s.put(t1, signs);
    int sl = tc > 10 && t1 < 3 ? 1 : 0, j = 0;
#pragma unroll
    for (int i = 15; i >= 0; i--)
        if (c[i] != 0) {
            if (j >= t1) sl = put_level(s, c[i], j == t1 && t1 < 3, sl);
            j++;
        }
    const int hi = 31 - __builtin_clz(mask), tz = hi + 1 - tc;
    if (tc < n) put_code(s, nc < 0 ? kTotalZerosCdc[tc - 1][tz] : kTotalZeros[tc - 1][tz]);
    int left = tz, prev = -1;
#pragma unroll
    for (int i = 15; i >= 0; i--)
        if (c[i] != 0) {
            if (prev >= 0 && left > 0) {
                const int run = prev - i - 1;
                put_code(s, kRunBefore[(left < 7 ? left : 7) - 1][run]);
                left -= run;
            }
            prev = i;
        }
}

// ------------------------------------------------------------- geometry
struct Geo {
    int wmb, total, ns;
    // slice of mb: the largest s with floor(s * total / ns) <= mb
    __device__ int slice_of(int mb) const { return (int)(((int64_t)(mb + 1) * ns - 1) / total); }
    __device__ int start_of(int s) const { return (int)((int64_t)s * total / ns); }
};

__device__ __forceinline__ int bx4(int k) { return (k & 1) | (k & 4) >> 1; }
__device__ __forceinline__ int by4(int k) { return (k & 2) >> 1 | (k & 8) >> 2; }
__device__ __forceinline__ int blk_at(int x, int y) { return (x & 1) | (y & 1) << 1 | (x & 2) << 1 | (y & 2) << 2; }

__device__ __forceinline__ int med3(int a, int b, int c) { return max(min(a, b), min(max(a, b), c)); }

// neighbour (dx, dy) of mb inside the picture and the slice starting at `start`
__device__ __forceinline__ bool nb_mv(const gpu_mv_t *mv, int wmb, int start, int mb, int dx, int dy, int &x, int &y)
{
    const int nx = mb % wmb + dx, ny = mb / wmb + dy;
    x = y = 0;
    if (nx < 0 || ny < 0 || nx >= wmb) return false;
    const int n = ny * wmb + nx;
    if (n < start) return false;
    x = mv[n].mvx;
    y = mv[n].mvy;
    return true;
}

// The reference's 16x16 predictor (src/encoder_h264.c:974-1003): C or D,
// A alone when B and C are missing, else the median with missing ones zero.
__device__ __forceinline__ void mv_pred(const gpu_mv_t *mv, int wmb, int start, int mb, int &px, int &py)
{
    int ax, ay, bx, by, cx, cy;
    const bool a = nb_mv(mv, wmb, start, mb, -1, 0, ax, ay);
    const bool b = nb_mv(mv, wmb, start, mb, 0, -1, bx, by);
    bool c = nb_mv(mv, wmb, start, mb, 1, -1, cx, cy);
    if (!c) c = nb_mv(mv, wmb, start, mb, -1, -1, cx, cy);
    if (!b && !c && a) {
        px = ax;
        py = ay;
        return;
    }
    px = med3(ax, bx, cx);
    py = med3(ay, by, cy);
}

// P_Skip vector of 8.4.1.1 (src/encoder_h264.c:1036-1058)
__device__ __forceinline__ void skip_mv(const gpu_mv_t *mv, int wmb, int start, int mb, int &px, int &py)
{
    int ax, ay, bx, by;
    const bool a = nb_mv(mv, wmb, start, mb, -1, 0, ax, ay);
    const bool b = nb_mv(mv, wmb, start, mb, 0, -1, bx, by);
    if (!a || !b || (ax == 0 && ay == 0) || (bx == 0 && by == 0)) {
        px = py = 0;
        return;
    }
    mv_pred(mv, wmb, start, mb, px, py);
}

// mode sanitising of src/encoder_h264.c:858-887
__device__ __forceinline__ int fix_mode(int m, bool top, bool left, int dcm, int hm, int vm, int pm)
{
    if (!top && !left) return dcm;
    if (!top && (m == vm || m == pm)) return left ? hm : dcm;
    if (!left && (m == hm || m == pm)) return top ? vm : dcm;
    return m;
}

// 16 levels of a 4x4 block, raster order
__device__ __forceinline__ void load_block(const int16_t *p, int (&r)[16])
{
    const int4 a = *(const int4 *)p, b = *(const int4 *)(p + 8);
    const int w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int i = 0; i < 8; i++) {
        r[2 * i] = (int16_t)(w[i] & 0xffff);
        r[2 * i + 1] = w[i] >> 16;
    }
}

__device__ __forceinline__ uint32_t nz_mask16(const int (&r)[16])
{
    uint32_t m = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) m |= (uint32_t)(r[i] != 0) << i;
    return m;
}

// chroma DC of component comp: the 2x2 Hadamard of the raw DC terms and
// its levels at the chroma QP
__device__ __forceinline__ void chroma_dc(const int32_t *dc, int mb, int comp, int qp, int (&lv)[4], bool &raw_nz)
{
    const int32_t *d = dc + (size_t)mb * 24 + 16 + comp * 4;
    const int c0 = d[0], c1 = d[1], c2 = d[2], c3 = d[3];
    const int h[4] = {c0 + c1 + c2 + c3, c0 - c1 + c2 - c3, c0 + c1 - c2 - c3, c0 - c1 - c2 + c3};
    const int qpc = chroma_qp(qp), per = qpc / 6, rem = qpc % 6;
    raw_nz = false;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        lv[i] = quant_dc_k5(h[i], per, rem);
        raw_nz |= h[i] != 0;
    }
}

// Intra16x16 luma DC levels in the order CAVLC scans them: Hadamard of the
// raw DC terms (halved, rounded), DC quantiser, transposed
// (src/encoder_h264.c:775-787)
__device__ __forceinline__ void luma_dc_levels(const int32_t *dc, int mb, int qp, int (&out)[16])
{
    const int32_t *d = dc + (size_t)mb * 24;
    int t[16];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int s01 = d[4 * i] + d[4 * i + 1], d01 = d[4 * i] - d[4 * i + 1];
        const int s23 = d[4 * i + 2] + d[4 * i + 3], d23 = d[4 * i + 2] - d[4 * i + 3];
        t[0 * 4 + i] = s01 + s23;
        t[1 * 4 + i] = s01 - s23;
        t[2 * 4 + i] = d01 - d23;
        t[3 * 4 + i] = d01 + d23;
    }
    const int per = qp / 6, rem = qp % 6;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int s01 = t[4 * i] + t[4 * i + 1], d01 = t[4 * i] - t[4 * i + 1];
        const int s23 = t[4 * i + 2] + t[4 * i + 3], d23 = t[4 * i + 2] - t[4 * i + 3];
        const int v[4] = {(s01 + s23 + 1) >> 1, (s01 - s23 + 1) >> 1, (d01 - d23 + 1) >> 1, (d01 + d23 + 1) >> 1};
#pragma unroll
        for (int k = 0; k < 4; k++) out[k * 4 + i] = quant_dc_k5(v[k], per, rem);  // natural [4i+k], transposed
    }
}

// MB info word: bit 0 coded, 1..4 luma cbp, 5..6 chroma cbp, 7..8 16x16
// mode, 9..10 chroma mode
constexpr uint32_t kCoded = 1u;

struct Work {
    uint8_t *tc;     // [nmb][32] TotalCoeff by lane (0 where not coded)
    uint32_t *info;  // [nmb]
    int2 *mvd;       // [nmb]
    uint32_t *bits;  // [nmb] MB bits without mb_skip_run
    uint32_t *off;   // [nmb] absolute bit offset of the MB's first bit
    uint32_t *run;   // [nmb] mb_skip_run written before the MB
    uint16_t *lane_off;  // [nmb][32] each lane's first bit relative to the MB's code (after the run)
};

__host__ __device__ inline size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

__host__ __device__ inline Work carve(void *base, int nmb)
{
    uint8_t *p = (uint8_t *)base;
    Work w;
    w.tc = p;
    p += align256((size_t)nmb * kTcStride);
    w.info = (uint32_t *)p;
    p += align256((size_t)nmb * 4);
    w.mvd = (int2 *)p;
    p += align256((size_t)nmb * 8);
    w.bits = (uint32_t *)p;
    p += align256((size_t)nmb * 4);
    w.off = (uint32_t *)p;
    p += align256((size_t)nmb * 4);
    w.run = (uint32_t *)p;
    p += align256((size_t)nmb * 4);
    w.lane_off = (uint16_t *)p;
    return w;
}

__host__ inline size_t work_bytes(int nmb)
{
    return align256((size_t)nmb * kTcStride) + 4 * align256((size_t)nmb * 4) + align256((size_t)nmb * 8) +
           align256((size_t)nmb * kLanesPerMb * 2);
}

// the 32-lane half of the wave's ballot that belongs to this lane's MB
__device__ __forceinline__ uint32_t mb_ballot(bool p)
{
    const uint64_t b = __ballot(p);
    return (uint32_t)(b >> (threadIdx.x & 32));
}

// ------------------------------------------------------------- 1. classify
__global__ __launch_bounds__(256) void cavlc_classify_kernel(const int16_t *__restrict__ levels,
                                                             const int32_t *__restrict__ dc,
                                                             const uint32_t *__restrict__ modes,
                                                             const gpu_mv_t *__restrict__ mvs,
                                                             bc250_h264_slice_cfg_t cfg, Work w)
{
    const int lane = threadIdx.x & 31;
    const int mb = blockIdx.x * 8 + (threadIdx.x >> 5);
    const Geo g{cfg.width_in_mbs, cfg.width_in_mbs * cfg.height_in_mbs, cfg.num_slices};
    const bool valid = mb < g.total;
    const int m = valid ? mb : g.total - 1;  // tail lanes mirror the last MB and store nothing
    const bool idr = cfg.is_idr != 0;
    // per-lane block statistics
    uint32_t msk = 0;
    bool cdc_nz = false, raw_nz = false;
    if (lane >= 2 && lane < 18) {
        const int k = lane - 2;
        int r[16];
        load_block(levels + ((size_t)m * 24 + by4(k) * 4 + bx4(k)) * 16, r);
        msk = nz_mask16(r);
    } else if (lane >= 20 && lane < 28) {
        int r[16];
        load_block(levels + ((size_t)m * 24 + 16 + (lane - 20)) * 16, r);
        msk = nz_mask16(r) & 0xfffeu;  // chroma AC
    } else if (lane == 18 || lane == 19) {
        int lv[4];
        chroma_dc(dc, m, lane - 18, cfg.qp, lv, raw_nz);
        cdc_nz = lv[0] | lv[1] | lv[2] | lv[3];
    }
    const uint32_t luma_any = mb_ballot(msk != 0) >> 2 & 0xffffu;              // bit k: luma block k
    const uint32_t luma_ac = mb_ballot((msk & 0xfffeu) != 0) >> 2 & 0xffffu;
    const bool cac = (mb_ballot(msk != 0) >> 20 & 0xffu) != 0;
    const bool cdc = (mb_ballot(cdc_nz) >> 18 & 3u) != 0;
    const bool raw = (mb_ballot(raw_nz) >> 18 & 3u) != 0;
    const int cbp_c = cac ? 2 : cdc ? 1 : 0;
    const int s = g.slice_of(m), start = g.start_of(s), mbx = m % g.wmb;
    uint32_t info = 0;
    int cbp_l = 0;
    if (idr) {
        const bool left = mbx > 0 && m - 1 >= start, top = m - g.wmb >= start;
        const int pm = fix_mode((int)(modes[m] & 3), top, left, 2, 1, 0, 3);
        const int cm = fix_mode((int)(modes[m] >> 2 & 3), top, left, 0, 1, 2, 3);
        cbp_l = luma_ac ? 15 : 0;
        info = kCoded | (uint32_t)cbp_l << 1 | (uint32_t)cbp_c << 5 | (uint32_t)pm << 7 | (uint32_t)cm << 9;
    } else {
#pragma unroll
        for (int q = 0; q < 4; q++) cbp_l |= ((luma_any >> (4 * q)) & 0xfu) ? 1 << q : 0;
        bool zc = !(raw || cac);
        if (cfg.relax_chroma_skip && luma_any == 0 && !zc && iabs(dc[(size_t)m * 2]) <= 1 &&
            iabs(dc[(size_t)m * 2 + 1]) <= 1)
            zc = true;
        int sx, sy;
        skip_mv(mvs, g.wmb, start, m, sx, sy);
        const bool skip = luma_any == 0 && zc && mvs[m].mvx == sx && mvs[m].mvy == sy;
        if (!skip) info = kCoded | (uint32_t)cbp_l << 1 | (uint32_t)cbp_c << 5;
        if (valid && lane == 0) {
            int px = 0, py = 0;
            if (!skip) mv_pred(mvs, g.wmb, start, m, px, py);
            w.mvd[m] = make_int2(mvs[m].mvx - px, mvs[m].mvy - py);
        }
    }
    // TotalCoeff a neighbour sees: 0 for blocks that are not coded
    int tc = 0;
    if (info & kCoded) {
        if (lane >= 2 && lane < 18) {
            const int k = lane - 2;
            if (idr) tc = cbp_l ? __builtin_popcount(msk & 0xfffeu) : 0;
            else tc = (cbp_l >> (k >> 2) & 1) ? __builtin_popcount(msk) : 0;
        } else if (lane >= 20 && lane < 28) {
            tc = cbp_c == 2 ? __builtin_popcount(msk) : 0;
        }
    }
    if (valid) {
        w.tc[(size_t)m * kTcStride + lane] = (uint8_t)tc;
        if (lane == 0) w.info[m] = info;
    }
}

// nC of the block a lane codes (9.2.1 with the slice as the bound)
__device__ __forceinline__ int lane_nc(const Work &w, const Geo &g, int mb, int lane)
{
    const int s = g.slice_of(mb), start = g.start_of(s), mbx = mb % g.wmb;
    const bool has_l = mbx > 0 && mb - 1 >= start, has_t = mb - g.wmb >= start;
    const uint8_t *own = w.tc + (size_t)mb * kTcStride;
    int na = -1, nb = -1;
    if (lane <= 17) {  // luma (lane 1, the DC block, uses blkIdx 0)
        const int k = lane <= 1 ? 0 : lane - 2, x = bx4(k), y = by4(k);
        if (x > 0) na = own[2 + blk_at(x - 1, y)];
        else if (has_l) na = own[2 + blk_at(3, y) - kTcStride];
        if (y > 0) nb = own[2 + blk_at(x, y - 1)];
        else if (has_t) nb = w.tc[(size_t)(mb - g.wmb) * kTcStride + 2 + blk_at(x, 3)];
    } else {  // chroma AC, lanes 20..27
        const int base = lane < 24 ? 20 : 24, k = lane - base;
        if (k & 1) na = own[lane - 1];
        else if (has_l) na = own[lane + 1 - kTcStride];
        if (k & 2) nb = own[lane - 2];
        else if (has_t) nb = w.tc[(size_t)(mb - g.wmb) * kTcStride + lane + 2];
    }
    if (na >= 0 && nb >= 0) return (na + nb + 1) >> 1;
    return na >= 0 ? na : nb >= 0 ? nb : 0;
}

// everything lane `lane` of macroblock mb writes (the run not included)
template <class S>
__device__ __forceinline__ void code_lane(S &s, const int16_t *__restrict__ levels, const int32_t *__restrict__ dc,
                          const bc250_h264_slice_cfg_t &cfg, const Work &w, const Geo &g, int mb, int lane)
{
    const uint32_t info = w.info[mb];
    if (!(info & kCoded)) return;
    const bool idr = cfg.is_idr != 0;
    const int cbp_l = info >> 1 & 15, cbp_c = info >> 5 & 3;
    if (lane == 0) {
        if (idr) {
            put_ue(s, 1u + (info >> 7 & 3) + 4u * cbp_c + 12u * (cbp_l != 0));
            put_ue(s, info >> 9 & 3);
            put_se(s, 0);
        } else {
            const int2 d = w.mvd[mb];
            const int cbp = cbp_l | cbp_c << 4;
            put_ue(s, 0);
            put_se(s, d.x);
            put_se(s, d.y);
            put_ue(s, kInterCbpCode[cbp]);
            if (cbp) put_se(s, 0);
        }
        return;
    }
    int c[16];
    int n = 16, nc = 0;
    if (lane == 1) {
        if (!idr) return;
        int dl[16];
        luma_dc_levels(dc, mb, cfg.qp, dl);
#pragma unroll
        for (int i = 0; i < 16; i++) c[i] = dl[kZigzag[i]];
        nc = lane_nc(w, g, mb, lane);
    } else if (lane < 18) {
        const int k = lane - 2;
        if (!(cbp_l >> (k >> 2) & 1)) return;
        int r[16];
        load_block(levels + ((size_t)mb * 24 + by4(k) * 4 + bx4(k)) * 16, r);
#pragma unroll
        for (int i = 0; i < 16; i++) c[i] = idr ? (i < 15 ? r[kZigzag[i + 1]] : 0) : r[kZigzag[i]];
        n = idr ? 15 : 16;
        nc = lane_nc(w, g, mb, lane);
    } else if (lane < 20) {
        if (cbp_c == 0) return;
        int lv[4];
        bool raw;
        chroma_dc(dc, mb, lane - 18, cfg.qp, lv, raw);
#pragma unroll
        for (int i = 0; i < 16; i++) c[i] = i < 4 ? lv[i] : 0;
        n = 4;
        nc = -1;
    } else if (lane < 28) {
        if (cbp_c != 2) return;
        int r[16];
        load_block(levels + ((size_t)mb * 24 + 16 + (lane - 20)) * 16, r);
#pragma unroll
        for (int i = 0; i < 16; i++) c[i] = i < 15 ? r[kZigzag[i + 1]] : 0;
        n = 15;
        nc = lane_nc(w, g, mb, lane);
    } else {
        return;
    }
    code_block(s, c, n, nc);
}

// ------------------------------------------------------------- 2. size
__global__ __launch_bounds__(256) void cavlc_size_kernel(const int16_t *__restrict__ levels,
                                                         const int32_t *__restrict__ dc,
                                                         bc250_h264_slice_cfg_t cfg, Work w)
{
    const int lane = threadIdx.x & 31;
    const int mb = blockIdx.x * 8 + (threadIdx.x >> 5);
    const Geo g{cfg.width_in_mbs, cfg.width_in_mbs * cfg.height_in_mbs, cfg.num_slices};
    const int m = mb < g.total ? mb : g.total - 1;
    CountSink s;
    code_lane(s, levels, dc, cfg, w, g, m, lane);
    // inclusive prefix over the MB's 32 lanes (a macroblock codes < 2^16 bits)
    uint32_t x = s.bits;
#pragma unroll
    for (int d = 1; d < 32; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 32);
        if (lane >= d) x += y;
    }
    if (mb < g.total) {
        w.lane_off[(size_t)mb * kLanesPerMb + lane] = (uint16_t)(x - s.bits);
        if (lane == 31) w.bits[mb] = x;
    }
}

// ------------------------------------------------------------- 3. scan
// the slice header (7.3.3) as the reference writes it; the caller sums or
// emits it
template <class S> __device__ __forceinline__ void slice_header(S &s, const bc250_h264_slice_cfg_t &cfg, int first_mb)
{
    const bool idr = cfg.is_idr != 0;
    put_ue(s, (uint32_t)first_mb);
    put_ue(s, idr ? 2u : 0u);
    put_ue(s, (uint32_t)cfg.pps_id);
    s.put(cfg.log2_max_frame_num_minus4 + 4, (uint32_t)cfg.frame_num);
    if (idr) put_ue(s, (uint32_t)cfg.idr_pic_id);
    s.put(cfg.log2_max_poc_lsb_minus4 + 4, (uint32_t)cfg.poc);
    s.put(idr ? 2 : 3, 0);
    put_se(s, cfg.qp - 26 - cfg.pic_init_qp_minus26);
    put_ue(s, (uint32_t)cfg.deblock_idc);
    if (cfg.deblock_idc != 1) {
        put_se(s, 0);
        put_se(s, 0);
    }
}

constexpr int kScanThreads = 1024;

// exclusive scan over the workgroup (sum, or max when MAX); `total` gets
// the reduction over all threads
template <bool MAX> __device__ uint32_t block_scan(uint32_t v, uint32_t *lds, uint32_t &total)
{
    constexpr int kWaves = kScanThreads / 64;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x = MAX ? max(x, y) : x + y;
    }
    if (lane == 63) lds[wv] = x;
    __syncthreads();
    if (wv == 0) {
        const uint32_t e = lane < kWaves ? lds[lane] : 0u;
        uint32_t z = e;
#pragma unroll
        for (int d = 1; d < kWaves; d <<= 1) {
            const uint32_t y = __shfl_up(z, d, 64);
            if (lane >= d) z = MAX ? max(z, y) : z + y;
        }
        const uint32_t before = __shfl_up(z, 1, 64);
        if (lane < kWaves) lds[kWaves + lane] = lane ? before : 0u;
        if (lane == kWaves - 1) lds[2 * kWaves] = z;
    }
    __syncthreads();
    const uint32_t base = lds[kWaves + wv];
    const uint32_t incl = MAX ? max(base, x) : base + x;
    const uint32_t before = __shfl_up(incl, 1, 64);
    total = lds[2 * kWaves];
    __syncthreads();  // lds is reused by the next scan
    return MAX ? (lane ? before : base) : incl - v;
}

// info: [4 s + 0] byte offset, [+1] byte length, [+2] header bits,
// [+3] trailing mb_skip_run (0 when the last MB is coded); [4 ns] total
// bytes, [4 ns + 1] overflow flag
__global__ __launch_bounds__(kScanThreads) void cavlc_scan_kernel(bc250_h264_slice_cfg_t cfg, Work w,
                                                                  uint32_t cap_bytes, uint32_t *info)
{
    __shared__ uint32_t lds[2 * (kScanThreads / 64) + 1];
    const Geo g{cfg.width_in_mbs, cfg.width_in_mbs * cfg.height_in_mbs, cfg.num_slices};
    const int t = threadIdx.x;
    const bool idr = cfg.is_idr != 0;
    uint32_t byte_base = 0;
    for (int s = 0; s < g.ns; s++) {
        const int a = g.start_of(s), b = g.start_of(s + 1);
        const int per = (b - a + kScanThreads - 1) / kScanThreads;
        const int lo = min(a + t * per, b), hi = min(lo + per, b);
        // 1 + the last coded MB of each chunk (0: none), scanned with max so
        // that every thread knows the coded MB before its chunk
        uint32_t last = 0;
        for (int i = lo; i < hi; i++)
            if (w.info[i] & kCoded) last = (uint32_t)(i + 1);
        uint32_t last_all;
        const uint32_t prev = block_scan<true>(last, lds, last_all);
        // mb_skip_run before each coded MB and the chunk's bits
        uint32_t sum = 0, p = prev;
        for (int i = lo; i < hi; i++) {
            if (!(w.info[i] & kCoded)) continue;
            const uint32_t r = (uint32_t)i - (p ? p : (uint32_t)a);
            w.run[i] = r;
            sum += w.bits[i] + (idr ? 0u : ue_bits(r));
            p = (uint32_t)(i + 1);
        }
        uint32_t body;
        const uint32_t base = block_scan<false>(sum, lds, body);
        CountSink hs;
        slice_header(hs, cfg, a);
        const uint32_t first = byte_base * 8u + hs.bits;
        uint32_t o = first + base;
        p = prev;
        for (int i = lo; i < hi; i++) {
            w.off[i] = o;
            if (!(w.info[i] & kCoded)) continue;
            o += w.bits[i] + (idr ? 0u : ue_bits(w.run[i]));
        }
        // trailing mb_skip_run, stop bit, alignment
        const uint32_t tail_run = idr ? 0u : (uint32_t)b - (last_all ? last_all : (uint32_t)a);
        const uint32_t tail_pos = first + body;
        const uint32_t end = tail_pos + (tail_run ? ue_bits(tail_run) : 0u) + 1u;
        const uint32_t bytes = ((end + 7) >> 3) - byte_base;
        if (t == 0) {
            info[4 * s + 0] = byte_base;
            info[4 * s + 1] = bytes;
            info[4 * s + 2] = tail_pos;
            info[4 * s + 3] = tail_run;
        }
        byte_base += bytes;
    }
    if (t == 0) {
        info[4 * g.ns] = byte_base;
        info[4 * g.ns + 1] = byte_base > cap_bytes ? 1u : 0u;
    }
}

// ------------------------------------------------------------- 4. fill
__global__ __launch_bounds__(256) void cavlc_fill_kernel(uint32_t *out, uint32_t cap_words, const uint32_t *info,
                                                         int ns)
{
    const uint32_t used = min((info[4 * ns] + 3) >> 2, cap_words);
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < used; i += gridDim.x * 256) out[i] = 0;
}

// ------------------------------------------------------------- 5. write
__global__ __launch_bounds__(256) void cavlc_write_kernel(const int16_t *__restrict__ levels,
                                                          const int32_t *__restrict__ dc,
                                                          bc250_h264_slice_cfg_t cfg, Work w, uint32_t *out,
                                                          uint32_t cap_words, uint32_t *info, int mb_blocks)
{
    const Geo g{cfg.width_in_mbs, cfg.width_in_mbs * cfg.height_in_mbs, cfg.num_slices};
    uint32_t *ovf = info + 4 * g.ns + 1;
    if ((int)blockIdx.x >= mb_blocks) {
        // slice headers, trailing runs and stop bits
        const int s = threadIdx.x;
        if (s >= g.ns) return;
        WriteSink h(out, cap_words, info[4 * s] * 8u, ovf);
        slice_header(h, cfg, g.start_of(s));
        h.finish();
        WriteSink e(out, cap_words, info[4 * s + 2], ovf);
        if (info[4 * s + 3]) put_ue(e, info[4 * s + 3]);
        e.put(1, 1);
        e.finish();
        return;
    }
    const int lane = threadIdx.x & 31;
    const int mb = blockIdx.x * 8 + (threadIdx.x >> 5);
    if (mb >= g.total || !(w.info[mb] & kCoded)) return;
    const bool p = !cfg.is_idr;
    const uint32_t run = p ? w.run[mb] : 0u, runbits = p ? ue_bits(run) : 0u;
    const uint32_t lo = w.lane_off[(size_t)mb * kLanesPerMb + lane];
    WriteSink s(out, cap_words, w.off[mb] + (lane ? runbits + lo : 0u), ovf);
    if (lane == 0 && p) put_ue(s, run);
    code_lane(s, levels, dc, cfg, w, g, mb, lane);
    s.finish();
}

// copies slice_info and the slices' bytes (as many as dst_cap allows) to dst
__global__ __launch_bounds__(256) void cavlc_copy_out_kernel(const uint32_t *__restrict__ src,
                                                             const uint32_t *__restrict__ info, int ns,
                                                             uint32_t *__restrict__ dst, uint32_t dst_words,
                                                             uint32_t *__restrict__ dst_info)
{
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t < (uint32_t)(4 * ns + 2)) dst_info[t] = info[t];
    const uint32_t words = min((info[4 * ns] + 3) >> 2, dst_words);
    for (uint32_t i = t; i < words; i += gridDim.x * 256) dst[i] = src[i];
}

}  // namespace

extern "C" int bc250_h264_cavlc_copy_out(const uint8_t *out, const uint32_t *slice_info, int num_slices,
                                         uint8_t *dst, size_t dst_cap, uint32_t *dst_info, void *stream)
{
    if (!out || !slice_info || !dst || !dst_info || num_slices < 1 || num_slices > 64 || ((uintptr_t)out & 3) ||
        ((uintptr_t)dst & 3))
        return -1;
    const uint32_t words = (uint32_t)std::min<size_t>(dst_cap / 4, (size_t)1 << 27);
    cavlc_copy_out_kernel<<<64, 256, 0, (hipStream_t)stream>>>((const uint32_t *)out, slice_info, num_slices,
                                                               (uint32_t *)dst, words, dst_info);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" size_t bc250_h264_cavlc_work_size(int width_in_mbs, int height_in_mbs)
{
    if (width_in_mbs <= 0 || height_in_mbs <= 0) return 0;
    return work_bytes(width_in_mbs * height_in_mbs);
}

extern "C" int bc250_h264_cavlc_slices(const int16_t *levels, const int32_t *dc, const uint32_t *modes,
                                       const gpu_mv_t *mvs, const bc250_h264_slice_cfg_t *cfg, uint8_t *out,
                                       size_t out_cap, uint32_t *slice_info, void *work, size_t work_size,
                                       void *stream)
{
    if (!cfg || !levels || !dc || !out || !slice_info || !work) return -1;
    const bc250_h264_slice_cfg_t c = *cfg;
    if (c.width_in_mbs <= 0 || c.height_in_mbs <= 0) return -1;
    const int64_t total = (int64_t)c.width_in_mbs * c.height_in_mbs;
    // bit offsets are 32-bit: 2^17 MBs (8K UHD is 129600) of at most ~20.3 kbit each stay below 2^32
    if (total > (1 << 17) || c.qp < 0 || c.qp > 51 || (c.is_idr != 0 && c.is_idr != 1)) return -1;
    if (c.num_slices < 1 || c.num_slices > 64 || c.num_slices > total) return -1;
    if (c.log2_max_frame_num_minus4 < 0 || c.log2_max_frame_num_minus4 > 12) return -1;
    if (c.log2_max_poc_lsb_minus4 < 0 || c.log2_max_poc_lsb_minus4 > 12) return -1;
    // PPS / slice header ranges of 7.4.2.2 and 7.4.3
    if (c.deblock_idc < 0 || c.deblock_idc > 2 || c.pps_id < 0 || c.pps_id > 255 || c.frame_num < 0 ||
        c.idr_pic_id < 0 || c.idr_pic_id > 65535 || c.pic_init_qp_minus26 < -26 || c.pic_init_qp_minus26 > 25 ||
        c.poc < 0 || (c.relax_chroma_skip != 0 && c.relax_chroma_skip != 1))
        return -1;
    if (c.is_idr ? !modes : !mvs) return -1;
    if (((uintptr_t)out & 3) || out_cap < 4) return -1;
    if (work_size < work_bytes((int)total)) return -1;
    hipStream_t st = (hipStream_t)stream;
    const Work w = carve(work, (int)total);
    const int blocks = (int)((total + 7) / 8);
    const uint32_t cap_words = (uint32_t)std::min<size_t>(out_cap / 4, (size_t)1 << 27);  // 512 MB of bits addressable
    cavlc_classify_kernel<<<blocks, 256, 0, st>>>(levels, dc, modes, mvs, c, w);
    cavlc_size_kernel<<<blocks, 256, 0, st>>>(levels, dc, c, w);
    cavlc_scan_kernel<<<1, kScanThreads, 0, st>>>(c, w, cap_words * 4u, slice_info);
    cavlc_fill_kernel<<<512, 256, 0, st>>>((uint32_t *)out, cap_words, slice_info, c.num_slices);
    cavlc_write_kernel<<<blocks + 1, 256, 0, st>>>(levels, dc, c, w, (uint32_t *)out, cap_words, slice_info,
                                                  blocks);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ---------------------------------------------------------------------------
// The allocation-free entry point.  Its work buffer is `scratch`, a
// file-scope device array of BC250_CAVLC_SCRATCH_BYTES that every GPU the
// library runs on holds until the process ends.  Host code never passes the
// array's name to a kernel (on the host it is a stand-in, not a device
// address): the device address comes from hipGetSymbolAddress, once per
// device with that device current, and is kept per device.  The writer's
// kernels write every work word they later read, so nothing in the array
// needs clearing between calls; calls that overlap on one device share the
// array and race.
static __device__ __align__(256) uint8_t scratch[BC250_CAVLC_SCRATCH_BYTES];
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices];
}  // namespace

extern "C" size_t bc250_h264_cavlc_noalloc_max_mbs(void)
{
    // the largest frame whose work fits: work_bytes grows with the MB count
    size_t lo = 0, hi = (size_t)1 << 17;
    while (lo < hi) {
        const size_t mid = (lo + hi + 1) / 2;
        if (work_bytes((int)mid) <= BC250_CAVLC_SCRATCH_BYTES) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

extern "C" int bc250_h264_cavlc_slices_noalloc(const int16_t *levels, const int32_t *dc, const uint32_t *modes,
                                               const gpu_mv_t *mvs, const bc250_h264_slice_cfg_t *cfg, uint8_t *out,
                                               size_t out_cap, uint32_t *slice_info, void *stream)
{
    if (!cfg || cfg->width_in_mbs <= 0 || cfg->height_in_mbs <= 0) return -1;
    const int64_t total = (int64_t)cfg->width_in_mbs * cfg->height_in_mbs;
    if (total > (int64_t)bc250_h264_cavlc_noalloc_max_mbs()) {
        fprintf(stderr, "[bc250-hip] bc250_h264_cavlc_slices_noalloc: %lld MBs exceed the %zu-MB limit of its "
                        "%d-byte scratch\n", (long long)total, bc250_h264_cavlc_noalloc_max_mbs(),
                BC250_CAVLC_SCRATCH_BYTES);
        return -1;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return -1;
    void *ptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (!g_scratch_ptr[dev]) {
            void *p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p) {
                fprintf(stderr, "[bc250-hip] hipGetSymbolAddress(scratch) failed on device %d: %s\n", dev,
                        hipGetErrorString(e));
                return -1;
            }
            g_scratch_ptr[dev] = p;
        }
        ptr = g_scratch_ptr[dev];
    }
    return bc250_h264_cavlc_slices(levels, dc, modes, mvs, cfg, out, out_cap, slice_info, ptr,
                                   BC250_CAVLC_SCRATCH_BYTES, stream);
}
