// hevc_dev.h - device helpers for the HEVC stages.
//
// Arithmetic follows H.265 and the reference encoder's chosen forward
// shifts (enc/src/hevc_intra.c), at bit depth bd:
//   4x4 forward  (1, 8) at 8 bits, (3, 8) at 10   inverse (7, 20 - bd)
//   8x8 forward  (bd - 6, 9)                      inverse (7, 20 - bd)
//   quant / dequant bdShift bd + log2(n) - 5, QP' = QP + 6 (bd - 8)
// with int16 saturation after every pass, quantisation
//   level = min(32767, floor(((|v| << shift) + off) / (ls << (q/6 + 4))))
// and the flat-list dequantiser of 8.6.3.
#pragma once
#include "bc250_dev.h"

namespace hevc {

// 8.6.4.2 transMatrix rows for nTbS 4 (DCT-II and the DST-VII of intra
// luma 4x4) and 8, and 8.5.3.3.3 interpolation filters: the standard's
// constants.
static __constant__ int8_t kDct4[16] = {64, 64, 64, 64, 83, 36, -36, -83, 64, -64, -64, 64, 36, -83, 83, -36};
static __constant__ int8_t kDst4[16] = {29, 55, 74, 84, 74, 74, 0, -74, 84, -29, -74, 55, 55, -84, 74, -29};
static __constant__ int8_t kDct8[64] = {64, 64,  64,  64,  64,  64,  64,  64,  89, 75,  50,  18,  -18, -50, -75, -89,
                                 83, 36,  -36, -83, -83, -36, 36,  83,  75, -18, -89, -50, 50,  89,  18,  -75,
                                 64, -64, -64, 64,  64,  -64, -64, 64,  50, -89, 18,  75,  -75, -18, 89,  -50,
                                 36, -83, 83,  -36, -36, 83,  -83, 36,  18, -50, 75,  -89, 89,  -75, 50,  -18};
static __constant__ int8_t kQpel[4][8] = {
    {0, 0, 0, 0, 0, 0, 0, 0}, {-1, 4, -10, 58, 17, -5, 1, 0}, {-1, 4, -11, 40, 40, -11, 4, -1}, {0, 1, -5, 17, 58, -10, 4, -1}};
static __constant__ int8_t kEpel[8][4] = {{0, 0, 0, 0},    {-2, 58, 10, -2}, {-4, 54, 16, -2}, {-6, 46, 28, -4},
                                   {-4, 36, 36, -4}, {-4, 28, 46, -6}, {-2, 16, 54, -4}, {-2, 10, 58, -2}};

BC_DEV int sat16(int v) { return v > 32767 ? 32767 : (v < -32768 ? -32768 : v); }

// Entry (k, n) of the N-point (N = 4..32) DCT-II matrix of 8.6.4.2.  Row k of
// the N-point matrix is row k * 32 / N of the 32-point one, whose entries are
// +-64 sqrt(2) cos(a pi / 64) at the angle index a = (2n + 1) k mod 128 folded
// into 0..32, rounded as the standard's table has them (row 0 is all 64).
// constexpr so that unrolled butterflies fold each entry to a constant.
BC_DEV constexpr int dct_coef(int k, int n, int N)
{
    const int kk = k * (32 / N);
    if (kk == 0) return 64;
    int a = ((2 * n + 1) * kk) & 127, sign = 1;
    if (a > 64) a = 128 - a;
    if (a > 32) {
        a = 64 - a;
        sign = -1;
    }
    const int odd[16] = {90, 90, 88, 85, 82, 78, 73, 67, 61, 54, 46, 38, 31, 22, 13, 4};
    const int by2[8] = {90, 87, 80, 70, 57, 43, 25, 9};
    const int by4[4] = {89, 75, 50, 18};
    const int mag = (a & 1)   ? odd[a >> 1]
                    : (a & 2) ? by2[a >> 2]
                    : (a & 4) ? by4[a >> 3]
                    : (a & 8) ? (a == 8 ? 83 : 36)
                              : 64;
    return sign * mag;
}

BC_DEV int level_scale(int r)
{
    return r == 0 ? 40 : r == 1 ? 45 : r == 2 ? 51 : r == 3 ? 57 : r == 4 ? 64 : 72;
}

BC_DEV int chroma_qp(int qp)
{
    const int q = clampi(qp, 0, 57);
    if (q < 30) return q;
    if (q > 43) return q - 6;
    const int mid[14] = {29, 30, 31, 32, 33, 33, 34, 34, 35, 35, 36, 36, 37, 37};  // Table 8-10
    return mid[q - 30];
}

// Per-QP quantiser constants (uniform across a launch).
struct Quant {
    int p, shift;
    uint32_t d, off, m;  // divisor, rounding offset, ceil(2^32 / d)
};

BC_DEV Quant make_quant(int q, int shift, int round_q12)
{
    Quant Q;
    Q.p = q / 6 + 4;
    Q.shift = shift;
    Q.d = (uint32_t)level_scale(q % 6);
    Q.off = (uint32_t)((((uint64_t)Q.d << Q.p) * (uint64_t)round_q12) / 12u);
    Q.m = 0xffffffffu / Q.d + 1u;  // ceil(2^32 / d) for d > 1, in 32-bit arithmetic
    return Q;
}

// floor(t / d) as a multiply-high: exact because t < 2^21 and m*d - 2^32 < d.
// v is an int16 value and shift <= 10, so mag << shift is a 24-bit product.
BC_DEV int quant(int v, const Quant &Q)
{
    const uint32_t mag = (uint32_t)(v < 0 ? -v : v);
    const uint32_t t = (__umul24(mag, 1u << Q.shift) + Q.off) >> Q.p;
    uint32_t l = __umulhi(t, Q.m);
    l = l > 32767u ? 32767u : l;
    return v < 0 ? -(int)l : (int)l;
}

// 8.6.3 flat-list scaling, kept in 32 bits by splitting on a >= bdShift.
BC_DEV int dequant(int c, int q, int bds)
{
    const int a = q / 6 + 4;
    const int prod = __mul24(c, level_scale(q % 6));  // int16 level x <= 72: 24-bit product is exact
    const int v = a >= bds ? prod * (1 << (a - bds)) : (prod + (1 << (bds - a - 1))) >> (bds - a);
    return sat16(v);
}

// One sample of an 8-bit plane or of a P010 plane (the 10-bit sample in the
// high bits of a 16-bit word); row is a byte pointer to the start of a row.
template <typename Pix>
BC_DEV int ldpix(const uint8_t *row, int x)
{
    if constexpr (sizeof(Pix) == 1) return row[x];
    else return ((const uint16_t *)row)[x] >> 6;
}

template <typename Pix>
BC_DEV void stpix(uint8_t *row, int x, int v)
{
    if constexpr (sizeof(Pix) == 1) row[x] = (uint8_t)v;
    else ((uint16_t *)row)[x] = (uint16_t)(v << 6);
}

// Four consecutive samples x .. x+3 of a row (x a multiple of 4): one dword
// (8-bit) or two (P010, where the sample is word >> 6).
template <typename Pix>
BC_DEV void ld4(const uint8_t *row, int x, int v[4])
{
    if constexpr (sizeof(Pix) == 1) {
        const uint32_t w = *(const uint32_t *)(row + x);
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = (w >> (8 * k)) & 255;
    } else {
        const uint2 w = *(const uint2 *)(row + 2 * x);
        v[0] = (w.x & 0xffffu) >> 6;
        v[1] = w.x >> 22;
        v[2] = (w.y & 0xffffu) >> 6;
        v[3] = w.y >> 22;
    }
}

template <typename Pix>
BC_DEV void st4(uint8_t *row, int x, const int v[4])
{
    if constexpr (sizeof(Pix) == 1) {
        *(uint32_t *)(row + x) = (uint32_t)v[0] | ((uint32_t)v[1] << 8) | ((uint32_t)v[2] << 16) | ((uint32_t)v[3] << 24);
    } else {
        *(uint2 *)(row + 2 * x) = make_uint2(((uint32_t)v[0] << 6) | ((uint32_t)v[1] << 22),
                                             ((uint32_t)v[2] << 6) | ((uint32_t)v[3] << 22));
    }
}

// First row of the slice holding row `row` when slices are the row ranges
// [floor(rows * s / ns), floor(rows * (s + 1) / ns)).
BC_DEV int slice_first_row(int row, int rows, int ns)
{
    ns = ns < 1 ? 1 : (ns > rows ? rows : ns);
    int r0 = 0;
    for (int s = 0; s < ns; s++) {
        const int b = (int)(((long long)rows * s) / ns);
        if (row >= b) r0 = b;
    }
    return r0;
}

// The same in CTU16 rows (16 luma rows) when the slices are ranges of CTU
// rows of 1 << ctu_log2 luma rows (ctu_log2 4 or 5).
BC_DEV int slice_first_row16(int row16, int rows16, int ns, int ctu_log2)
{
    const int s = ctu_log2 - 4;
    return slice_first_row(row16 >> s, (rows16 + s) >> s, ns) << s;
}

// The plain N-point matrix product of the batch transforms (one 256-thread
// workgroup per block, the matrix and both passes in LDS), defined in
// hevc_inter.hip.
int launch_nxn_matrix(const int16_t *in, int n, int log2n, int q, int bd, int round_q12, bool inverse,
                      int16_t *out, hipStream_t stream);

}  // namespace hevc
