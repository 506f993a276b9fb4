// hevc_t32.hip - the 32x32 HEVC transform pair (forward + quant, dequant +
// inverse) in two exact forms for gfx950:
//
//  * register partial butterflies on the VALU: one lane per column, then
//    per row, two blocks per wave;
//  * a split-int8 product on the matrix cores: each int16 operand is split
//    into its signed high byte and its low byte biased by -128, so that
//    A x = 256 (A hi) + (A lo) + 128 rowsum(A), every part an exact
//    v_mfma_i32_32x32x32_i8 with int32 accumulation; one wave per block.
//
// Both equal the plain matrix product of 8.6.4.2 with rounding and int16
// saturation after each pass (columns first), forward shifts (bd - 4, 11),
// inverse shifts (7, 20 - bd), quantiser / dequantiser shift bd.
#include "bc250_hevc.h"
#include "hevc_dev.h"

using namespace hevc;

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// y = M x for the 32-point matrix by even/odd decomposition: M[k][31 - n] =
// (-1)^k M[k][n], so odd rows need x[n] - x[31 - n] only and even rows are the
// same split one level down on x[n] + x[31 - n], to two points.
BC_DEV void fwd32(const int x[32], int y[32])
{
    int e[16], o[16];
#pragma unroll
    for (int n = 0; n < 16; n++) {
        e[n] = x[n] + x[31 - n];
        o[n] = x[n] - x[31 - n];
    }
#pragma unroll
    for (int k = 1; k < 32; k += 2) {
        int s = 0;
#pragma unroll
        for (int n = 0; n < 16; n++) s += dct_coef(k, n, 32) * o[n];
        y[k] = s;
    }
    int ee[8], eo[8];
#pragma unroll
    for (int n = 0; n < 8; n++) {
        ee[n] = e[n] + e[15 - n];
        eo[n] = e[n] - e[15 - n];
    }
#pragma unroll
    for (int k = 2; k < 32; k += 4) {
        int s = 0;
#pragma unroll
        for (int n = 0; n < 8; n++) s += dct_coef(k, n, 32) * eo[n];
        y[k] = s;
    }
    int eee[4], eeo[4];
#pragma unroll
    for (int n = 0; n < 4; n++) {
        eee[n] = ee[n] + ee[7 - n];
        eeo[n] = ee[n] - ee[7 - n];
    }
#pragma unroll
    for (int k = 4; k < 32; k += 8) {
        int s = 0;
#pragma unroll
        for (int n = 0; n < 4; n++) s += dct_coef(k, n, 32) * eeo[n];
        y[k] = s;
    }
    const int e4[2] = {eee[0] + eee[3], eee[1] + eee[2]}, o4[2] = {eee[0] - eee[3], eee[1] - eee[2]};
#pragma unroll
    for (int k = 0; k < 32; k += 8) {
        const int *v = (k & 8) ? o4 : e4;
        y[k] = dct_coef(k, 0, 32) * v[0] + dct_coef(k, 1, 32) * v[1];
    }
}

// y = M^T x, the same decomposition run backwards.
BC_DEV void inv32(const int x[32], int y[32])
{
    int o[16], eo[8], eeo[4], o4[2], e4[2];
#pragma unroll
    for (int n = 0; n < 16; n++) {
        int s = 0;
#pragma unroll
        for (int k = 1; k < 32; k += 2) s += dct_coef(k, n, 32) * x[k];
        o[n] = s;
    }
#pragma unroll
    for (int n = 0; n < 8; n++) {
        int s = 0;
#pragma unroll
        for (int k = 2; k < 32; k += 4) s += dct_coef(k, n, 32) * x[k];
        eo[n] = s;
    }
#pragma unroll
    for (int n = 0; n < 4; n++) {
        int s = 0;
#pragma unroll
        for (int k = 4; k < 32; k += 8) s += dct_coef(k, n, 32) * x[k];
        eeo[n] = s;
    }
#pragma unroll
    for (int n = 0; n < 2; n++) {
        o4[n] = dct_coef(8, n, 32) * x[8] + dct_coef(24, n, 32) * x[24];
        e4[n] = dct_coef(0, n, 32) * x[0] + dct_coef(16, n, 32) * x[16];
    }
    int eee[4], ee[8], e[16];
#pragma unroll
    for (int n = 0; n < 2; n++) {
        eee[n] = e4[n] + o4[n];
        eee[3 - n] = e4[n] - o4[n];
    }
#pragma unroll
    for (int n = 0; n < 4; n++) {
        ee[n] = eee[n] + eeo[n];
        ee[7 - n] = eee[n] - eeo[n];
    }
#pragma unroll
    for (int n = 0; n < 8; n++) {
        e[n] = ee[n] + eo[n];
        e[15 - n] = ee[n] - eo[n];
    }
#pragma unroll
    for (int n = 0; n < 16; n++) {
        y[n] = e[n] + o[n];
        y[31 - n] = e[n] - o[n];
    }
}

BC_DEV uint32_t pack16(int lo, int hi) { return (uint32_t)(lo & 0xffff) | (uint32_t)hi << 16; }

// Dequantises the two int16 coefficients of a word.
BC_DEV uint32_t dequant2(uint32_t w, int q, int bds)
{
    return pack16(dequant((int16_t)(w & 0xffff), q, bds), dequant((int16_t)(w >> 16), q, bds));
}

BC_DEV uint4 dequant8(uint4 v, int q, int bds)
{
    return uint4{dequant2(v.x, q, bds), dequant2(v.y, q, bds), dequant2(v.z, q, bds), dequant2(v.w, q, bds)};
}

// ---------------------------------------------------------------------------
// Butterflies: lane l works on block 2 p + (l >> 5) of pair p, column then row
// l & 31.  LDS per block: the 32x32 input (64 B rows) and the intermediate
// (80 B rows); the second block of each buffer sits 64 B past a multiple of
// 256 B so that the two half-waves address different banks.

constexpr int kBfX = 2048 + 64;  // input tile stride between the two blocks
constexpr int kBfT = 2560 + 64;  // intermediate tile stride
constexpr int kTP = 80;          // intermediate row pitch (bytes)

template <bool INVERSE>
__global__ __launch_bounds__(64) void hevc_t32_bfly_kernel(const int16_t *__restrict__ in, int n, int q, int bd,
                                                           int round_q12, int16_t *__restrict__ out)
{
    __shared__ __align__(16) uint8_t xs[2 * kBfX];
    __shared__ __align__(16) uint8_t ts[2 * kBfT];
    const int l = threadIdx.x, m = l & 31, h = l >> 5;
    const int s1 = INVERSE ? 7 : bd - 4, s2 = INVERSE ? 20 - bd : 11;
    const int qd = clampi(q, 0, 63);
    const Quant Q = make_quant(clampi(q, 0, bd == 8 ? 51 : 63), bd, round_q12);
    const int pairs = (n + 1) >> 1;
    for (int p = blockIdx.x; p < pairs; p += gridDim.x) {
        const int b0 = 2 * p, nb = min(2, n - b0);
        // 4 KB of two consecutive blocks: chunk i of lane l is bytes
        // 1024 i + 16 l, i.e. block i >> 1, its half i & 1.
#pragma unroll
        for (int i = 0; i < 4; i++) {
            uint4 v = uint4{0, 0, 0, 0};
            if ((i >> 1) < nb) v = *(const uint4 *)((const uint8_t *)in + (size_t)b0 * 2048 + i * 1024 + l * 16);
            if (INVERSE) v = dequant8(v, qd, bd);
            *(uint4 *)(xs + (i >> 1) * kBfX + (i & 1) * 1024 + l * 16) = v;
        }
        __syncthreads();
        int x[32], y[32];
        const uint8_t *col = xs + h * kBfX + m * 2;
#pragma unroll
        for (int r = 0; r < 32; r++) x[r] = *(const int16_t *)(col + r * 64);
        if (INVERSE) inv32(x, y);
        else fwd32(x, y);
        uint8_t *tcol = ts + h * kBfT + m * 2;
#pragma unroll
        for (int i = 0; i < 32; i++)
            *(int16_t *)(tcol + i * kTP) = (int16_t)sat16((y[i] + (1 << (s1 - 1))) >> s1);
        __syncthreads();
        const uint4 *trow = (const uint4 *)(ts + h * kBfT + m * kTP);
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const uint4 v = trow[c];
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int j = 0; j < 4; j++) {
                x[c * 8 + 2 * j] = (int16_t)(w[j] & 0xffff);
                x[c * 8 + 2 * j + 1] = (int16_t)(w[j] >> 16);
            }
        }
        if (INVERSE) inv32(x, y);
        else fwd32(x, y);
        if (h < nb) {
            uint4 *dst = (uint4 *)(out + (size_t)(b0 + h) * 1024 + m * 32);
#pragma unroll
            for (int c = 0; c < 4; c++) {
                int v[8];
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    v[j] = sat16((y[c * 8 + j] + (1 << (s2 - 1))) >> s2);
                    if (!INVERSE) v[j] = quant(v[j], Q);
                }
                dst[c] = uint4{pack16(v[0], v[1]), pack16(v[2], v[3]), pack16(v[4], v[5]), pack16(v[6], v[7])};
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Split-int8 matrix-core form, one wave per block.
//
// v_mfma_i32_32x32x32_i8 lane layout: lane l supplies row l & 31 of A and
// column l & 31 of B, each as 16 bytes of K; lanes l and l + 32 hold the two
// halves of K.  Any assignment of K to (l >> 5, byte) works as long as A and
// B share it.  Result element k of lane l is D[drow(k, l >> 5)][l & 31],
// drow(k, h) = 8 (k >> 2) + 4 h + (k & 3).
//
// M' = M (forward) or M^T (inverse).
// Pass 1: T^T = X^T M'^T: A = X^T (lane c holds column c of the staged input,
// K = rows 16 h .. 16 h + 15), B = M'^T (constant).  Lane i then holds
// T[i][c] for the 16 columns c = drow(k, h).
// Pass 2: Y^T = M' T^T: B = T^T with exactly those values, K = c assigned to
// (h, byte k) by c = drow(k, h); A = M' in the same K order (constant).  Lane
// i holds Y[i][j] for j = drow(k, h).  No exchange between the passes.

constexpr int kMfHalf = 1088;  // byte offset of input rows 16..31 (64 B gap -> other banks)

BC_DEV int drow(int k, int h) { return (k >> 2) * 8 + h * 4 + (k & 3); }

// Splits 16 int16 (8 words, K in increasing order) into the signed high bytes
// and the low bytes minus 128.
BC_DEV void split16(const uint32_t p[8], v4i &hi, v4i &lo)
{
#pragma unroll
    for (int d = 0; d < 4; d++) {
        hi[d] = (int)__builtin_amdgcn_perm(p[2 * d + 1], p[2 * d], 0x07050301u);
        lo[d] = (int)(__builtin_amdgcn_perm(p[2 * d + 1], p[2 * d], 0x06040200u) ^ 0x80808080u);
    }
}

// Two int32 saturated to int16 and packed (low half = a).
BC_DEV uint32_t sat_pack(int a, int b)
{
    typedef short s2 __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(uint32_t, (s2)__builtin_amdgcn_cvt_pk_i16(a, b));
}

// Waves per workgroup; each wave works on its own blocks with its own LDS.
#ifndef BC250_T32_WAVES
#define BC250_T32_WAVES 4
#endif
constexpr int kMfWaves = BC250_T32_WAVES;

// Orders LDS stores before later LDS loads of other lanes of the same wave: a
// wave's LDS operations complete in issue order, so only the compiler must not
// move memory operations across this point.
BC_DEV void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Byte-packed 16 K values of row (col) `row` of M' = M or M^T, K given by kof.
template <bool INVERSE, typename KOf>
BC_DEV v4i pack_rows(int row, KOf kof, int &sum)
{
    v4i r;
    sum = 0;
#pragma unroll
    for (int d = 0; d < 4; d++) {
        uint32_t w = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int kk = kof(4 * d + j);
            const int c = INVERSE ? dct_coef(kk, row, 32) : dct_coef(row, kk, 32);
            sum += c;
            w |= (uint32_t)(c & 255) << (8 * j);
        }
        r[d] = (int)w;
    }
    return r;
}

template <bool INVERSE>
__global__ __launch_bounds__(64 * kMfWaves) void hevc_t32_mfma_kernel(const int16_t *__restrict__ in, int n, int q,
                                                                      int bd, int round_q12,
                                                                      int16_t *__restrict__ out)
{
    __shared__ __align__(16) uint8_t xs_all[kMfWaves][2 * 1024 + 64];
    __shared__ int rs[32];
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, m = l & 31, h = l >> 5;
    uint8_t *xs = xs_all[wv];
    // pass-1 B: row m of M', K = 16 h + t; pass-2 A: row m of M', K = drow(t, h)
    int part, unused;
    const v4i b1 = pack_rows<INVERSE>(m, [h](int t) { return 16 * h + t; }, part);
    const v4i a2 = pack_rows<INVERSE>(m, [h](int t) { return drow(t, h); }, unused);
    const int rowsum = part + __shfl_xor(part, 32);  // sum of row m of M'
    if (wv == 0 && h == 0) rs[m] = rowsum;
    __syncthreads();
    const int s1 = INVERSE ? 7 : bd - 4, s2 = INVERSE ? 20 - bd : 11;
    // x = 256 hi + (lo + 128): the 128 sum(M' row) term and the rounding
    // offset enter through the accumulator.
    const int c1 = 128 * rowsum + (1 << (s1 - 1));
    int c2[16];
#pragma unroll
    for (int k = 0; k < 16; k++) c2[k] = 128 * rs[drow(k, h)] + (1 << (s2 - 1));
    const int qd = clampi(q, 0, 63);
    const Quant Q = make_quant(clampi(q, 0, bd == 8 ? 51 : 63), bd, round_q12);

    const uint8_t *src = (const uint8_t *)in;
    const int stride = gridDim.x * kMfWaves;
    int b = blockIdx.x * kMfWaves + wv;
    if (b >= n) return;
    uint4 v0 = *(const uint4 *)(src + (size_t)b * 2048 + l * 16);
    uint4 v1 = *(const uint4 *)(src + (size_t)b * 2048 + 1024 + l * 16);
    for (; b < n; b += stride) {
        uint4 cur0 = v0, cur1 = v1;
        // the next block's load is issued before this block's work (the last
        // iteration reloads block n - 1: no branch for the loads to sink into)
        const int nb = min(b + stride, n - 1);
        v0 = *(const uint4 *)(src + (size_t)nb * 2048 + l * 16);
        v1 = *(const uint4 *)(src + (size_t)nb * 2048 + 1024 + l * 16);
        if (INVERSE) {
            cur0 = dequant8(cur0, qd, bd);
            cur1 = dequant8(cur1, qd, bd);
        }
        wave_lds_sync();
        *(uint4 *)(xs + l * 16) = cur0;
        *(uint4 *)(xs + kMfHalf + l * 16) = cur1;
        wave_lds_sync();
        uint32_t p[8];
        const uint8_t *col = xs + h * kMfHalf + m * 2;
#pragma unroll
        for (int t = 0; t < 8; t++)
            p[t] = (uint32_t) * (const uint16_t *)(col + 2 * t * 64) |
                   (uint32_t) * (const uint16_t *)(col + (2 * t + 1) * 64) << 16;
        v4i hi, lo;
        split16(p, hi, lo);
        v16i acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(hi, b1, v16i{}, 0, 0, 0);
#pragma unroll
        for (int k = 0; k < 16; k++) acc[k] = (acc[k] << 8) + c1;
        acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(lo, b1, acc, 0, 0, 0);
#pragma unroll
        for (int t = 0; t < 8; t++) p[t] = sat_pack(acc[2 * t] >> s1, acc[2 * t + 1] >> s1);
        split16(p, hi, lo);
        acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(a2, hi, v16i{}, 0, 0, 0);
#pragma unroll
        for (int k = 0; k < 16; k++) acc[k] = (acc[k] << 8) + c2[k];
        acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(a2, lo, acc, 0, 0, 0);
        int16_t *dst = out + (size_t)b * 1024 + m * 32 + h * 4;
#pragma unroll
        for (int g = 0; g < 4; g++) {
            uint2 w;
            if (INVERSE) {
                w = uint2{sat_pack(acc[4 * g] >> s2, acc[4 * g + 1] >> s2),
                          sat_pack(acc[4 * g + 2] >> s2, acc[4 * g + 3] >> s2)};
            } else {
                int y[4];
#pragma unroll
                for (int j = 0; j < 4; j++) y[j] = quant(sat16(acc[4 * g + j] >> s2), Q);
                w = uint2{pack16(y[0], y[1]), pack16(y[2], y[3])};
            }
            *(uint2 *)(dst + g * 8) = w;
        }
    }
}

int launch_ok() { return hipGetLastError() == hipSuccess ? 0 : -1; }

// Resident waves, each looping over blocks.
#ifndef BC250_T32_GRID
#define BC250_T32_GRID 4096
#endif
constexpr int kT32Grid = BC250_T32_GRID;

template <bool INVERSE>
int launch_t32(const int16_t *in, int n, int q, int bd, int round_q12, int method, int16_t *out, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    if (method == BC250_T32_MATRIX) return launch_nxn_matrix(in, n, 5, q, bd, round_q12, INVERSE, out, st);
    if (method == BC250_T32_BUTTERFLY) {
        const int pairs = (n + 1) / 2;
        hevc_t32_bfly_kernel<INVERSE><<<min(pairs, kT32Grid), 64, 0, st>>>(in, n, q, bd, round_q12, out);
    } else {
        const int groups = min((n + kMfWaves - 1) / kMfWaves, kT32Grid / kMfWaves);
        hevc_t32_mfma_kernel<INVERSE><<<groups, 64 * kMfWaves, 0, st>>>(in, n, q, bd, round_q12, out);
    }
    return launch_ok();
}

}  // namespace

extern "C" int bc250_hevc_tq32_batch(const int16_t *res, int n, int qp, int bd, int round_q12, int method,
                                     int16_t *out, void *stream)
{
    if (n < 0 || (bd != 8 && bd != 10) || round_q12 < 0 || round_q12 > 12) return -1;
    if (method < BC250_T32_BUTTERFLY || method > BC250_T32_MATRIX) return -1;
    if (n == 0) return 0;
    return launch_t32<false>(res, n, qp, bd, round_q12, method, out, stream);
}

extern "C" int bc250_hevc_dqit32_batch(const int16_t *coef, int n, int qp, int bd, int method, int16_t *out,
                                       void *stream)
{
    if (n < 0 || (bd != 8 && bd != 10)) return -1;
    if (method < BC250_T32_BUTTERFLY || method > BC250_T32_MATRIX) return -1;
    if (n == 0) return 0;
    return launch_t32<true>(coef, n, qp, bd, 0, method, out, stream);
}
