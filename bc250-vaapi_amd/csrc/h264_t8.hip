// h264_t8.hip - H.264 8x8 transform pair for gfx950, batched, one wave per
// 8x8 block (lane = coefficient):
//   forward   Y = C X C^T with the 8x8 integer basis C (exact, no rounding)
//   quant     level = floor((|Y| * 2^14 + off) / step), step(i,j) =
//             n_i n_j v8[qp%6][class] 2^(qp/6), n = diag(C C^T), off =
//             round_q12/12 of a step (the algebraic inverse of the decoder)
//   dequant   8.5.13.1, flat weight scale 16; inverse 8.5.13.2, (x+32)>>6
// The decoder side is the reference decoder's scale8 + h264d_idct8_add
// (enc/src/h264_recon.c:87-93,192-210).
#include "bc250_dev.h"
#include "bc250_stages.h"

namespace {
__constant__ int8_t kC8[64] = {8,  8,   8,  8,   8,  8,   8,  8,  12, 10,  6,   3,  -3, -6,  -10, -12,
                               8,  4,   -4, -8,  -8, -4,  4,  8,  10, -3,  -12, -6, 6,  12,  3,   -10,
                               8,  -8,  -8, 8,   8,  -8,  -8, 8,  6,  -12, 3,   10, -10, -3, 12,  -6,
                               4,  -8,  8,  -4,  -4, 8,   -8, 4,  3,  -6,  10,  -12, 12, -10, 6,  -3};
__constant__ int16_t kN8[8] = {512, 578, 320, 578, 512, 578, 320, 578};
__constant__ uint8_t kV8[6][6] = {{20, 18, 32, 19, 25, 24}, {22, 19, 35, 21, 28, 26}, {26, 23, 42, 24, 33, 31},
                                  {28, 25, 45, 26, 35, 33}, {32, 28, 51, 30, 40, 38}, {36, 32, 58, 34, 46, 43}};

// 8.5.9 position class of (row i, column j)
BC_DEV int class8(int i, int j)
{
    if (!(i & 3) && !(j & 3)) return 0;
    if ((i & 1) && (j & 1)) return 1;
    if ((i & 3) == 2 && (j & 3) == 2) return 2;
    if ((!(i & 3) && (j & 1)) || ((i & 1) && !(j & 3))) return 3;
    if ((!(i & 3) && (j & 3) == 2) || ((i & 3) == 2 && !(j & 3))) return 4;
    return 5;
}

// 8.5.13.2 one-dimensional inverse of v[0..7] at stride s, in place
BC_DEV void inv8_1d(int *v, int s)
{
    const int d0 = v[0], d1 = v[s], d2 = v[2 * s], d3 = v[3 * s], d4 = v[4 * s], d5 = v[5 * s], d6 = v[6 * s],
              d7 = v[7 * s];
    const int e0 = d0 + d4, e1 = -d3 + d5 - d7 - (d7 >> 1), e2 = d0 - d4, e3 = d1 + d7 - d3 - (d3 >> 1);
    const int e4 = (d2 >> 1) - d6, e5 = -d1 + d7 + d5 + (d5 >> 1), e6 = d2 + (d6 >> 1), e7 = d3 + d5 + d1 + (d1 >> 1);
    const int f0 = e0 + e6, f1 = e1 + (e7 >> 2), f2 = e2 + e4, f3 = e3 + (e5 >> 2);
    const int f4 = e2 - e4, f5 = (e3 >> 2) - e5, f6 = e0 - e6, f7 = e7 - (e1 >> 2);
    v[0] = f0 + f7;
    v[s] = f2 + f5;
    v[2 * s] = f4 + f3;
    v[3 * s] = f6 + f1;
    v[4 * s] = f6 - f1;
    v[5 * s] = f4 - f3;
    v[6 * s] = f2 - f5;
    v[7 * s] = f0 - f7;
}
}  // namespace

__global__ __launch_bounds__(256) void h264_tq8x8_kernel(const int16_t *__restrict__ res, int n, int qp,
                                                         int round_q12, int16_t *__restrict__ out)
{
    __shared__ int X[4][64];
    __shared__ long long T[4][64];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, b = blockIdx.x * 4 + wv;
    const int i = lane >> 3, j = lane & 7;
    X[wv][lane] = b < n ? res[(size_t)b * 64 + lane] : 0;
    __syncthreads();
    long long t = 0;
#pragma unroll
    for (int r = 0; r < 8; r++) t += (long long)kC8[i * 8 + r] * X[wv][r * 8 + j];
    T[wv][lane] = t;
    __syncthreads();
    long long y = 0;
#pragma unroll
    for (int c = 0; c < 8; c++) y += (long long)kC8[j * 8 + c] * T[wv][i * 8 + c];
    const unsigned long long step =
        ((unsigned long long)kN8[i] * (unsigned long long)kN8[j] * kV8[qp % 6][class8(i, j)]) << (qp / 6);
    const unsigned long long off = step * (unsigned long long)round_q12 / 12ull;
    const unsigned long long mag = (unsigned long long)(y < 0 ? -y : y);
    unsigned long long l = ((mag << 14) + off) / step;
    l = l > 32767ull ? 32767ull : l;
    if (b < n) out[(size_t)b * 64 + lane] = (int16_t)(y < 0 ? -(long long)l : (long long)l);
}

__global__ __launch_bounds__(256) void h264_dqit8x8_kernel(const int16_t *__restrict__ lv, int n, int qp,
                                                           int16_t *__restrict__ out)
{
    __shared__ int D[4][64];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, b = blockIdx.x * 4 + wv;
    const int per = qp / 6;
    const int c = b < n ? lv[(size_t)b * 64 + lane] : 0;
    const int ls = 16 * kV8[qp % 6][class8(lane >> 3, lane & 7)];
    D[wv][lane] = per >= 6 ? (c * ls) * (1 << (per - 6)) : (c * ls + (1 << (5 - per))) >> (6 - per);
    __syncthreads();
    if (lane < 8) inv8_1d(&D[wv][lane * 8], 1);  // rows
    __syncthreads();
    if (lane < 8) inv8_1d(&D[wv][lane], 8);      // columns
    __syncthreads();
    if (b < n) out[(size_t)b * 64 + lane] = (int16_t)((D[wv][lane] + 32) >> 6);
}

extern "C" int bc250_h264_tq8x8_batch(const int16_t *res, int n, int qp, int round_q12, int16_t *out, void *stream)
{
    if (n < 0 || qp < 0 || qp > 51 || round_q12 < 0 || round_q12 > 12) return -1;
    if (n == 0) return 0;
    h264_tq8x8_kernel<<<(n + 3) / 4, 256, 0, (hipStream_t)stream>>>(res, n, qp, round_q12, out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int bc250_h264_dqit8x8_batch(const int16_t *lv, int n, int qp, int16_t *out, void *stream)
{
    if (n < 0 || qp < 0 || qp > 51) return -1;
    if (n == 0) return 0;
    h264_dqit8x8_kernel<<<(n + 3) / 4, 256, 0, (hipStream_t)stream>>>(lv, n, qp, out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
