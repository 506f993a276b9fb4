// h264_pred.hip - H.264 intra 4x4 and 8x8 luma prediction (8.3.1.2,
// 8.3.2.2) for gfx950, batched: the per-block function of the reference
// decoder (enc/src/h264_pred.c:45,321) over many blocks at once, one lane
// per predicted sample.  Neighbours (and, for 8x8, their 8.3.2.2.1 low-pass
// filtered forms) sit in LDS per block so every lane can index them.
#include "bc250_dev.h"
#include "bc250_stages.h"

namespace {
BC_DEV int f3(int a, int b, int c) { return (a + 2 * b + c + 2) >> 2; }
BC_DEV int f2(int a, int b) { return (a + b + 1) >> 1; }

// One sample of mode 0..8 for an n x n block; R points at the block's
// references with R[0] = p[-1,-1], R[1 + i] = p[i,-1] (i < 2n) and
// R[1 + 2n + i] = p[-1,i] (i < n).
BC_DEV int dir_sample(int n, int mode, int x, int y, const int *R, int dc)
{
    const int *top = R + 1, *left = R + 1 + 2 * n;
    auto T = [&](int i) { return i < 0 ? R[0] : top[i]; };
    auto L = [&](int i) { return i < 0 ? R[0] : left[i]; };
    const int last = 2 * n - 1;
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
        return f3(top[0], R[0], left[0]);
    case 5: {
        const int z = 2 * x - y;
        if (z >= 0 && !(z & 1)) return f2(T(x - (y >> 1) - 1), top[x - (y >> 1)]);
        if (z > 0) return f3(T(x - (y >> 1) - 2), T(x - (y >> 1) - 1), top[x - (y >> 1)]);
        if (z == -1) return f3(left[0], R[0], top[0]);
        return f3(L(y - 2 * x - 1), L(y - 2 * x - 2), L(y - 2 * x - 3));
    }
    case 6: {
        const int z = 2 * y - x;
        if (z >= 0 && !(z & 1)) return f2(L(y - (x >> 1) - 1), left[y - (x >> 1)]);
        if (z > 0) return f3(L(y - (x >> 1) - 2), L(y - (x >> 1) - 1), left[y - (x >> 1)]);
        if (z == -1) return f3(left[0], R[0], top[0]);
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
}

BC_DEV int dc_of(int n, const int *R, bool at, bool al)
{
    int st = 0, sl = 0;
    for (int i = 0; i < n; i++) {
        st += R[1 + i];
        sl += R[1 + 2 * n + i];
    }
    const int sh = n == 4 ? 2 : 3;
    if (at && al) return (st + sl + n) >> (sh + 1);
    if (al) return (sl + (n >> 1)) >> sh;
    if (at) return (st + (n >> 1)) >> sh;
    return 128;
}
}  // namespace

// 16 blocks of 4x4 per 256-thread workgroup.  nbr[16*b]: p[-1,0..3],
// p[-1,-1], p[0..7,-1]; ctl[b]: mode | top<<4 | left<<5.
__global__ __launch_bounds__(256) void h264_pred4x4_kernel(const uint8_t *__restrict__ nbr,
                                                           const uint8_t *__restrict__ ctl, int n,
                                                           uint8_t *__restrict__ out)
{
    __shared__ int R[16][13];
    const int lb = threadIdx.x >> 4, k = threadIdx.x & 15, b = blockIdx.x * 16 + lb;
    const bool ok = b < n;
    if (ok && k < 13) {
        // R layout: [0] corner, [1..8] top, [9..12] left
        const uint8_t *s = nbr + (size_t)b * 16;
        R[lb][k] = k == 0 ? s[4] : (k <= 8 ? s[4 + k] : s[k - 9]);
    }
    __syncthreads();
    if (!ok) return;
    const int c = ctl[b], mode = c & 15;
    const int dc = dc_of(4, R[lb], (c >> 4) & 1, (c >> 5) & 1);
    out[(size_t)b * 16 + k] = (uint8_t)dir_sample(4, mode, k & 3, k >> 2, R[lb], dc);
}

// 4 blocks of 8x8 per 256-thread workgroup.  nbr[32*b]: p[-1,0..7],
// p[-1,-1], p[0..15,-1]; ctl[b]: mode | top<<4 | left<<5 | corner<<6 |
// top-right<<7.
__global__ __launch_bounds__(256) void h264_pred8x8_kernel(const uint8_t *__restrict__ nbr,
                                                           const uint8_t *__restrict__ ctl, int n,
                                                           uint8_t *__restrict__ out)
{
    __shared__ int raw[4][25], R[4][25];
    const int lb = threadIdx.x >> 6, k = threadIdx.x & 63, b = blockIdx.x * 4 + lb;
    const bool ok = b < n;
    const int c = ok ? ctl[b] : 0;
    const bool at = (c >> 4) & 1, al = (c >> 5) & 1, ac = (c >> 6) & 1, atr = (c >> 7) & 1;
    if (ok && k < 25) {
        // raw layout as R: [0] corner, [1..16] top, [17..24] left; a missing
        // top-right run repeats p[7,-1] before filtering
        const uint8_t *s = nbr + (size_t)b * 32;
        int v = k == 0 ? s[8] : (k <= 16 ? s[8 + k] : s[k - 17]);
        if (k >= 9 && k <= 16 && at && !atr) v = s[8 + 8];
        raw[lb][k] = v;
    }
    __syncthreads();
    if (ok && k < 25) {  // 8.3.2.2.1
        const int *w = raw[lb];
        int v = w[k];
        if (k == 0) {
            if (ac && at && al) v = f3(w[1], w[0], w[17]);
        } else if (k <= 16) {
            const int i = k - 1;
            if (at) {
                if (i == 0) v = ac ? f3(w[0], w[1], w[2]) : (3 * w[1] + w[2] + 2) >> 2;
                else if (i == 15) v = (w[15] + 3 * w[16] + 2) >> 2;
                else v = f3(w[k - 1], w[k], w[k + 1]);
            }
        } else {
            const int i = k - 17;
            if (al) {
                if (i == 0) v = ac ? f3(w[0], w[17], w[18]) : (3 * w[17] + w[18] + 2) >> 2;
                else if (i == 7) v = (w[23] + 3 * w[24] + 2) >> 2;
                else v = f3(w[k - 1], w[k], w[k + 1]);
            }
        }
        R[lb][k] = v;
    }
    __syncthreads();
    if (!ok) return;
    const int dc = dc_of(8, R[lb], at, al);
    out[(size_t)b * 64 + k] = (uint8_t)dir_sample(8, c & 15, k & 7, k >> 3, R[lb], dc);
}

extern "C" int bc250_h264_pred4x4_batch(const uint8_t *nbr, const uint8_t *ctl, int n, uint8_t *out, void *stream)
{
    if (n < 0) return -1;
    if (n == 0) return 0;
    h264_pred4x4_kernel<<<(n + 15) / 16, 256, 0, (hipStream_t)stream>>>(nbr, ctl, n, out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int bc250_h264_pred8x8_batch(const uint8_t *nbr, const uint8_t *ctl, int n, uint8_t *out, void *stream)
{
    if (n < 0) return -1;
    if (n == 0) return 0;
    h264_pred8x8_kernel<<<(n + 3) / 4, 256, 0, (hipStream_t)stream>>>(nbr, ctl, n, out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
