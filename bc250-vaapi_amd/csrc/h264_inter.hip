// h264_inter.hip - P16x16 macroblock encode for gfx950, fused: quarter-pel
// luma / eighth-pel chroma motion compensation, residual, forward 4x4 core
// transform, inter quantisation, chroma DC Hadamard path, dequantisation,
// inverse transform and reconstruction, one wave per macroblock.  Only the
// products the host consumes (levels, dc, nz, modes) and the reconstruction
// leave the chip; residual, prediction and coefficients stay in LDS.
#include "bc250_dev.h"
#include "bc250_stages.h"

__global__ __launch_bounds__(64) void h264_p_encode_kernel(
    const uint8_t *__restrict__ cy, const uint8_t *__restrict__ cuv, int cpy, int cpuv,
    const uint8_t *__restrict__ ry, const uint8_t *__restrict__ ruv, int rpy, int rpuv,
    uint8_t *__restrict__ oy, uint8_t *__restrict__ ouv, int opy, int opuv, int wc, int hc, int mbw,
    int nmb, int qp, const uint4 *__restrict__ mvs, int16_t *__restrict__ levels,
    int32_t *__restrict__ dc, uint32_t *__restrict__ nz, uint32_t *__restrict__ modes)
{
    __shared__ uint8_t w[24 * 24];
    __shared__ uint8_t wc8[2][10 * 10];
    __shared__ int pred[24][16];
    __shared__ int res[24][16];
    __shared__ int lv[24][16];
    __shared__ int cdc[2][4];
    __shared__ int cdin[8];

    const int mb = xcd_item(blockIdx.x, nmb);
    if (mb >= nmb) return;
    const int lane = threadIdx.x;
    const int mbx = mb % mbw, mby = mb / mbw;
    const int X0 = mbx * 16, Y0 = mby * 16, U0 = mbx * 8, V0 = mby * 8;
    const int hcc = hc / 2, wcc = wc / 2;
    const uint4 mv = mvs[mb];
    const int mvx = (int)mv.x, mvy = (int)mv.y;
    const int ivx = mvx >> 2, ivy = mvy >> 2, fx = mvx & 3, fy = mvy & 3;
    const int cvx = mvx >> 3, cvy = mvy >> 3, cfx = mvx & 7, cfy = mvy & 7;

    // luma window: block + 4 on each side, replicated edges
    for (int i = lane; i < 24 * 24; i += 64) {
        int yy = i / 24, xx = i - yy * 24;
        int gx = clampi(X0 + ivx - 4 + xx, 0, wc - 1), gy = clampi(Y0 + ivy - 4 + yy, 0, hc - 1);
        w[i] = ry[(size_t)gy * rpy + gx];
    }
    // chroma windows: 9x9 integer samples per component
    for (int i = lane; i < 2 * 81; i += 64) {
        int comp = i / 81, k = i - comp * 81, yy = k / 9, xx = k - yy * 9;
        int gx = clampi(U0 + cvx + xx, 0, wcc - 1), gy = clampi(V0 + cvy + yy, 0, hcc - 1);
        wc8[comp][yy * 10 + xx] = ruv[(size_t)gy * rpuv + 2 * gx + comp];
    }
    __syncthreads();

    // prediction + residual: 256 luma samples, 128 chroma samples
    for (int p = lane; p < 256; p += 64) {
        int x = p & 15, y = p >> 4;
        int v = qpel_win(w, 24, x + 4, y + 4, fx, fy);
        int b = (y >> 2) * 4 + (x >> 2), k = (y & 3) * 4 + (x & 3);
        pred[b][k] = v;
        res[b][k] = (int)cy[(size_t)(Y0 + y) * cpy + X0 + x] - v;
    }
    for (int p = lane; p < 128; p += 64) {
        int comp = p >> 6, q = p & 63, u = q & 7, v = q >> 3;
        const uint8_t *s = &wc8[comp][v * 10 + u];
        int A = s[0], B = s[1], C = s[10], D = s[11];
        int pv = ((8 - cfx) * (8 - cfy) * A + cfx * (8 - cfy) * B + (8 - cfx) * cfy * C + cfx * cfy * D + 32) >> 6;
        int b = 16 + comp * 4 + (v >> 2) * 2 + (u >> 2), k = (v & 3) * 4 + (u & 3);
        pred[b][k] = pv;
        res[b][k] = (int)cuv[(size_t)(V0 + v) * cpuv + 2 * (U0 + u) + comp] - pv;
    }
    __syncthreads();

    const int qpc = chroma_qp(qp);
    int co0 = 0;
    if (lane < 24) {
        const int b = lane, q = b < 16 ? qp : qpc;
        const int per = q / 6, rem = q % 6, f = (1 << (15 + per)) / 6;
        int r[16], co[16];
#pragma unroll
        for (int k = 0; k < 16; k++) r[k] = res[b][k];
        fwd4x4(r, co);
        co0 = co[0];
        uint32_t msk = 0, pk[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int k = 0; k < 16; k++) {
            int l = quant_ac(co[k], k, per, rem, f);
            lv[b][k] = l;
            pk[k >> 1] |= (uint32_t)(uint16_t)(int16_t)l << (16 * (k & 1));
            msk |= (l != 0 ? 1u : 0u) << k;
        }
        uint4 *dst = (uint4 *)(levels + ((size_t)mb * 24 + b) * 16);
        dst[0] = make_uint4(pk[0], pk[1], pk[2], pk[3]);
        dst[1] = make_uint4(pk[4], pk[5], pk[6], pk[7]);
        dc[mb * 24 + b] = co[0];
        nz[mb * 24 + b] = msk;
    }
    // chroma DC inputs to LDS (lanes 16..23 hold blocks 16..23)
    if (lane >= 16 && lane < 24) cdin[lane - 16] = co0;
    if (lane == 0) modes[mb] = 2u;
    __syncthreads();
    if (lane < 2) {
        const int per = qpc / 6, rem = qpc % 6;
        int din[4], raw[4], l4[4], f4[4];
#pragma unroll
        for (int s = 0; s < 4; s++) din[s] = cdin[lane * 4 + s];
        hadamard2x2(din, raw);
#pragma unroll
        for (int s = 0; s < 4; s++) l4[s] = quant_dc_k5(raw[s], per, rem);
        hadamard2x2(l4, f4);
#pragma unroll
        for (int s = 0; s < 4; s++) cdc[lane][s] = dequant_dc_chroma(f4[s], per, rem);
    }
    __syncthreads();

    if (lane < 16) {
        const int b = lane, per = qp / 6, rem = qp % 6;
        int d[16], r[16];
#pragma unroll
        for (int k = 0; k < 16; k++) d[k] = dequant_ac(lv[b][k], k, per, rem);
        inv4x4(d, r);
        const int bx = b & 3, by = b >> 2;
#pragma unroll
        for (int row = 0; row < 4; row++) {
            uint32_t word = 0;
#pragma unroll
            for (int c = 0; c < 4; c++) word |= (uint32_t)clip255(pred[b][row * 4 + c] + r[row * 4 + c]) << (8 * c);
            *(uint32_t *)(oy + (size_t)(Y0 + by * 4 + row) * opy + X0 + bx * 4) = word;
        }
    } else if (lane < 20) {
        // one lane per chroma quadrant reconstructs Cb and Cr together
        const int s = lane - 16, per = qpc / 6, rem = qpc % 6;
        int rr[2][16];
#pragma unroll
        for (int comp = 0; comp < 2; comp++) {
            const int b = 16 + comp * 4 + s;
            int d[16];
            d[0] = cdc[comp][s];
#pragma unroll
            for (int k = 1; k < 16; k++) d[k] = dequant_ac(lv[b][k], k, per, rem);
            inv4x4(d, rr[comp]);
        }
        const int ux = U0 + (s & 1) * 4, vy = V0 + (s >> 1) * 4;
#pragma unroll
        for (int row = 0; row < 4; row++) {
            uint32_t lo = 0, hi = 0;
#pragma unroll
            for (int c = 0; c < 4; c++) {
                int k = row * 4 + c;
                uint32_t cbv = (uint32_t)clip255(pred[16 + s][k] + rr[0][k]);
                uint32_t crv = (uint32_t)clip255(pred[20 + s][k] + rr[1][k]);
                uint32_t pair = cbv | (crv << 8);
                if (c < 2) lo |= pair << (16 * c);
                else hi |= pair << (16 * (c - 2));
            }
            *(uint2 *)(ouv + (size_t)(vy + row) * opuv + 2 * ux) = make_uint2(lo, hi);
        }
    }
}

extern "C" int bc250_launch_p_encode(const uint8_t *cy, const uint8_t *cuv, int cpy, int cpuv,
                                     const uint8_t *ry, const uint8_t *ruv, int rpy, int rpuv, uint8_t *oy,
                                     uint8_t *ouv, int opy, int opuv, int wc, int hc, int qp,
                                     const uint32_t *mvs, int16_t *levels, int32_t *dc, uint32_t *nz,
                                     uint32_t *modes, void *stream)
{
    if (wc <= 0 || hc <= 0 || (wc & 15) || (hc & 15) || qp < 0 || qp > 51 || (opy & 3) || (opuv & 7) ||
        (cpy & 3) || (cpuv & 7))
        return -1;
    const int mbw = wc / 16, nmb = mbw * (hc / 16);
    const int grid = ((nmb + 7) / 8) * 8;
    h264_p_encode_kernel<<<grid, 64, 0, (hipStream_t)stream>>>(cy, cuv, cpy, cpuv, ry, ruv, rpy, rpuv, oy, ouv,
                                                                opy, opuv, wc, hc, mbw, nmb, qp,
                                                                (const uint4 *)mvs, levels, dc, nz, modes);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
