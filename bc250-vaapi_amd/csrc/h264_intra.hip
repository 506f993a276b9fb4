// h264_intra.hip - I-slice encode as ONE persistent wavefront kernel for
// gfx950 (replaces the reference's 187 per-diagonal dispatches at 1080p).
//
// Each workgroup is a single wave that takes MB rows from an atomic ticket
// (so a row's producer always started before its consumer) and walks its row
// left to right.  Before MB (x,y) it waits until row y-1 has published
// progress > x, i.e. MB (x,y-1) (and hence (x-1,y-1)) is reconstructed; the
// left neighbour was produced by this same wave and is carried in LDS.
// Per MB: I16x16 SAD mode decision (DC, V, H, Plane, strict <), residual,
// forward 4x4, intra AC quantisation of positions 1..15, luma DC Hadamard /
// quant / inverse / dequant, reconstruction; chroma DC prediction per
// quadrant with the same chain.  Waits are bounded and report on *err.
#include "bc250_dev.h"
#include "bc250_stages.h"

__global__ __launch_bounds__(64) void h264_i_encode_kernel(
    const uint8_t *__restrict__ cy, const uint8_t *__restrict__ cuv, int cpy, int cpuv,
    uint8_t *__restrict__ oy, uint8_t *__restrict__ ouv, int opy, int opuv, int mbw, int mbh, int qp,
    int num_slices, int16_t *__restrict__ levels, int32_t *__restrict__ dc, uint32_t *__restrict__ nz,
    uint32_t *__restrict__ modes, int *sync /* [0] ticket, [1] err, [2..] progress */, unsigned spin_limit)
{
    __shared__ int s_row;
    __shared__ int res[16][16];   // [block][pos]
    __shared__ int pred[16][16];
    __shared__ int acl[16][16];
    __shared__ int dcin[16], dcv[16];
    __shared__ uint8_t leftY[16];
    __shared__ uint8_t leftC[2][8];
    __shared__ int cres[8][16], cacl[8][16], cdin[8], cdcv[8];
    __shared__ int cpred[2][4];

    const int lane = threadIdx.x;
    int *ticket = sync, *err = sync + 1, *progress = sync + 2;
    if (lane == 0) s_row = atomicAdd(ticket, 1);
    __syncthreads();
    const int y = s_row;
    if (y >= mbh) return;
    const int total = mbw * mbh;
    const int qpc = chroma_qp(qp);
    const int per = qp / 6, rem = qp % 6, fi = (1 << (15 + per)) / 3;
    const int perc = qpc / 6, remc = qpc % 6, fic = (1 << (15 + perc)) / 3;

    // lane -> one 4-pixel row of one luma 4x4 block
    const int lr = lane >> 2, lc0 = (lane & 3) * 4;
    const int lb = (lr >> 2) * 4 + (lane & 3), lk = (lr & 3) * 4;

    for (int x = 0; x < mbw; x++) {
        const int m = y * mbw + x;
        const int X0 = x * 16, Y0 = y * 16, U0 = x * 8, V0 = y * 8;
        if (y > 0) {
            if (lane == 0) wait_at_least(&progress[y - 1], x + 1, err, spin_limit);
            __syncthreads();
            __threadfence();
        }
        const int sl = slice_of(m, total, num_slices);
        const bool te = y > 0, le = x > 0;
        const bool ta = te && slice_of(m - mbw, total, num_slices) == sl;
        const bool la = le && slice_of(m - 1, total, num_slices) == sl;
        const bool ca = te && le && slice_of(m - mbw - 1, total, num_slices) == sl;
        const bool plane_ok = ta && la && ca;

        // neighbours: top row + corner from row y-1 (global), left from LDS
        int top = 128, left = 128, corner = 128;
        if (lane < 16) {
            if (te) top = oy[(size_t)(Y0 - 1) * opy + X0 + lane];
            if (le) left = leftY[lane];
        }
        if (te && le) corner = oy[(size_t)(Y0 - 1) * opy + X0 - 1];
        // broadcast arrays through shuffles
        int sum_t = 0, sum_l = 0;
        for (int i = 0; i < 16; i++) { sum_t += __shfl(top, i, 64); sum_l += __shfl(left, i, 64); }
        const int dcp = (ta && la) ? (sum_t + sum_l + 16) >> 5 : ta ? (sum_t + 8) >> 4 : la ? (sum_l + 8) >> 4 : 128;
        int pa = 0, pb = 0, pc = 0;
        if (plane_ok) {
            int H = 0, V = 0;
            for (int k = 0; k < 8; k++) {
                int tl = 6 - k >= 0 ? __shfl(top, 6 - k, 64) : corner;
                int ll = 6 - k >= 0 ? __shfl(left, 6 - k, 64) : corner;
                H += (k + 1) * (__shfl(top, 8 + k, 64) - tl);
                V += (k + 1) * (__shfl(left, 8 + k, 64) - ll);
            }
            pa = 16 * (__shfl(left, 15, 64) + __shfl(top, 15, 64));
            pb = (5 * H + 32) >> 6;
            pc = (5 * V + 32) >> 6;
        }
        // current samples of this lane
        const uint32_t cw = *(const uint32_t *)(cy + (size_t)(Y0 + lr) * cpy + X0 + lc0);
        const int lft = __shfl(left, lr, 64);
        int tp[4], s[4];
        uint32_t sdc = 0, sv = 0, sh = 0, sp = 0;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            s[i] = (cw >> (8 * i)) & 255;
            tp[i] = __shfl(top, lc0 + i, 64);
            sdc += iabs(s[i] - dcp);
            sv += iabs(s[i] - tp[i]);
            sh += iabs(s[i] - lft);
            if (plane_ok) sp += iabs(s[i] - clip255((pa + pb * (lc0 + i - 7) + pc * (lr - 7) + 16) >> 5));
        }
        sdc = wave_sum_u32(sdc);
        sv = wave_sum_u32(sv);
        sh = wave_sum_u32(sh);
        sp = wave_sum_u32(sp);
        uint32_t best = sdc, mode = 2;
        if (ta && sv < best) { best = sv; mode = 0; }
        if (la && sh < best) { best = sh; mode = 1; }
        if (plane_ok && sp < best) { best = sp; mode = 3; }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            int p = mode == 0 ? tp[i] : mode == 1 ? lft : mode == 3
                    ? clip255((pa + pb * (lc0 + i - 7) + pc * (lr - 7) + 16) >> 5) : dcp;
            pred[lb][lk + i] = p;
            res[lb][lk + i] = s[i] - p;
        }
        if (lane == 0) modes[m] = mode;
        __syncthreads();

        // luma: forward transform + intra AC quant (lanes 0..15, one block each)
        if (lane < 16) {
            int r[16], co[16];
#pragma unroll
            for (int k = 0; k < 16; k++) r[k] = res[lane][k];
            fwd4x4(r, co);
            dcin[lane] = co[0];
            dc[m * 24 + lane] = co[0];
            uint32_t msk = 0, pk[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            acl[lane][0] = 0;
#pragma unroll
            for (int k = 1; k < 16; k++) {
                int l = quant_ac(co[k], k, per, rem, fi);
                acl[lane][k] = l;
                pk[k >> 1] |= (uint32_t)(uint16_t)(int16_t)l << (16 * (k & 1));
                msk |= (l != 0 ? 1u : 0u) << k;
            }
            uint4 *dst = (uint4 *)(levels + ((size_t)m * 24 + lane) * 16);
            dst[0] = make_uint4(pk[0], pk[1], pk[2], pk[3]);
            dst[1] = make_uint4(pk[4], pk[5], pk[6], pk[7]);
            nz[m * 24 + lane] = msk;
        }
        __syncthreads();
        if (lane == 0) {
            int raw[16], l16[16], f[16];
            hadamard4x4(dcin, raw, true);
#pragma unroll
            for (int b = 0; b < 16; b++) l16[b] = quant_dc_k5(raw[b], per, rem);
            hadamard4x4(l16, f, false);
#pragma unroll
            for (int b = 0; b < 16; b++) dcv[b] = dequant_dc_luma(f[b], per, rem);
        }
        __syncthreads();
        if (lane < 16) {
            int d[16], r[16];
            d[0] = dcv[lane];
#pragma unroll
            for (int k = 1; k < 16; k++) d[k] = dequant_ac(acl[lane][k], k, per, rem);
            inv4x4(d, r);
            const int bx = lane & 3, by = lane >> 2;
#pragma unroll
            for (int row = 0; row < 4; row++) {
                uint32_t word = 0;
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    int v = clip255(pred[lane][row * 4 + c] + r[row * 4 + c]);
                    word |= (uint32_t)v << (8 * c);
                    if (bx == 3 && c == 3) leftY[by * 4 + row] = (uint8_t)v;
                }
                *(uint32_t *)(oy + (size_t)(Y0 + by * 4 + row) * opy + X0 + bx * 4) = word;
            }
        }

        // chroma: DC prediction per quadrant (H.264 8.3.4.1-3)
        int ctop = 128, cleft = 128;
        if (lane < 16) {
            const int comp = lane >> 3, i = lane & 7;
            if (ta) ctop = ouv[(size_t)(V0 - 1) * opuv + 2 * (U0 + i) + comp];
            if (la) cleft = leftC[comp][i];
        }
        {
            // every lane sums for component (lane & 1); lanes 0 and 1 keep the result
            const int comp = lane & 1;
            int t0 = 0, t1 = 0, l0 = 0, l1 = 0;
            for (int i = 0; i < 4; i++) {
                t0 += __shfl(ctop, comp * 8 + i, 64);
                t1 += __shfl(ctop, comp * 8 + 4 + i, 64);
                l0 += __shfl(cleft, comp * 8 + i, 64);
                l1 += __shfl(cleft, comp * 8 + 4 + i, 64);
            }
            if (lane < 2) {
                cpred[comp][0] = (ta && la) ? (t0 + l0 + 4) >> 3 : ta ? (t0 + 2) >> 2 : la ? (l0 + 2) >> 2 : 128;
                cpred[comp][1] = ta ? (t1 + 2) >> 2 : la ? (l0 + 2) >> 2 : 128;
                cpred[comp][2] = la ? (l1 + 2) >> 2 : ta ? (t0 + 2) >> 2 : 128;
                cpred[comp][3] = (ta && la) ? (t1 + l1 + 4) >> 3 : ta ? (t1 + 2) >> 2 : la ? (l1 + 2) >> 2 : 128;
            }
        }
        __syncthreads();
        // residual: lane -> one 4-sample row of one chroma block (32 lanes)
        if (lane < 32) {
            const int comp = lane >> 4, sq = (lane >> 2) & 3, row = lane & 3;
            const int u0 = U0 + (sq & 1) * 4, v0 = V0 + (sq >> 1) * 4 + row;
            const uint2 w2 = *(const uint2 *)(cuv + (size_t)v0 * cpuv + 2 * u0);
#pragma unroll
            for (int c = 0; c < 4; c++) {
                uint32_t word = c < 2 ? w2.x : w2.y;
                int sv2 = (word >> (16 * (c & 1) + 8 * comp)) & 255;
                cres[comp * 4 + sq][row * 4 + c] = sv2 - cpred[comp][sq];
            }
        }
        __syncthreads();
        if (lane < 8) {
            const int b = 16 + lane;
            int r[16], co[16];
#pragma unroll
            for (int k = 0; k < 16; k++) r[k] = cres[lane][k];
            fwd4x4(r, co);
            cdin[lane] = co[0];
            dc[m * 24 + b] = co[0];
            uint32_t msk = 0, pk[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            cacl[lane][0] = 0;
#pragma unroll
            for (int k = 1; k < 16; k++) {
                int l = quant_ac(co[k], k, perc, remc, fic);
                cacl[lane][k] = l;
                pk[k >> 1] |= (uint32_t)(uint16_t)(int16_t)l << (16 * (k & 1));
                msk |= (l != 0 ? 1u : 0u) << k;
            }
            uint4 *dst = (uint4 *)(levels + ((size_t)m * 24 + b) * 16);
            dst[0] = make_uint4(pk[0], pk[1], pk[2], pk[3]);
            dst[1] = make_uint4(pk[4], pk[5], pk[6], pk[7]);
            nz[m * 24 + b] = msk;
        }
        __syncthreads();
        if (lane < 2) {
            int din[4], raw[4], l4[4], f4[4];
#pragma unroll
            for (int q = 0; q < 4; q++) din[q] = cdin[lane * 4 + q];
            hadamard2x2(din, raw);
#pragma unroll
            for (int q = 0; q < 4; q++) l4[q] = quant_dc_k5(raw[q], perc, remc);
            hadamard2x2(l4, f4);
#pragma unroll
            for (int q = 0; q < 4; q++) cdcv[lane * 4 + q] = dequant_dc_chroma(f4[q], perc, remc);
        }
        __syncthreads();
        if (lane < 4) {
            const int sq = lane;
            int rr[2][16];
#pragma unroll
            for (int comp = 0; comp < 2; comp++) {
                int d[16];
                d[0] = cdcv[comp * 4 + sq];
#pragma unroll
                for (int k = 1; k < 16; k++) d[k] = dequant_ac(cacl[comp * 4 + sq][k], k, perc, remc);
                inv4x4(d, rr[comp]);
            }
            const int ux = U0 + (sq & 1) * 4, vy = V0 + (sq >> 1) * 4;
#pragma unroll
            for (int row = 0; row < 4; row++) {
                uint32_t lo = 0, hi = 0;
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    int k = row * 4 + c;
                    uint32_t cbv = (uint32_t)clip255(cpred[0][sq] + rr[0][k]);
                    uint32_t crv = (uint32_t)clip255(cpred[1][sq] + rr[1][k]);
                    if (c < 2) lo |= (cbv | (crv << 8)) << (16 * c);
                    else hi |= (cbv | (crv << 8)) << (16 * (c - 2));
                    if ((sq & 1) && c == 3) {
                        leftC[0][(sq >> 1) * 4 + row] = (uint8_t)cbv;
                        leftC[1][(sq >> 1) * 4 + row] = (uint8_t)crv;
                    }
                }
                *(uint2 *)(ouv + (size_t)(vy + row) * opuv + 2 * ux) = make_uint2(lo, hi);
            }
        }
        // publish MB (x,y): every lane's stores are made visible first
        __threadfence();
        __syncthreads();
        if (lane == 0) flag_store_release(&progress[y], x + 1);
    }
}

extern "C" int bc250_launch_i_encode(const uint8_t *cy, const uint8_t *cuv, int cpy, int cpuv, uint8_t *oy,
                                     uint8_t *ouv, int opy, int opuv, int wc, int hc, int qp, int num_slices,
                                     int16_t *levels, int32_t *dc, uint32_t *nz, uint32_t *modes, int *sync,
                                     unsigned spin_limit, void *stream)
{
    if (wc <= 0 || hc <= 0 || (wc & 15) || (hc & 15) || qp < 0 || qp > 51 || (opy & 3) || (opuv & 7) ||
        (cpy & 3) || (cpuv & 7) || num_slices < 1)
        return -1;
    const int mbw = wc / 16, mbh = hc / 16;
    h264_i_encode_kernel<<<mbh, 64, 0, (hipStream_t)stream>>>(cy, cuv, cpy, cpuv, oy, ouv, opy, opuv, mbw, mbh,
                                                               qp, num_slices, levels, dc, nz, modes, sync,
                                                               spin_limit);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
