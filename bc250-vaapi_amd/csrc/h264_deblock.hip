// h264_deblock.hip - H.264 8.7 deblocking (luma + chroma, bS 1..4 incl. the
// strong filter) as ONE persistent wavefront kernel for gfx950.
//
// The standard filters macroblocks in raster order, each MB's vertical
// edges then its horizontal edges, on samples already modified by earlier
// MBs.  MB (x,y) reads/writes samples that MB (x+1,y-1) modified last (its
// left edge touches the right columns of MB (x,y-1)), so the exact order
// is reproduced by a wavefront where row y runs two MBs behind row y-1.
// Each workgroup is one wave that takes a row from an atomic ticket and keeps
// a sliding tile of the previous and current MB (+4 rows above) in LDS:
//   luma tile  [20][32]: rows Y0-4..Y0+15, cols X0-16..X0+15
//   chroma     [2][10][16]: rows V0-2..V0+7, cols U0-8..U0+7
// Step x filters MB x in LDS, writes back MB x-1 (now final for this row)
// and the rows of MB (x,y-1) its top edge changed, then publishes
// progress[y] = x (MBs 0..x-1 final).  Row y+1 starts MB x once
// progress[y] >= x+1, i.e. after MB x+1 of row y has filtered its left edge.
// Waits are bounded and report on *err.
#include "bc250_dev.h"
#include "bc250_stages.h"

namespace {
__constant__ uint8_t kAlpha[52] = {
    0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 4, 4, 5, 6, 7, 8, 9, 10, 12, 13,
    15, 17, 20, 22, 25, 28, 32, 36, 40, 45, 50, 56, 63, 71, 80, 90, 101, 113, 127, 144, 162, 182,
    203, 226, 255, 255};
__constant__ uint8_t kBeta[52] = {
    0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4,
    6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13, 14, 14, 15, 15, 16, 16, 17, 17, 18, 18};
__constant__ uint8_t kTc0[52][3] = {
    {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0},
    {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0},
    {0, 0, 0}, {0, 0, 1}, {0, 0, 1}, {0, 0, 1}, {0, 0, 1}, {0, 1, 1}, {0, 1, 1}, {1, 1, 1},
    {1, 1, 1}, {1, 1, 1}, {1, 1, 1}, {1, 1, 2}, {1, 1, 2}, {1, 1, 2}, {1, 1, 2}, {1, 2, 3},
    {1, 2, 3}, {2, 2, 3}, {2, 2, 4}, {2, 3, 4}, {2, 3, 4}, {3, 3, 5}, {3, 4, 6}, {3, 4, 6},
    {4, 5, 7}, {4, 5, 8}, {4, 6, 9}, {5, 7, 10}, {6, 8, 11}, {6, 8, 13}, {7, 10, 14}, {8, 11, 16},
    {9, 12, 18}, {10, 13, 20}, {11, 15, 23}, {13, 17, 25}};

constexpr int LP = 32;  // luma tile pitch
constexpr int CP = 16;  // chroma tile pitch

__device__ __forceinline__ void luma_line(uint8_t *q, int d, int bs, int qp)
{
    const int alpha = kAlpha[qp], beta = kBeta[qp];
    int p0 = q[-d], p1 = q[-2 * d], p2 = q[-3 * d], p3 = q[-4 * d];
    int q0 = q[0], q1 = q[d], q2 = q[2 * d], q3 = q[3 * d];
    if (!(iabs(p0 - q0) < alpha && iabs(p1 - p0) < beta && iabs(q1 - q0) < beta)) return;
    int ap = iabs(p2 - p0), aq = iabs(q2 - q0);
    if (bs < 4) {
        int tc0 = kTc0[qp][bs - 1];
        int tc = tc0 + (ap < beta) + (aq < beta);
        int delta = clampi((((q0 - p0) << 2) + (p1 - q1) + 4) >> 3, -tc, tc);
        q[-d] = (uint8_t)clip255(p0 + delta);
        q[0] = (uint8_t)clip255(q0 - delta);
        if (ap < beta) q[-2 * d] = (uint8_t)(p1 + clampi((p2 + ((p0 + q0 + 1) >> 1) - (p1 << 1)) >> 1, -tc0, tc0));
        if (aq < beta) q[d] = (uint8_t)(q1 + clampi((q2 + ((p0 + q0 + 1) >> 1) - (q1 << 1)) >> 1, -tc0, tc0));
    } else {
        const bool narrow = iabs(p0 - q0) < ((alpha >> 2) + 2);
        if (ap < beta && narrow) {
            q[-d] = (uint8_t)((p2 + 2 * p1 + 2 * p0 + 2 * q0 + q1 + 4) >> 3);
            q[-2 * d] = (uint8_t)((p2 + p1 + p0 + q0 + 2) >> 2);
            q[-3 * d] = (uint8_t)((2 * p3 + 3 * p2 + p1 + p0 + q0 + 4) >> 3);
        } else {
            q[-d] = (uint8_t)((2 * p1 + p0 + q1 + 2) >> 2);
        }
        if (aq < beta && narrow) {
            q[0] = (uint8_t)((p1 + 2 * p0 + 2 * q0 + 2 * q1 + q2 + 4) >> 3);
            q[d] = (uint8_t)((p0 + q0 + q1 + q2 + 2) >> 2);
            q[2 * d] = (uint8_t)((2 * q3 + 3 * q2 + q1 + q0 + p0 + 4) >> 3);
        } else {
            q[0] = (uint8_t)((2 * q1 + q0 + p1 + 2) >> 2);
        }
    }
}

__device__ __forceinline__ void chroma_line(uint8_t *q, int d, int bs, int qpc)
{
    const int alpha = kAlpha[qpc], beta = kBeta[qpc];
    int p0 = q[-d], p1 = q[-2 * d], q0 = q[0], q1 = q[d];
    if (!(iabs(p0 - q0) < alpha && iabs(p1 - p0) < beta && iabs(q1 - q0) < beta)) return;
    if (bs < 4) {
        int tc = kTc0[qpc][bs - 1] + 1;
        int delta = clampi((((q0 - p0) << 2) + (p1 - q1) + 4) >> 3, -tc, tc);
        q[-d] = (uint8_t)clip255(p0 + delta);
        q[0] = (uint8_t)clip255(q0 - delta);
    } else {
        q[-d] = (uint8_t)((2 * p1 + p0 + q1 + 2) >> 2);
        q[0] = (uint8_t)((2 * q1 + q0 + p1 + 2) >> 2);
    }
}
}  // namespace

__global__ __launch_bounds__(64) void h264_deblock_kernel(
    uint8_t *__restrict__ y_, uint8_t *__restrict__ uv, int py, int puv, int mbw, int mbh, int qp,
    int is_intra, const uint32_t *__restrict__ nz, const uint4 *__restrict__ mvs, int num_slices, int idc,
    int *sync /* [0] ticket, [1] err, [2..] progress */, unsigned spin_limit)
{
    __shared__ uint8_t L[20 * LP];
    __shared__ uint8_t C[2][10 * CP];
    __shared__ int bsv[2][4][4];  // [dir][edge][segment]
    __shared__ int s_row;

    const int lane = threadIdx.x;
    int *ticket = sync, *err = sync + 1, *progress = sync + 2;
    if (lane == 0) s_row = atomicAdd(ticket, 1);
    __syncthreads();
    const int yrow = s_row;
    if (yrow >= mbh) return;
    const int total = mbw * mbh;
    const int qpc = chroma_qp(qp);
    const int Y0 = yrow * 16, V0 = yrow * 8;

    for (int x = 0; x <= mbw; x++) {
        const int X0 = x * 16, U0 = x * 8;
        if (x < mbw) {
            const int m = yrow * mbw + x;
            if (yrow > 0) {
                if (lane == 0) wait_at_least(&progress[yrow - 1], x + 1, err, spin_limit);
                __syncthreads();
                __threadfence();
            }
            const int sl = slice_of(m, total, num_slices);
            const bool skip_left = x == 0 || (idc == 2 && slice_of(m - 1, total, num_slices) != sl);
            const bool skip_top = yrow == 0 || (idc == 2 && slice_of(m - mbw, total, num_slices) != sl);

            // ---- load MB x body (+ 4 rows / 2 chroma rows above)
            {
                const int r = lane >> 2, c = (lane & 3) * 4;
                *(uint32_t *)&L[(4 + r) * LP + 16 + c] = *(const uint32_t *)(y_ + (size_t)(Y0 + r) * py + X0 + c);
                if (lane < 16 && yrow > 0)
                    *(uint32_t *)&L[r * LP + 16 + c] = *(const uint32_t *)(y_ + (size_t)(Y0 - 4 + r) * py + X0 + c);
                if (lane < 40) {
                    // chroma: 8 body rows + 2 rows above, 4 dwords (2 Cb/Cr pairs each) per row
                    const int cr = lane >> 2, cc = (lane & 3) * 2;  // tile row 0..9, chroma col 0..7 step 2
                    if (cr >= 2 || yrow > 0) {
                        uint32_t v = *(const uint32_t *)(uv + (size_t)(V0 - 2 + cr) * puv + 2 * (U0 + cc));
                        C[0][cr * CP + 8 + cc] = (uint8_t)v;
                        C[1][cr * CP + 8 + cc] = (uint8_t)(v >> 8);
                        C[0][cr * CP + 9 + cc] = (uint8_t)(v >> 16);
                        C[1][cr * CP + 9 + cc] = (uint8_t)(v >> 24);
                    }
                }
            }
            // ---- boundary strengths (H.264 8.7.2.1)
            if (lane < 32) {
                const int dir = lane >> 4, e = (lane >> 2) & 3, i = lane & 3;
                int bs = 0;
                if (is_intra) {
                    bs = e == 0 ? 4 : 3;
                } else {
                    const int pm = e ? m : (dir == 0 ? m - 1 : m - mbw);
                    const int bq = dir == 0 ? i * 4 + e : e * 4 + i;
                    const int bp = dir == 0 ? (e ? i * 4 + e - 1 : i * 4 + 3) : (e ? (e - 1) * 4 + i : 12 + i);
                    const bool edge_exists = e != 0 || (dir == 0 ? x > 0 : yrow > 0);
                    if (edge_exists) {
                        if (nz[pm * 24 + bp] || nz[m * 24 + bq]) bs = 2;
                        else {
                            uint4 a = mvs[pm], b = mvs[m];
                            if (iabs((int)a.x - (int)b.x) >= 4 || iabs((int)a.y - (int)b.y) >= 4) bs = 1;
                        }
                    }
                }
                bsv[dir][e][i] = bs;
            }
            __syncthreads();

            // ---- vertical edges, left to right
            for (int e = 0; e < 4; e++) {
                if (!(e == 0 && skip_left)) {
                    if (lane < 16) {
                        int bs = bsv[0][e][lane >> 2];
                        if (bs) luma_line(&L[(4 + lane) * LP + 16 + 4 * e], 1, bs, qp);
                    } else if (lane < 32 && (e == 0 || e == 2)) {
                        const int comp = (lane - 16) >> 3, k = (lane - 16) & 7;
                        int bs = bsv[0][e][k >> 1];
                        if (bs) chroma_line(&C[comp][(2 + k) * CP + 8 + 2 * e], 1, bs, qpc);
                    }
                }
                __syncthreads();
            }
            // ---- horizontal edges, top to bottom
            for (int e = 0; e < 4; e++) {
                if (!(e == 0 && skip_top)) {
                    if (lane < 16) {
                        int bs = bsv[1][e][lane >> 2];
                        if (bs) luma_line(&L[(4 + 4 * e) * LP + 16 + lane], LP, bs, qp);
                    } else if (lane < 32 && (e == 0 || e == 2)) {
                        const int comp = (lane - 16) >> 3, k = (lane - 16) & 7;
                        int bs = bsv[1][e][k >> 1];
                        if (bs) chroma_line(&C[comp][(2 + 2 * e) * CP + 8 + k], CP, bs, qpc);
                    }
                }
                __syncthreads();
            }
            // ---- rows of MB (x,y-1) changed by the top edge
            if (!skip_top && lane < 12) {
                const int r = 1 + (lane >> 2), c = (lane & 3) * 4;
                *(uint32_t *)(y_ + (size_t)(Y0 - 4 + r) * py + X0 + c) = *(const uint32_t *)&L[r * LP + 16 + c];
            }
            if (!skip_top && lane >= 16 && lane < 20) {
                const int cc = (lane - 16) * 2;
                uint32_t v = (uint32_t)C[0][CP + 8 + cc] | ((uint32_t)C[1][CP + 8 + cc] << 8) |
                             ((uint32_t)C[0][CP + 9 + cc] << 16) | ((uint32_t)C[1][CP + 9 + cc] << 24);
                *(uint32_t *)(uv + (size_t)(V0 - 1) * puv + 2 * (U0 + cc)) = v;
            }
        }
        // ---- write back MB x-1 (final for this row)
        if (x > 0) {
            const int r = lane >> 2, c = (lane & 3) * 4;
            *(uint32_t *)(y_ + (size_t)(Y0 + r) * py + X0 - 16 + c) = *(const uint32_t *)&L[(4 + r) * LP + c];
            if (lane < 32) {
                const int cr = lane >> 2, cc = (lane & 3) * 2;
                uint32_t v = (uint32_t)C[0][(2 + cr) * CP + cc] | ((uint32_t)C[1][(2 + cr) * CP + cc] << 8) |
                             ((uint32_t)C[0][(2 + cr) * CP + 1 + cc] << 16) |
                             ((uint32_t)C[1][(2 + cr) * CP + 1 + cc] << 24);
                *(uint32_t *)(uv + (size_t)(V0 + cr) * puv + 2 * (U0 - 8 + cc)) = v;
            }
        }
        __syncthreads();
        // ---- slide: MB x becomes the left neighbour
        {
            const int r = lane >> 2, c = (lane & 3) * 4;
            *(uint32_t *)&L[(4 + r) * LP + c] = *(const uint32_t *)&L[(4 + r) * LP + 16 + c];
            if (lane < 40) {
                const int cr = lane >> 2, cc = (lane & 3) * 2;
                C[0][cr * CP + cc] = C[0][cr * CP + 8 + cc];
                C[0][cr * CP + cc + 1] = C[0][cr * CP + 9 + cc];
                C[1][cr * CP + cc] = C[1][cr * CP + 8 + cc];
                C[1][cr * CP + cc + 1] = C[1][cr * CP + 9 + cc];
            }
        }
        __threadfence();
        __syncthreads();
        if (lane == 0) flag_store_release(&progress[yrow], x);
    }
}

extern "C" int bc250_launch_deblock(uint8_t *y, uint8_t *uv, int py, int puv, int wc, int hc, int qp,
                                    int is_intra, const uint32_t *nz, const uint32_t *mvs, int num_slices, int idc,
                                    int *sync, unsigned spin_limit, void *stream)
{
    if (wc <= 0 || hc <= 0 || (wc & 15) || (hc & 15) || qp < 0 || qp > 51 || (py & 3) || (puv & 3) ||
        num_slices < 1 || idc < 0 || idc > 2)
        return -1;
    if (idc == 1) return 0;
    const int mbw = wc / 16, mbh = hc / 16;
    h264_deblock_kernel<<<mbh, 64, 0, (hipStream_t)stream>>>(y, uv, py, puv, mbw, mbh, qp, is_intra, nz,
                                                              (const uint4 *)mvs, num_slices, idc, sync,
                                                              spin_limit);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
