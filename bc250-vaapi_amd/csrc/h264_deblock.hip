// h264_deblock.hip - H.264 8.7 deblocking (luma + chroma, bS 1..4 incl. the
// strong filter) for gfx950: a fully parallel boundary-strength kernel, then
// ONE persistent wavefront kernel that filters in the standard's order.
//
// The standard filters macroblocks in raster order, each MB's vertical
// edges then its horizontal edges, on samples already modified by earlier
// MBs.  MB (x,y) reads samples that MB (x+1,y-1) modified last (its left
// edge touches the right columns of MB (x,y-1)), so the exact order is a
// wavefront in which row y runs two MBs behind row y-1.  Luma and chroma
// never read each other, so they run as two independent wavefronts.
//
// Work layout: a workgroup is 2*ROWS waves for ROWS MB rows (row group from
// an atomic ticket), one wave per SIMD: the first ROWS waves filter the luma
// of rows ROWS*g.., the others their chroma.  A wave keeps a sliding LDS tile (previous MB, current MB,
// rows above).  Per MB it makes two passes: one lane per line through the
// four vertical edges in registers, then one lane per column through the
// four horizontal edges.  When MB x-1 is final for its row (after MB x
// filtered its left edge) the wave
//   - hands the bottom rows that the next row's top edge reads and partly
//     rewrites (luma 12..15, chroma 6..7) to the wave below: through an LDS
//     ring inside the workgroup, or to the next workgroup as system-coherent
//     64-bit records {tag = (epoch << 16) | (x + 1), 4 samples} that the
//     reader polls until every tag matches, so the writer never waits for
//     its stores; the launch epoch makes the previous frame's record at the
//     same place a mismatch,
//   - stores the other rows to the frame; the wave below stores the rows it
//     rewrote (luma 13..15, chroma 7), so each frame byte has one writer.
// The records are also cleared, by system-scope stores of the
// boundary-strength kernel, before each frame.  No cache writeback or
// invalidate is issued.  Waits are bounded and report on *err.
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

#ifndef BC250_DBK_ROWS
#define BC250_DBK_ROWS 2
#endif
constexpr int ROWS = BC250_DBK_ROWS;  // MB rows per workgroup (a power of two)
constexpr int LP = 32;         // luma tile pitch: cols 0..15 MB x-1, 16..31 MB x
constexpr int CP = 16;         // chroma tile pitch: cols 0..7 MB x-1, 8..15 MB x
constexpr int RING = 16;       // LDS hand-off slots per producing wave
constexpr int LHAND = 16;      // luma hand-off dwords: rows 12..15, 16 bytes each
constexpr int CHAND = 8;       // chroma hand-off dwords: Cb rows 6..7, Cr rows 6..7 (8 bytes each)
constexpr int GHAND = LHAND + CHAND;

// One luma line across an edge (8.7.2.3 / 8.7.2.4): v[0..7] = p3 p2 p1 p0
// q0 q1 q2 q3, filtered in place.  Written without branches: every candidate
// output is computed, conditions are combined with '&', and each '?:' picks
// between two ready values, so it lowers to v_cndmask.  (A branchy form makes
// the compiler copy the whole line at every join.)  can4 = false promises bS
// < 4 (a compile-time constant at every call), which drops the strong form.
__device__ __forceinline__ void luma_filter(int *v, int bs, int alpha, int beta, int tc0, bool can4)
{
    const int p3 = v[0], p2 = v[1], p1 = v[2], p0 = v[3], q0 = v[4], q1 = v[5], q2 = v[6], q3 = v[7];
    const int d0 = iabs(p0 - q0);
    const bool f = (bs != 0) & (d0 < alpha) & (iabs(p1 - p0) < beta) & (iabs(q1 - q0) < beta);
    const bool apb = iabs(p2 - p0) < beta, aqb = iabs(q2 - q0) < beta;
    const bool st = can4 && bs >= 4;
    // bS < 4
    const int tc = tc0 + (int)apb + (int)aqb;
    const int delta = clampi((((q0 - p0) << 2) + (p1 - q1) + 4) >> 3, -tc, tc);
    const int avg = (p0 + q0 + 1) >> 1;
    const int w_p0 = clip255(p0 + delta), w_q0 = clip255(q0 - delta);
    const int c_p1 = p1 + clampi((p2 + avg - (p1 << 1)) >> 1, -tc0, tc0);
    const int c_q1 = q1 + clampi((q2 + avg - (q1 << 1)) >> 1, -tc0, tc0);
    // bS == 4
    const bool narrow = d0 < ((alpha >> 2) + 2);
    const bool sp = apb & narrow, sq = aqb & narrow;
    const int sp_p0 = (p2 + 2 * p1 + 2 * p0 + 2 * q0 + q1 + 4) >> 3, wp_p0 = (2 * p1 + p0 + q1 + 2) >> 2;
    const int sp_p1 = (p2 + p1 + p0 + q0 + 2) >> 2;
    const int sp_p2 = (2 * p3 + 3 * p2 + p1 + p0 + q0 + 4) >> 3;
    const int sq_q0 = (p1 + 2 * p0 + 2 * q0 + 2 * q1 + q2 + 4) >> 3, wq_q0 = (2 * q1 + q0 + p1 + 2) >> 2;
    const int sq_q1 = (p0 + q0 + q1 + q2 + 2) >> 2;
    const int sq_q2 = (2 * q3 + 3 * q2 + q1 + q0 + p0 + 4) >> 3;
    const int s_p0 = sp ? sp_p0 : wp_p0, s_q0 = sq ? sq_q0 : wq_q0;
    const int n_p1 = apb ? c_p1 : p1, n_q1 = aqb ? c_q1 : q1;
    const int o_p0 = st ? s_p0 : w_p0, o_q0 = st ? s_q0 : w_q0;
    const int o_p1 = st ? (sp ? sp_p1 : p1) : n_p1, o_q1 = st ? (sq ? sq_q1 : q1) : n_q1;
    const int o_p2 = (st & sp) ? sp_p2 : p2, o_q2 = (st & sq) ? sq_q2 : q2;
    v[1] = f ? o_p2 : p2;
    v[2] = f ? o_p1 : p1;
    v[3] = f ? o_p0 : p0;
    v[4] = f ? o_q0 : q0;
    v[5] = f ? o_q1 : q1;
    v[6] = f ? o_q2 : q2;
}

// One chroma line across an edge (chromaEdgeFlag = 1): v[0..3] = p1 p0 q0
// q1, filtered in place, branch-free like luma_filter (same can4).
__device__ __forceinline__ void chroma_filter(int *v, int bs, int alpha, int beta, int tc0, bool can4)
{
    const int p1 = v[0], p0 = v[1], q0 = v[2], q1 = v[3];
    const bool f = (bs != 0) & (iabs(p0 - q0) < alpha) & (iabs(p1 - p0) < beta) & (iabs(q1 - q0) < beta);
    const int tc = tc0 + 1;
    const int delta = clampi((((q0 - p0) << 2) + (p1 - q1) + 4) >> 3, -tc, tc);
    const bool st = can4 && bs >= 4;
    const int s_p0 = (2 * p1 + p0 + q1 + 2) >> 2, s_q0 = (2 * q1 + q0 + p1 + 2) >> 2;
    const int w_p0 = clip255(p0 + delta), w_q0 = clip255(q0 - delta);
    const int o_p0 = st ? s_p0 : w_p0, o_q0 = st ? s_q0 : w_q0;
    v[1] = f ? o_p0 : p0;
    v[2] = f ? o_q0 : q0;
}

// bS words: per direction two words, edges 0..1 then 2..3, 4 bits per (edge, segment).
__device__ __forceinline__ int bs_get(uint32_t lo, uint32_t hi, int e, int i)
{
    return ((e < 2 ? lo : hi) >> (4 * ((e & 1) * 4 + i))) & 15;
}

struct alignas(16) LumaLds {
    uint8_t T[20 * LP];  // rows 0..3 above, 4..19 body
};
struct alignas(16) ChromaLds {
    uint8_t T[2][10 * CP];  // [Cb, Cr] rows 0..1 above, 2..9 body
};
}  // namespace

// ---------------------------------------------------------------------------
// Boundary strengths (H.264 8.7.2.1), every MB in parallel.  bs[m] = {vertical
// edges 0..1, vertical edges 2..3, horizontal edges 0..1, horizontal edges
// 2..3}, 4 bits per (edge e, segment i) at bit 4*(4*(e&1)+i) of its word.
// Edges that are not filtered (picture border, slice border with idc 2) get 0.
__global__ void h264_bs_kernel(int mbw, int mbh, int is_intra, const uint32_t *__restrict__ nz,
                               const uint4 *__restrict__ mvs, int num_slices, int idc, uint4 *__restrict__ bs,
                               uint64_t *__restrict__ ghand)
{
    const int m = blockIdx.x * 8 + (threadIdx.x >> 5);
    const int j = threadIdx.x & 31;
    const int nmb = mbw * mbh;
    if (m >= nmb) return;
    const int x = m % mbw, y = m / mbw;
    if (j < GHAND) st_sys64(ghand + (size_t)m * GHAND + j, 0);  // no valid tag
    const int dir = j >> 4, e = (j >> 2) & 3, i = j & 3;
    int b = 0;
    bool exists = true;
    if (e == 0) {
        const int nb = dir == 0 ? m - 1 : m - mbw;
        exists = dir == 0 ? x > 0 : y > 0;
        if (exists && idc == 2 && slice_of(nb, nmb, num_slices) != slice_of(m, nmb, num_slices)) exists = false;
    }
    if (exists) {
        if (is_intra) {
            b = e == 0 ? 4 : 3;
        } else {
            const int pm = e ? m : (dir == 0 ? m - 1 : m - mbw);
            const int bq = dir == 0 ? i * 4 + e : e * 4 + i;
            const int bp = dir == 0 ? (e ? i * 4 + e - 1 : i * 4 + 3) : (e ? (e - 1) * 4 + i : 12 + i);
            if (nz[pm * 24 + bp] || nz[m * 24 + bq]) {
                b = 2;
            } else {
                const uint4 a = mvs[pm], c = mvs[m];
                if (iabs((int)a.x - (int)c.x) >= 4 || iabs((int)a.y - (int)c.y) >= 4) b = 1;
            }
        }
    }
    // pack 8 nibbles per word with a reduction over the 8 lanes of (dir, e>>1)
    uint32_t word = (uint32_t)b << (4 * ((e & 1) * 4 + i));
#pragma unroll
    for (int o = 4; o > 0; o >>= 1) word |= __shfl_xor(word, o, 64);
    uint32_t *out = (uint32_t *)&bs[m];
    if ((j & 7) == 0) out[j >> 3] = word;
}

// ---------------------------------------------------------------------------
// INTRA: bS is 4 on MB edges (edge 0) and 3 inside; otherwise bS <= 2, so
// each edge's filter form is known at compile time.
template <bool INTRA>
__global__ __launch_bounds__(128 * ROWS) void h264_deblock_kernel(
    uint8_t *__restrict__ y_, uint8_t *__restrict__ uv, int py, int puv, int mbw, int mbh, int qp,
    const uint4 *__restrict__ bsw, int *sync /* [0] ticket, [1] err, [2] epoch */,
    uint64_t *ghand /* [mbh][mbw][GHAND] {tag, value} */, unsigned spin_limit)
{
    __shared__ LumaLds ll[ROWS];
    __shared__ ChromaLds cl[ROWS];
    __shared__ uint32_t lring[ROWS - 1][RING][LHAND];
    __shared__ uint32_t cring[ROWS - 1][RING][CHAND];
    __shared__ int prog[2][ROWS], cons[2][ROWS];
    __shared__ int s_group;

    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const int chroma = wv >= ROWS, r4 = wv & (ROWS - 1);
    int *err = sync + 1;
    const uint32_t ep = wf_epoch((const uint32_t *)sync + 2);
    if (tid == 0) s_group = atomicAdd(sync, 1);
    if (tid < ROWS) prog[0][tid] = prog[1][tid] = cons[0][tid] = cons[1][tid] = 0;
    __syncthreads();
    const int yrow = s_group * ROWS + r4;
    if (yrow >= mbh) return;
    volatile int *vprog = prog[chroma], *vcons = cons[chroma];
    const bool last_row = yrow == mbh - 1;
    const bool lds_below = r4 < ROWS - 1 && !last_row;
    const bool lds_above = r4 > 0;
    const int qpx = chroma ? chroma_qp(qp) : qp;
    const int alpha = kAlpha[qpx], beta = kBeta[qpx];
    // tC0 for bS 1..3 in registers: a per-lane table lookup would be a vector
    // memory load, and its wait would also wait for this wave's frame stores.
    const int tc1 = kTc0[qpx][0], tc2 = kTc0[qpx][1], tc3 = kTc0[qpx][2];
    auto tc0_of = [&](int b) { return (b == 1 ? tc1 : 0) | (b == 2 ? tc2 : 0) | (b == 3 ? tc3 : 0); };
    const int Y0 = yrow * 16, V0 = yrow * 8;
    const int nhand = chroma ? CHAND : LHAND;
    uint64_t *gh = ghand + (chroma ? LHAND : 0);
    uint8_t *LT = ll[r4].T;
    uint8_t(*CT)[10 * CP] = cl[r4].T;

    // one-MB-ahead prefetch: body (luma: 1 dword per lane; chroma: lanes 0..31) and bS words
    uint32_t pre = 0;
    uint4 pbs = make_uint4(0, 0, 0, 0);
    auto prefetch = [&](int x) {
        if (!chroma) pre = *(const uint32_t *)(y_ + (size_t)(Y0 + (lane >> 2)) * py + x * 16 + (lane & 3) * 4);
        else if (lane < 32) pre = *(const uint32_t *)(uv + (size_t)(V0 + (lane >> 2)) * puv + x * 16 + (lane & 3) * 4);
        pbs = bsw[yrow * mbw + x];
    };
    prefetch(0);

    for (int x = 0; x <= mbw; x++) {
        const int X0 = x * 16;
        const uint4 bs = pbs;
        if (x < mbw) {
            // ---- body of MB x into the tile
            if (!chroma) {
                *(uint32_t *)&LT[(4 + (lane >> 2)) * LP + 16 + (lane & 3) * 4] = pre;
            } else if (lane < 32) {
                const int r = lane >> 2, c = (lane & 3) * 2;
                CT[0][(2 + r) * CP + 8 + c] = (uint8_t)pre;
                CT[1][(2 + r) * CP + 8 + c] = (uint8_t)(pre >> 8);
                CT[0][(2 + r) * CP + 9 + c] = (uint8_t)(pre >> 16);
                CT[1][(2 + r) * CP + 9 + c] = (uint8_t)(pre >> 24);
            }
            // ---- rows above: MB (x, y-1) must be final; take its hand-off
            if (yrow > 0) {
                uint32_t h = 0;
                if (lds_above) {
                    unsigned n = 0;
                    while (vprog[r4 - 1] < x + 1) {
                        __builtin_amdgcn_s_sleep(1);
                        if (++n > spin_limit) { atomicOr(err, 1); break; }
                    }
                    wave_lds_sync();
                    if (lane < nhand) h = chroma ? cring[r4 - 1][x % RING][lane] : lring[r4 - 1][x % RING][lane];
                    wave_lds_sync();
                    if (lane == 0) vcons[r4] = x + 1;
                } else {
                    const uint64_t *src = gh + ((size_t)(yrow - 1) * mbw + x) * GHAND + lane;
                    const uint32_t tag = wf_tag(ep, x + 1);
                    unsigned n = 0;
                    uint64_t q = 0;
                    for (;;) {
                        bool ok = true;
                        if (lane < nhand) {
                            q = ld_sys64(src);
                            ok = (uint32_t)(q >> 32) == tag;
                        }
                        if (__all(ok)) break;
                        __builtin_amdgcn_s_sleep(1);
                        if (++n > spin_limit) {
                            if (lane == 0) atomicOr(err, 1);
                            break;
                        }
                    }
                    h = (uint32_t)q;
                }
                if (!chroma) {
                    if (lane < LHAND) *(uint32_t *)&LT[(lane >> 2) * LP + 16 + (lane & 3) * 4] = h;
                } else if (lane < CHAND) {
                    const int comp = lane >> 2, r = (lane >> 1) & 1, c = (lane & 1) * 4;
                    *(uint32_t *)&CT[comp][r * CP + 8 + c] = h;
                }
            }
            if (x + 1 < mbw) prefetch(x + 1);
            wave_lds_sync();

            if (!chroma) {
                // ---- vertical edges: lane = line, tile cols 12..31 in registers
                if (lane < 16) {
                    uint8_t *row = &LT[(4 + lane) * LP + 12];
                    int v[20];
#pragma unroll
                    for (int d = 0; d < 5; d++) {
                        const uint32_t wd = *(const uint32_t *)(row + 4 * d);
#pragma unroll
                        for (int k = 0; k < 4; k++) v[4 * d + k] = (wd >> (8 * k)) & 255;
                    }
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        const int b = bs_get(bs.x, bs.y, e, lane >> 2);
                        if (__any(b != 0)) luma_filter(&v[4 * e], b, alpha, beta, tc0_of(b), INTRA && e == 0);
                    }
#pragma unroll
                    for (int d = 0; d < 5; d++)
                        *(uint32_t *)(row + 4 * d) = (uint32_t)v[4 * d] | ((uint32_t)v[4 * d + 1] << 8) |
                                                     ((uint32_t)v[4 * d + 2] << 16) | ((uint32_t)v[4 * d + 3] << 24);
                }
                wave_lds_sync();
                // ---- horizontal edges: lane = column, tile rows 0..19 in registers
                if (lane < 16) {
                    uint8_t *col = &LT[16 + lane];
                    int v[20];
#pragma unroll
                    for (int k = 0; k < 20; k++) v[k] = col[k * LP];
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        const int b = bs_get(bs.z, bs.w, e, lane >> 2);
                        if (__any(b != 0)) luma_filter(&v[4 * e], b, alpha, beta, tc0_of(b), INTRA && e == 0);
                    }
#pragma unroll
                    for (int k = 1; k < 19; k++) col[k * LP] = (uint8_t)v[k];
                }
            } else {
                // ---- chroma: lane = (component, line), tile cols 4..15
                if (lane < 16) {
                    const int comp = lane >> 3, k = lane & 7;
                    uint8_t *row = &CT[comp][(2 + k) * CP + 4];
                    int v[12];
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        const uint32_t wd = *(const uint32_t *)(row + 4 * d);
#pragma unroll
                        for (int q = 0; q < 4; q++) v[4 * d + q] = (wd >> (8 * q)) & 255;
                    }
#pragma unroll
                    for (int e = 0; e < 4; e += 2) {
                        const int b = bs_get(bs.x, bs.y, e, k >> 1);
                        chroma_filter(&v[2 + 2 * e], b, alpha, beta, tc0_of(b), INTRA && e == 0);
                    }
#pragma unroll
                    for (int d = 0; d < 3; d++)
                        *(uint32_t *)(row + 4 * d) = (uint32_t)v[4 * d] | ((uint32_t)v[4 * d + 1] << 8) |
                                                     ((uint32_t)v[4 * d + 2] << 16) | ((uint32_t)v[4 * d + 3] << 24);
                }
                wave_lds_sync();
                if (lane < 16) {
                    const int comp = lane >> 3, k = lane & 7;
                    uint8_t *col = &CT[comp][8 + k];
                    int v[10];
#pragma unroll
                    for (int q = 0; q < 10; q++) v[q] = col[q * CP];
#pragma unroll
                    for (int e = 0; e < 4; e += 2) {
                        const int b = bs_get(bs.z, bs.w, e, k >> 1);
                        chroma_filter(&v[2 * e], b, alpha, beta, tc0_of(b), INTRA && e == 0);
                    }
#pragma unroll
                    for (int q = 1; q < 9; q++) col[q * CP] = (uint8_t)v[q];
                }
            }
            wave_lds_sync();
        }
        // ---- MB x-1 final for this row: hand its bottom rows down first
        if (x > 0 && !last_row) {
            uint32_t h = 0;
            if (!chroma) {
                if (lane < LHAND) h = *(const uint32_t *)&LT[(16 + (lane >> 2)) * LP + (lane & 3) * 4];
            } else if (lane < CHAND) {
                const int comp = lane >> 2, r = (lane >> 1) & 1, c = (lane & 1) * 4;
                h = *(const uint32_t *)&CT[comp][(8 + r) * CP + c];
            }
            if (lds_below) {
                unsigned n = 0;
                while (vcons[r4 + 1] < x - RING) {
                    __builtin_amdgcn_s_sleep(1);
                    if (++n > spin_limit) { atomicOr(err, 2); break; }
                }
                if (lane < nhand) {
                    if (chroma) cring[r4][(x - 1) % RING][lane] = h;
                    else lring[r4][(x - 1) % RING][lane] = h;
                }
                wave_lds_sync();
                if (lane == 0) vprog[r4] = x;
            } else if (lane < nhand) {
                st_sys64(gh + ((size_t)yrow * mbw + x - 1) * GHAND + lane, ((uint64_t)wf_tag(ep, x) << 32) | h);
            }
        }
        // ---- frame stores: MB x-1 (rows this wave owns) and the rewritten rows of MB (x, y-1)
        if (!chroma) {
            if (x > 0) {
                const int r = lane >> 2, c = (lane & 3) * 4;
                if (r < (last_row ? 16 : 13))
                    *(uint32_t *)(y_ + (size_t)(Y0 + r) * py + X0 - 16 + c) = *(const uint32_t *)&LT[(4 + r) * LP + c];
            }
            if (x < mbw && yrow > 0 && lane < 12) {
                const int r = 1 + (lane >> 2), c = (lane & 3) * 4;
                *(uint32_t *)(y_ + (size_t)(Y0 - 4 + r) * py + X0 + c) = *(const uint32_t *)&LT[r * LP + 16 + c];
            }
        } else {
            if (x > 0 && lane < 32) {
                const int r = lane >> 2, c = (lane & 3) * 2;
                if (r < (last_row ? 8 : 7)) {
                    const uint32_t v = (uint32_t)CT[0][(2 + r) * CP + c] | ((uint32_t)CT[1][(2 + r) * CP + c] << 8) |
                                       ((uint32_t)CT[0][(2 + r) * CP + 1 + c] << 16) |
                                       ((uint32_t)CT[1][(2 + r) * CP + 1 + c] << 24);
                    *(uint32_t *)(uv + (size_t)(V0 + r) * puv + X0 - 16 + 2 * c) = v;
                }
            }
            if (x < mbw && yrow > 0 && lane >= 32 && lane < 36) {
                const int c = (lane - 32) * 2;
                const uint32_t v = (uint32_t)CT[0][CP + 8 + c] | ((uint32_t)CT[1][CP + 8 + c] << 8) |
                                   ((uint32_t)CT[0][CP + 9 + c] << 16) | ((uint32_t)CT[1][CP + 9 + c] << 24);
                *(uint32_t *)(uv + (size_t)(V0 - 1) * puv + X0 + 2 * c) = v;
            }
        }
        wave_lds_sync();
        // ---- slide: MB x becomes the left neighbour
        if (!chroma) {
            const int r = lane >> 2, c = (lane & 3) * 4;
            *(uint32_t *)&LT[(4 + r) * LP + c] = *(const uint32_t *)&LT[(4 + r) * LP + 16 + c];
        } else if (lane < 20) {
            const int cr = lane >> 1, comp = lane & 1;
            *(uint2 *)&CT[comp][cr * CP] = *(const uint2 *)&CT[comp][cr * CP + 8];
        }
        wave_lds_sync();
    }
}

extern "C" size_t bc250_deblock_work_bytes(int wc, int hc)
{
    const size_t nmb = (size_t)(wc / 16) * (hc / 16);
    return 256 + nmb * GHAND * 8 + nmb * 16;
}

extern "C" int bc250_launch_deblock(uint8_t *y, uint8_t *uv, int py, int puv, int wc, int hc, int qp,
                                    int is_intra, const uint32_t *nz, const uint32_t *mvs, int num_slices, int idc,
                                    int *sync, unsigned spin_limit, void *stream)
{
    if (wc <= 0 || hc <= 0 || (wc & 15) || (hc & 15) || qp < 0 || qp > 51 || (py & 3) || (puv & 3) ||
        num_slices < 1 || idc < 0 || idc > 2 || wc / 16 >= 65535)
        return -1;
    if (idc == 1) return 0;
    hipStream_t st = (hipStream_t)stream;
    const int mbw = wc / 16, mbh = hc / 16, nmb = mbw * mbh;
    // workspace: 256 B of flags (ticket, error, epoch) | hand-off records | bS words
    uint64_t *ghand = (uint64_t *)((uint8_t *)sync + 256);
    uint4 *bsw = (uint4 *)(ghand + (size_t)nmb * GHAND);
    wavefront_begin_kernel<<<1, 64, 0, st>>>(sync, 2, (uint32_t *)sync + 2);
    h264_bs_kernel<<<(nmb + 7) / 8, 256, 0, st>>>(mbw, mbh, is_intra, nz, (const uint4 *)mvs, num_slices, idc, bsw,
                                                  ghand);
    const int groups = (mbh + ROWS - 1) / ROWS;
    if (is_intra)
        h264_deblock_kernel<true><<<groups, 128 * ROWS, 0, st>>>(y, uv, py, puv, mbw, mbh, qp, bsw, sync, ghand,
                                                                 spin_limit);
    else
        h264_deblock_kernel<false><<<groups, 128 * ROWS, 0, st>>>(y, uv, py, puv, mbw, mbh, qp, bsw, sync, ghand,
                                                                  spin_limit);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
