// learned_dev.h - device functions shared by the learned proposer's kernels
// (me_cnn.hip, me_learned.hip): the head of one 16x16 block and the
// candidate ordering, so that every launch that proposes candidates computes
// the same bits.
#pragma once
#include "bc250_dev.h"
#include "bc250_learned.h"

// The 4 hypotheses of one block's head outputs zb ordered by confidence
// (ties: lower index), in whole pels, clamped to the range, into cand[m].
__device__ __forceinline__ void cand_from_z(const float *zb, int range, int m, int2 *__restrict__ cand)
{
    int ord[4] = {0, 1, 2, 3};
    // stable selection by confidence, highest first
    for (int a = 0; a < 4; a++)
        for (int b = a + 1; b < 4; b++)
            if (zb[8 + ord[b]] > zb[8 + ord[a]]) { const int t = ord[a]; ord[a] = ord[b]; ord[b] = t; }
    for (int k = 0; k < 4; k++) {
        const int h = ord[k];
        const int dx = clampi((int)rintf(zb[2 * h] * 16.f), -range, range);
        const int dy = clampi((int)rintf(zb[2 * h + 1] * 16.f), -range, range);
        cand[(size_t)m * 4 + k] = make_int2(dx, dy);
    }
}

// The head's neighbour input k (0..15) of MB (mx, my): the previous
// picture's integer 16x16 vector at MB offset ((k & 3) - 1, (k >> 2) - 1),
// clamped to the picture (prev16 null: zero), pel / 16.
__device__ __forceinline__ void neighbour_input(const uint4 *__restrict__ prev16, int mbw, int mbh, int mx, int my,
                                                int k, float *o)
{
    const int nx = clampi(mx + (k & 3) - 1, 0, mbw - 1), ny = clampi(my + (k >> 2) - 1, 0, mbh - 1);
    const uint4 v = prev16 ? prev16[ny * mbw + nx] : make_uint4(0, 0, 0, 0);
    o[0] = (float)(int)v.x * (1.f / 16.f);
    o[1] = (float)(int)v.y * (1.f / 16.f);
}

// Output o of the head for one block: pb its 32 cell sums (16-byte
// aligned), mv its 32 neighbour inputs; the sums and fused multiply-adds in
// the order of cnn_head_kernel.
__device__ __forceinline__ float head_output(const float *__restrict__ wh, int o, const float *pb, const float *mv)
{
    const float *row = wh + o * (BC250_CNN_NFEAT + 1);
    const float4 *p4 = (const float4 *)pb;
    float sum = row[BC250_CNN_NFEAT];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const float4 v = p4[q];
        sum = fmaf(row[4 * q], v.x * BC250_CNN_POOL_SCALE, sum);
        sum = fmaf(row[4 * q + 1], v.y * BC250_CNN_POOL_SCALE, sum);
        sum = fmaf(row[4 * q + 2], v.z * BC250_CNN_POOL_SCALE, sum);
        sum = fmaf(row[4 * q + 3], v.w * BC250_CNN_POOL_SCALE, sum);
    }
#pragma unroll
    for (int i = 0; i < 32; i++) sum = fmaf(row[32 + i], mv[i], sum);
    return sum;
}

// The candidates of the (up to) 16 MBs of tile (tx, ty) of a picture mbw x
// mbh MBs, from the tile's cell sums spool [16][32] (LDS), by a workgroup of
// nthreads threads; smv [16][32] and sz [16][12] are LDS scratch.  Ends with
// every thread past a barrier.
__device__ __forceinline__ void propose_tile(const float *spool, const float *__restrict__ wh,
                                             const uint4 *__restrict__ prev16, int mbw, int mbh, int tx, int ty,
                                             int range, int2 *__restrict__ cand, float (*smv)[32],
                                             float (*sz)[BC250_CNN_NOUT], int nthreads)
{
    const int t = threadIdx.x;
    for (int i = t; i < 16 * 16; i += nthreads) {
        const int j = i >> 4, k = i & 15, mx = tx * 4 + (j & 3), my = ty * 4 + (j >> 2);
        if (mx < mbw && my < mbh) neighbour_input(prev16, mbw, mbh, mx, my, k, &smv[j][2 * k]);
    }
    __syncthreads();
    for (int i = t; i < 16 * BC250_CNN_NOUT; i += nthreads) {
        const int j = i / BC250_CNN_NOUT, o = i % BC250_CNN_NOUT, mx = tx * 4 + (j & 3), my = ty * 4 + (j >> 2);
        if (mx < mbw && my < mbh) sz[j][o] = head_output(wh, o, spool + j * 32, smv[j]);
    }
    __syncthreads();
    if (t < 16) {
        const int mx = tx * 4 + (t & 3), my = ty * 4 + (t >> 2);
        if (mx < mbw && my < mbh) cand_from_z(sz[t], range, my * mbw + mx, cand);
    }
}

// The search record of key k of the full search (q 0: 16x16, 1..4: the 8x8
// quadrants) at range `range`: the key is cost * 2^15 + raster index.
__device__ __forceinline__ uint4 fs_record(uint32_t k, int q, int range, int lambda8)
{
    const int nc = 2 * range + 1;
    const uint32_t idx = k & 0x7fffu, cost = k >> 15;
    const int dx = (int)(idx % nc) - range, dy = (int)(idx / nc) - range;
    const uint32_t l1 = (uint32_t)(iabs(dx) + iabs(dy));
    const uint32_t sad = cost - (q == 0 ? 4u : 1u) * (uint32_t)lambda8 * l1;
    return make_uint4((uint32_t)dx, (uint32_t)dy, cost, sad);
}
