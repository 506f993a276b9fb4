// me_learned.hip - the integer motion search of the learned form: the
// network's proposals refined in +-W windows, with the plain full search as
// the fallback, for gfx950.
//
//   1. cnn_mvin_kernel       : the head's neighbour-vector input of every
//                              block: the previous picture's integer 16x16
//                              vectors at MB offsets -1..2 in x and y
//   2. cnn_candidates_kernel : per MB the network's hypotheses ordered by
//                              confidence (ties: lower index), rounded to
//                              whole pels and clamped to the search range
//   3. cand_search_kernel    : per MB the (2W+1)^2 positions around each of
//                              the first n candidates, with the full
//                              search's costs and argmin keys (cost * 2^15 +
//                              raster index in the +-R square): wherever the
//                              full search's winner lies in a window, the
//                              result is the full search's, bit for bit
//   4. fallback_select_kernel: MBs whose 16x16 SAD exceeds a bound from
//                              their neighbours' SADs (and every MB above an
//                              absolute ceiling) go to a list, which
//                              me_fullsearch_kernel then searches in full
#include "bc250_dev.h"
#include "bc250_learned.h"
#include "bc250_stages.h"
#include "learned_dev.h"

namespace {

constexpr int CW = 4;  // window radius of the candidate search (pels)
constexpr int WS = 16 + 2 * CW;       // window side (24)
constexpr int WP = WS / 4 + 1;        // window pitch in dwords (7, odd)

__device__ __forceinline__ void mvin_block(const uint4 *__restrict__ prev16, int mbw, int mbh, int tw, int tile,
                                           int j, float *__restrict__ o)
{
    const int mx = (tile % tw) * 4 + (j & 3), my = (tile / tw) * 4 + (j >> 2);
    for (int k = 0; k < 16; k++) {
        const int nx = clampi(mx + (k & 3) - 1, 0, mbw - 1), ny = clampi(my + (k >> 2) - 1, 0, mbh - 1);
        const uint4 v = prev16 ? prev16[ny * mbw + nx] : make_uint4(0, 0, 0, 0);
        o[2 * k] = (float)(int)v.x * (1.f / 16.f);
        o[2 * k + 1] = (float)(int)v.y * (1.f / 16.f);
    }
}

__global__ void cnn_mvin_kernel(const uint4 *__restrict__ prev16, int mbw, int mbh, int tw, int ntiles,
                                float *__restrict__ mvin)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;  // (tile, block)
    if (b >= ntiles * 16) return;
    mvin_block(prev16, mbw, mbh, tw, b / 16, b % 16, mvin + (size_t)b * 32);
}

// Candidates of MB m of a picture whose tiles start at z (tw tiles per row).
__device__ __forceinline__ void candidates_mb(const float *__restrict__ z, int mbw, int tw, int range, int m,
                                              int2 *__restrict__ cand)
{
    const int mx = m % mbw, my = m / mbw;
    const int tile = (my >> 2) * tw + (mx >> 2), j = (my & 3) * 4 + (mx & 3);
    cand_from_z(z + ((size_t)tile * 16 + j) * BC250_CNN_NOUT, range, m, cand);
}

__global__ void cnn_candidates_kernel(const float *__restrict__ z, int mbw, int mbh, int tw, int range,
                                      int2 *__restrict__ cand)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= mbw * mbh) return;
    candidates_mb(z, mbw, tw, range, m, cand);
}

// One wave per MB.  The current block is wave-uniform (scalar registers);
// the windows of all candidates (edge-replicated like the full search's)
// are staged in LDS at once: a window inside the picture as aligned dwords,
// every lane's loads issued before its stores, one touching the edge sample
// by sample.  Lane 9 k + ox (k < ncand, ox = 0 .. 8) then owns the nine
// positions (ox, oy = 0 .. 8) of window k: it reads each window row once and
// adds it to every position whose block covers that row, so a row's LDS
// reads and byte alignment serve up to nine positions.
__device__ __forceinline__ void cand_search_mb(const uint8_t *__restrict__ cur, int cpitch,
                                               const uint8_t *__restrict__ ref, int rpitch, int wc, int hc, int mbw,
                                               int mb, int range, int lambda8, const int2 *__restrict__ cand,
                                               int ncand, uint4 *__restrict__ out16, uint4 *__restrict__ out8,
                                               uint32_t *win)
{
    constexpr int NO = 2 * CW + 1, WIN = WS * WP, NROW = 4 * WS;
    const int lane = threadIdx.x;
    const int x0 = (mb % mbw) * 16, y0 = (mb / mbw) * 16;
    const int nc = 2 * range + 1;
    uint32_t cw[16][4];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const uint32_t *p = (const uint32_t *)(cur + (size_t)(y0 + r) * cpitch + x0);
#pragma unroll
        for (int c = 0; c < 4; c++) cw[r][c] = __builtin_amdgcn_readfirstlane(p[c]);
    }
    int2 cv[4];
#pragma unroll
    for (int k = 0; k < 4; k++) cv[k] = k < ncand ? cand[(size_t)mb * 4 + k] : make_int2(0, 0);
    // window k's origin: the block at vector (c.x - W, c.y - W); rows of all
    // windows, two per lane
    const bool aligned = !(((uintptr_t)ref | (uintptr_t)rpitch) & 3);
    uint32_t raw[2][7];
    int rsh[2];
    bool fast[2];
#pragma unroll
    for (int it = 0; it < 2; it++) {
        const int i = lane + 64 * it, k = i / WS, r = i % WS;
        const int2 c = k == 0 ? cv[0] : k == 1 ? cv[1] : k == 2 ? cv[2] : cv[3];
        const int gx = x0 + c.x - CW, gy = y0 + c.y - CW + r;
        fast[it] = i < ncand * WS && aligned && gx >= 0 && gx + WS <= wc && gy >= 0 && gy < hc;
        rsh[it] = gx & 3;
        if (fast[it]) {
            const uint32_t *src = (const uint32_t *)(ref + (size_t)gy * rpitch + (gx & ~3));
#pragma unroll
            for (int d = 0; d < 6; d++) raw[it][d] = src[d];
            raw[it][6] = rsh[it] ? src[6] : 0u;  // only when the row reaches into it
        }
    }
#pragma unroll
    for (int it = 0; it < 2; it++) {
        const int i = lane + 64 * it, k = i / WS, r = i % WS;
        if (i >= ncand * WS || i >= NROW) continue;
        uint32_t *dst = win + k * WIN + r * WP;
        if (fast[it]) {
#pragma unroll
            for (int d = 0; d < 6; d++) dst[d] = __builtin_amdgcn_alignbyte(raw[it][d + 1], raw[it][d], rsh[it]);
        } else {
            const int2 c = k == 0 ? cv[0] : k == 1 ? cv[1] : k == 2 ? cv[2] : cv[3];
            const uint8_t *row = ref + (size_t)clampi(y0 + c.y - CW + r, 0, hc - 1) * rpitch;
#pragma unroll
            for (int d = 0; d < 6; d++) {
                const int gx = x0 + c.x - CW + 4 * d;
                dst[d] = (uint32_t)row[clampi(gx, 0, wc - 1)] | ((uint32_t)row[clampi(gx + 1, 0, wc - 1)] << 8) |
                         ((uint32_t)row[clampi(gx + 2, 0, wc - 1)] << 16) |
                         ((uint32_t)row[clampi(gx + 3, 0, wc - 1)] << 24);
            }
        }
        dst[WS / 4] = 0;  // pad dword
    }
    __syncthreads();
    uint32_t b16 = 0xffffffffu, bq[4] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu};
    if (lane < ncand * NO) {
        const int k = lane / NO, ox = lane % NO;
        const int2 c = k == 0 ? cv[0] : k == 1 ? cv[1] : k == 2 ? cv[2] : cv[3];
        const int sh = ox & 3, d0 = ox >> 2;
        uint32_t q[NO][4];
#pragma unroll
        for (int oy = 0; oy < NO; oy++)
#pragma unroll
            for (int h = 0; h < 4; h++) q[oy][h] = 0;
        const uint32_t *wk = win + k * WIN + d0;
#pragma unroll
        for (int r = 0; r < WS; r++) {
            uint32_t d[5];
#pragma unroll
            for (int i = 0; i < 5; i++) d[i] = wk[r * WP + i];
            uint32_t s[4];
#pragma unroll
            for (int i = 0; i < 4; i++) s[i] = __builtin_amdgcn_alignbyte(d[i + 1], d[i], sh);
#pragma unroll
            for (int oy = 0; oy < NO; oy++) {
                const int cr = r - oy;  // the block row this window row meets at offset oy
                if (cr >= 0 && cr < 16) {
                    const int h = (cr >> 3) * 2;
                    q[oy][h] = __builtin_amdgcn_sad_u8(s[1], cw[cr][1], __builtin_amdgcn_sad_u8(s[0], cw[cr][0], q[oy][h]));
                    q[oy][h + 1] =
                        __builtin_amdgcn_sad_u8(s[3], cw[cr][3], __builtin_amdgcn_sad_u8(s[2], cw[cr][2], q[oy][h + 1]));
                }
            }
        }
        const int dx = c.x - CW + ox;
#pragma unroll
        for (int oy = 0; oy < NO; oy++) {
            const int dy = c.y - CW + oy;
            if (dx < -range || dx > range || dy < -range || dy > range) continue;
            const uint32_t l1 = (uint32_t)(iabs(dx) + iabs(dy));
            const uint32_t idx = (uint32_t)((dy + range) * nc + (dx + range));
            const uint32_t sad16 = q[oy][0] + q[oy][1] + q[oy][2] + q[oy][3];
            b16 = min(b16, ((sad16 + 4u * lambda8 * l1) << 15) | idx);
#pragma unroll
            for (int qq = 0; qq < 4; qq++) bq[qq] = min(bq[qq], ((q[oy][qq] + (uint32_t)lambda8 * l1) << 15) | idx);
        }
    }
    b16 = wave_min_u32(b16);
#pragma unroll
    for (int qq = 0; qq < 4; qq++) bq[qq] = wave_min_u32(bq[qq]);
    if (lane < 5) {
        const uint32_t key = lane == 0 ? b16 : bq[lane - 1];
        const uint32_t idx = key & 0x7fffu, cost = key >> 15;
        const int dx = (int)(idx % nc) - range, dy = (int)(idx / nc) - range;
        const uint32_t l1 = (uint32_t)(iabs(dx) + iabs(dy));
        const uint32_t sad = cost - (lane == 0 ? 4u : 1u) * (uint32_t)lambda8 * l1;
        const uint4 o = make_uint4((uint32_t)dx, (uint32_t)dy, cost, sad);
        if (lane == 0) out16[mb] = o;
        else out8[4 * mb + lane - 1] = o;
    }
}

__global__ __launch_bounds__(64) void cand_search_kernel(const uint8_t *__restrict__ cur, int cpitch,
                                                         const uint8_t *__restrict__ ref, int rpitch, int wc, int hc,
                                                         int mbw, int nmb, int range, int lambda8,
                                                         const int2 *__restrict__ cand, int ncand,
                                                         uint4 *__restrict__ out16, uint4 *__restrict__ out8,
                                                         int *__restrict__ clear_count,
                                                         uint32_t *__restrict__ clear_keys)
{
    __shared__ uint32_t win[4 * WS * WP];
    const int mb = xcd_item(blockIdx.x, nmb);
    if (mb >= nmb) return;
    // the fallback list of this picture starts empty (its previous use
    // ended before this launch; its next starts after it)
    if (clear_count && mb == 0 && threadIdx.x == 0) *clear_count = 0;
    // clear_keys: every MB's split-search keys start empty, so that a set
    // key after the search marks an MB searched in full
    if (clear_keys && threadIdx.x < 5) clear_keys[5 * mb + threadIdx.x] = 0xffffffffu;
    cand_search_mb(cur, cpitch, ref, rpitch, wc, hc, mbw, mb, range, lambda8, cand, ncand, out16, out8, win);
}

__global__ void zero1_kernel(int *p) { *p = 0; }

// The fallback rule: an MB goes to the full search when its 16x16 SAD
// exceeds (alpha_q4 / 16) * m + beta, m the median (lower middle) of its
// left, right, top and bottom neighbours' SADs, or exceeds t_high.
__device__ __forceinline__ bool fallback_mb(const uint4 *__restrict__ r16, int mbw, int mbh, int alpha_q4, int beta,
                                            int t_high, int m)
{
    const int mx = m % mbw, my = m / mbw;
    uint32_t nb[4];
    int n = 0;
    if (mx > 0) nb[n++] = r16[m - 1].w;
    if (mx + 1 < mbw) nb[n++] = r16[m + 1].w;
    if (my > 0) nb[n++] = r16[m - mbw].w;
    if (my + 1 < mbh) nb[n++] = r16[m + mbw].w;
    for (int a = 0; a < n; a++)
        for (int b = a + 1; b < n; b++)
            if (nb[b] < nb[a]) { const uint32_t t = nb[a]; nb[a] = nb[b]; nb[b] = t; }
    const uint32_t med = n ? nb[(n - 1) / 2] : 0u;
    const uint32_t sad = r16[m].w;
    const uint64_t bound = ((uint64_t)alpha_q4 * med >> 4) + (uint64_t)beta;
    return sad > bound || sad > (uint32_t)t_high;
}

// keys (optional): the split full search's keys, 5 per MB of the picture
// (batch), those of a listed MB set to the empty key as it is listed.
__device__ __forceinline__ void list_mb(int *__restrict__ list, int *__restrict__ count, uint32_t *__restrict__ keys,
                                        int m)
{
    const int i = atomicAdd(count, 1);
    list[i] = m;
    if (keys) {
#pragma unroll
        for (int q = 0; q < 5; q++) keys[5 * m + q] = 0xffffffffu;
    }
}

__global__ void fallback_select_kernel(const uint4 *__restrict__ r16, int mbw, int mbh, int alpha_q4, int beta,
                                       int t_high, int *__restrict__ list, int *__restrict__ count,
                                       uint32_t *__restrict__ keys)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= mbw * mbh) return;
    if (fallback_mb(r16, mbw, mbh, alpha_q4, beta, t_high, m)) list_mb(list, count, keys, m);
}

// ---- batch forms: the picture of each batch-wide tile / MB from the descriptor

__global__ __launch_bounds__(64) void cand_search_batch_kernel(const bc250_me_batch_t b, int ncand,
                                                               int *__restrict__ clear_count)
{
    __shared__ uint32_t win[4 * WS * WP];
    const int g = xcd_item(blockIdx.x, b.nmb_total);
    if (g >= b.nmb_total) return;
    if (clear_count && g == 0 && threadIdx.x == 0) *clear_count = 0;
    const int f = batch_frame_of_mb(b, g);
    const bc250_batch_frame_t &fr = b.f[f];
    cand_search_mb(fr.cur, fr.cpitch, fr.ref, fr.rpitch, fr.wc, fr.hc, fr.wc / 16, g - fr.mb_base, b.range, b.lambda8,
                   (const int2 *)fr.cand, ncand, (uint4 *)fr.out16, (uint4 *)fr.out8, win);
}

__global__ void fallback_select_batch_kernel(const bc250_me_batch_t b, int alpha_q4, int beta, int t_high,
                                             int *__restrict__ list, int *__restrict__ count,
                                             uint32_t *__restrict__ keys)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= b.nmb_total) return;
    const int f = batch_frame_of_mb(b, g);
    const bc250_batch_frame_t &fr = b.f[f];
    if (fallback_mb((const uint4 *)fr.out16, fr.wc / 16, fr.hc / 16, alpha_q4, beta, t_high, g - fr.mb_base))
        list_mb(list, count, keys, g);
}

// ---- proposals in one launch: per MB its neighbour-vector features, the
// head's 12 outputs (the order and fused multiply-adds of cnn_head_kernel)
// and the candidates (cand_from_z).  A workgroup takes PMB consecutive MBs
// (batch-wide for a batch), one thread per (MB, neighbour) for the features,
// per (MB, output) for the head and per MB for the candidates.
constexpr int PMB = 16;

struct ProposeSrc {       // one picture's inputs and candidate array
    const float *pool;    // [tile][16][32] of the picture
    const uint4 *prev16;
    int2 *cand;
    int mbw, mbh, m;      // m: the local MB, -1 for none
};

template <typename GetSrc>
__device__ __forceinline__ void propose_group(const float *__restrict__ wh, int range, int g0, GetSrc src_of)
{
    __shared__ float smv[PMB][32];
    __shared__ float sz[PMB][BC250_CNN_NOUT];
    const int t = threadIdx.x;
    for (int i = t; i < PMB * 16; i += PMB * BC250_CNN_NOUT) {
        const int j = i >> 4, k = i & 15;
        const ProposeSrc s = src_of(g0 + j);
        if (s.m >= 0) neighbour_input(s.prev16, s.mbw, s.mbh, s.m % s.mbw, s.m / s.mbw, k, &smv[j][2 * k]);
    }
    __syncthreads();
    if (t < PMB * BC250_CNN_NOUT) {
        const int j = t / BC250_CNN_NOUT, o = t % BC250_CNN_NOUT;
        const ProposeSrc s = src_of(g0 + j);
        if (s.m >= 0) {
            const int tw = (s.mbw + 3) / 4, mx = s.m % s.mbw, my = s.m / s.mbw;
            const int tile = (my >> 2) * tw + (mx >> 2), jb = (my & 3) * 4 + (mx & 3);
            sz[j][o] = head_output(wh, o, s.pool + ((size_t)tile * 16 + jb) * 32, smv[j]);
        }
    }
    __syncthreads();
    if (t < PMB) {
        const ProposeSrc s = src_of(g0 + t);
        if (s.m >= 0) cand_from_z(sz[t], range, s.m, s.cand);
    }
}

__global__ __launch_bounds__(PMB * BC250_CNN_NOUT) void cnn_propose_kernel(const float *__restrict__ pool,
                                                                          const float *__restrict__ wh,
                                                                          const uint4 *__restrict__ prev16, int mbw,
                                                                          int mbh, int range, int2 *__restrict__ cand)
{
    const int nmb = mbw * mbh;
    propose_group(wh, range, blockIdx.x * PMB, [&](int m) {
        return ProposeSrc{pool, prev16, cand, mbw, mbh, m < nmb ? m : -1};
    });
}

}  // namespace

extern "C" int bc250_cnn_mvin(const uint32_t *prev16, int mbw, int mbh, float *mvin, void *stream)
{
    if (mbw <= 0 || mbh <= 0 || !mvin) return -1;
    const int tw = (mbw + 3) / 4, nt = tw * ((mbh + 3) / 4);
    cnn_mvin_kernel<<<(nt * 16 + 127) / 128, 128, 0, (hipStream_t)stream>>>((const uint4 *)prev16, mbw, mbh, tw, nt,
                                                                           mvin);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int bc250_cnn_candidates(const float *z, int mbw, int mbh, int range, int *cand, void *stream)
{
    if (mbw <= 0 || mbh <= 0 || !z || !cand || range < CW || range > 64) return -1;
    const int nmb = mbw * mbh;
    cnn_candidates_kernel<<<(nmb + 127) / 128, 128, 0, (hipStream_t)stream>>>(z, mbw, mbh, (mbw + 3) / 4, range,
                                                                             (int2 *)cand);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int bc250_launch_cand_search(const uint8_t *cur, int cpitch, const uint8_t *ref, int rpitch, int wc,
                                        int hc, int range, int lambda8, const int *cand, int ncand, uint32_t *out16,
                                        uint32_t *out8, void *stream)
{
    if (wc <= 0 || hc <= 0 || (wc & 15) || (hc & 15) || (cpitch & 3) || lambda8 < 0 || lambda8 > 15 ||
        ncand < 1 || ncand > 4 || !cand || range < CW || range > 64 || (2 * range + 1) * (2 * range + 1) > 32768)
        return -1;
    const int mbw = wc / 16, nmb = mbw * (hc / 16);
    cand_search_kernel<<<((nmb + 7) / 8) * 8, 64, 0, (hipStream_t)stream>>>(
        cur, cpitch, ref, rpitch, wc, hc, mbw, nmb, range, lambda8, (const int2 *)cand, ncand, (uint4 *)out16,
        (uint4 *)out8, nullptr, nullptr);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int bc250_fallback_select(const uint32_t *r16, int mbw, int mbh, int alpha_q4, int beta, int t_high,
                                     int *list, int *count, void *stream)
{
    if (mbw <= 0 || mbh <= 0 || !r16 || !list || !count || alpha_q4 < 0 || beta < 0 || t_high < 0) return -1;
    const int nmb = mbw * mbh;
    hipStream_t st = (hipStream_t)stream;
    zero1_kernel<<<1, 1, 0, st>>>(count);
    fallback_select_kernel<<<(nmb + 127) / 128, 128, 0, st>>>((const uint4 *)r16, mbw, mbh, alpha_q4, beta, t_high,
                                                              list, count, nullptr);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int bc250_cnn_propose(const float *pool, const float *params, const uint32_t *prev16, int mbw, int mbh,
                                 int range, int *cand, void *stream)
{
    if (!pool || !params || !cand || mbw <= 0 || mbh <= 0 || range < CW || range > 64) return -1;
    const int nmb = mbw * mbh;
    cnn_propose_kernel<<<(nmb + PMB - 1) / PMB, PMB * BC250_CNN_NOUT, 0, (hipStream_t)stream>>>(
        pool, params + bc250_cnn_w_offset(4), (const uint4 *)prev16, mbw, mbh, range, (int2 *)cand);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

static int learned_search(const uint8_t *cur, int cpitch, const uint8_t *ref, int rpitch, int wc, int hc, int range,
                          int lambda8, const float *params, const uint16_t *packed, const uint32_t *prev16, int ncand,
                          int alpha_q4, int beta, int t_high, float *pool, int *cand, int *list, int *count,
                          uint32_t *keys, uint32_t *out16, uint32_t *out8, void *stream, bool records)
{
    if (wc <= 0 || hc <= 0 || (wc & 15) || (hc & 15) || (cpitch & 3) || lambda8 < 0 || lambda8 > 15 ||
        ncand < 1 || ncand > 4 || range < CW || range > 64 || !params || !packed || !pool || !cand || !list ||
        !count || !keys || !out16 || !out8 || alpha_q4 < 0 || beta < 0 || t_high < 0)
        return -1;
    hipStream_t st = (hipStream_t)stream;
    const int mbw = wc / 16, mbh = hc / 16, nmb = mbw * mbh;
    int r = bc250_cnn_forward_propose_picture(cur, cpitch, ref, rpitch, wc, hc, packed, params, pool, prev16, range,
                                              cand, st);
    if (r == 0) {
        cand_search_kernel<<<((nmb + 7) / 8) * 8, 64, 0, st>>>(cur, cpitch, ref, rpitch, wc, hc, mbw, nmb, range,
                                                               lambda8, (const int2 *)cand, ncand, (uint4 *)out16,
                                                               (uint4 *)out8, count, records ? nullptr : keys);
        fallback_select_kernel<<<(nmb + 127) / 128, 128, 0, st>>>((const uint4 *)out16, mbw, mbh, alpha_q4, beta,
                                                                  t_high, list, count, keys);
        r = hipGetLastError() == hipSuccess ? 0 : -1;
    }
    if (r == 0)
        r = bc250_launch_me_fullsearch_list_split(cur, cpitch, ref, rpitch, wc, hc, range, lambda8, out16, out8, list,
                                                  count, keys, records ? 1 : 3, st);
    return r;
}

extern "C" int bc250_learned_search_picture(const uint8_t *cur, int cpitch, const uint8_t *ref, int rpitch, int wc,
                                            int hc, int range, int lambda8, const float *params,
                                            const uint16_t *packed, const uint32_t *prev16, int ncand, int alpha_q4,
                                            int beta, int t_high, float *pool, int *cand, int *list, int *count,
                                            uint32_t *keys, uint32_t *out16, uint32_t *out8, void *stream)
{
    return learned_search(cur, cpitch, ref, rpitch, wc, hc, range, lambda8, params, packed, prev16, ncand, alpha_q4,
                          beta, t_high, pool, cand, list, count, keys, out16, out8, stream, true);
}

extern "C" int bc250_learned_search_picture_keys(const uint8_t *cur, int cpitch, const uint8_t *ref, int rpitch,
                                                 int wc, int hc, int range, int lambda8, const float *params,
                                                 const uint16_t *packed, const uint32_t *prev16, int ncand,
                                                 int alpha_q4, int beta, int t_high, float *pool, int *cand,
                                                 int *list, int *count, uint32_t *keys, uint32_t *out16,
                                                 uint32_t *out8, void *stream)
{
    return learned_search(cur, cpitch, ref, rpitch, wc, hc, range, lambda8, params, packed, prev16, ncand, alpha_q4,
                          beta, t_high, pool, cand, list, count, keys, out16, out8, stream, false);
}

// ---------------------------------------------------------------------------
// The batched learned search.  Work layout (each part 256-byte aligned):
// block sums, the fallback list, its count, the split search's keys.
namespace {
struct LearnedWork {
    float *pool;
    int *list, *count;
    uint32_t *keys;
    size_t bytes;
};
size_t al256(size_t v) { return (v + 255) & ~(size_t)255; }
LearnedWork carve_learned(void *base, int nt, int nmb)
{
    uint8_t *p = (uint8_t *)base;
    LearnedWork w;
    size_t o = 0;
    auto take = [&](size_t n) { uint8_t *q = p ? p + o : nullptr; o += al256(n); return q; };
    w.pool = (float *)take((size_t)nt * 16 * 32 * 4);
    w.list = (int *)take((size_t)nmb * 4);
    w.count = (int *)take(4);
    w.keys = (uint32_t *)take((size_t)nmb * 5 * 4);
    w.bytes = o;
    return w;
}
}  // namespace

extern "C" size_t bc250_learned_work_bytes(int ntiles, int nmb)
{
    if (ntiles <= 0 || nmb <= 0) return 0;
    return carve_learned(nullptr, ntiles, nmb).bytes;
}

extern "C" const int *bc250_learned_fallback_count(const void *work, int ntiles, int nmb)
{
    if (!work || ntiles <= 0 || nmb <= 0) return nullptr;
    return carve_learned((void *)work, ntiles, nmb).count;
}

extern "C" int bc250_learned_search_batch(const bc250_me_batch_t *b, const float *params, const uint16_t *packed,
                                          int ncand, int alpha_q4, int beta, int t_high, void *work, int work_tiles,
                                          int work_mbs, void *stream)
{
    if (!b || b->n < 1 || b->n > BC250_BATCH_MAX || b->nmb_total <= 0 || b->ntile_total <= 0 || !params ||
        !packed || !work || ncand < 1 || ncand > 4 || b->range < CW || b->range > 64 ||
        b->ntile_total > work_tiles || b->nmb_total > work_mbs)
        return -1;
    const LearnedWork w = carve_learned(work, work_tiles, work_mbs);
    hipStream_t st = (hipStream_t)stream;
    int r = bc250_cnn_forward_propose_batch(b, packed, params, w.pool, st);
    if (r == 0) {
        cand_search_batch_kernel<<<((b->nmb_total + 7) / 8) * 8, 64, 0, st>>>(*b, ncand, w.count);
        fallback_select_batch_kernel<<<(b->nmb_total + 127) / 128, 128, 0, st>>>(*b, alpha_q4, beta, t_high, w.list,
                                                                                 w.count, w.keys);
        r = hipGetLastError() == hipSuccess ? 0 : -1;
    }
    if (r == 0) r = bc250_launch_me_fullsearch_batch_split(b, w.list, w.count, w.keys, 1, st);
    return r;
}
