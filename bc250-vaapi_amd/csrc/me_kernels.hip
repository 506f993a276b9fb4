// me_kernels.hip - +-R integer full-search motion estimation and the
// half/quarter-pel refinement, hand-written for gfx950.
//
// Integer search (me_fullsearch_kernel): one 256-thread workgroup per
// macroblock.  The reference window of the MB (edge-replicated) is staged in
// LDS; the 16x16 current block is wave-uniform and lives in SGPRs.  Every
// lane owns a tile of 4 (dx) x 9 (dy) candidates (at R = 64: 33 x 15 = 495
// tiles, two per thread; 4-wave workgroups load the 4 SIMDs of a CU evenly,
// 4 workgroups per CU): for each reference row it reads 5 aligned dwords from LDS,
// builds the byte-shifted copies with v_alignbyte_b32, and feeds v_sad_u8 /
// v_sad_hi_u8 into two registers per candidate that hold the four 8x8
// quadrant costs as 16-bit halves (72 VGPRs), so the 16x16 cost and the four
// 8x8 costs come from one pass.  Accumulators start at lambda8*|mv|_1, which makes
//   cost8 = SAD8 + lambda8*|mv|_1      cost16 = SAD16 + 4*lambda8*|mv|_1.
// The argmin key is cost*2^15 + raster index of (dx,dy) (ties -> lowest
// raster index), reduced per lane, across the wave with DPP shuffles and
// across waves through LDS.  Measured on MI355X at 1080p, +-64: 0.417 ms per
// frame (this shape; 0.406 ms with the one-instruction argmin keys below)
// against 0.425 ms for 32-bit accumulators and 0.544 ms for 4x8 tiles on
// 3-wave workgroups, which leave SIMDs unevenly loaded.
#include "bc250_dev.h"
#include "bc250_learned.h"
#include "bc250_stages.h"
#include "learned_dev.h"

namespace {

// Launch shape, overridable at build time for tuning sweeps: threads per
// workgroup, dy candidates per lane tile, and the waves per SIMD the
// register budget must allow.
#ifndef BC250_ME_THREADS
#define BC250_ME_THREADS 256
#endif
#ifndef BC250_ME_DY
#define BC250_ME_DY 9
#endif
#ifndef BC250_ME_WAVES_PER_SIMD
#define BC250_ME_WAVES_PER_SIMD 4
#endif
constexpr int ME_THREADS = BC250_ME_THREADS;
constexpr int DY = BC250_ME_DY;  // dy candidates per lane task

template <int R>
struct FsGeom {
    static constexpr int NC = 2 * R + 1;          // candidates per axis
    static constexpr int NG = (NC + 3) / 4;       // dx groups of 4
    static constexpr int NGY = (NC + DY - 1) / DY;  // dy groups of DY
    static constexpr int NT = NG * NGY;           // lane tasks per MB
    static constexpr int NR = NGY * DY + 15;      // window rows
    static constexpr int WD = NG + 4;             // window dwords per row
    static constexpr int PITCH = WD | 1;          // odd dword pitch
    static constexpr int SLOTS = (NT + ME_THREADS - 1) / ME_THREADS;
    static_assert(R % 4 == 0, "range must be a multiple of 4");
    static_assert(NC * NC <= 32768, "raster index must fit 15 bits");
};

// Argmin key of a 16-bit cost half, cost * 64 + li, in one v_mad_u32_u16
// (op_sel picks the high half); the compiler builds it from a mask or shift
// and an or, two to three instructions.
BC_DEV uint32_t mad_lo16(uint32_t x, uint32_t li)
{
    uint32_t d;
    asm("v_mad_u32_u16 %0, %1, 64, %2" : "=v"(d) : "v"(x), "v"(li));
    return d;
}
BC_DEV uint32_t mad_hi16(uint32_t x, uint32_t li)
{
    uint32_t d;
    asm("v_mad_u32_u16 %0, %1, 64, %2 op_sel:[1,0,0,0]" : "=v"(d) : "v"(x), "v"(li));
    return d;
}
}  // namespace

// One lane task of the full search (the split list search's copy of the
// loop body in fullsearch_mb, which keeps its own inline so that its
// register allocation stays spill-free): the 4 (dx) x DY (dy) candidates
// from (dx0, dy0), window rows from wrow (the task's first row and dword),
// keys folded into best16 / bq.
template <int R>
__device__ __forceinline__ void fs_task(const uint32_t *__restrict__ wrow, const uint32_t (&cw)[16][4], int lambda8,
                                        int dx0, int dy0, uint32_t &best16, uint32_t (&bq)[4])
{
    using Gm = FsGeom<R>;
    // Two 8x8-quadrant costs per register: the left quadrant in bits 0..15
    // (v_sad_u8), the right one in bits 16..31 (v_sad_hi_u8); acc[j][k][0]
    // rows 0-7, [1] rows 8-15.  Both start at lambda8*|dx| + lambda8*|dy|,
    // plus 20480 per axis outside the range: a valid quadrant costs at most
    // 16320 + 15*128 = 18240 < 20480, and no half exceeds 2*20480 + 2055 +
    // 16320 < 2^16, so the halves never carry into each other.
    uint32_t ax[4], by[DY];
#pragma unroll
    for (int k = 0; k < 4; k++)
        ax[k] = (uint32_t)(lambda8 * iabs(dx0 + k)) + (dx0 + k > R ? 20480u : 0u);
#pragma unroll
    for (int j = 0; j < DY; j++)
        by[j] = (uint32_t)(lambda8 * iabs(dy0 + j)) + (dy0 + j > R ? 20480u : 0u);
    uint32_t acc[DY][4][2];
#pragma unroll
    for (int j = 0; j < DY; j++)
#pragma unroll
        for (int k = 0; k < 4; k++) acc[j][k][0] = acc[j][k][1] = (ax[k] + by[j]) * 0x10001u;
#pragma unroll
    for (int r = 0; r < DY + 15; r++) {
        uint32_t d[5];
#pragma unroll
        for (int i = 0; i < 5; i++) d[i] = wrow[r * Gm::PITCH + i];
        uint32_t s[4][4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            s[0][c] = d[c];
            s[1][c] = __builtin_amdgcn_alignbyte(d[c + 1], d[c], 1);
            s[2][c] = __builtin_amdgcn_alignbyte(d[c + 1], d[c], 2);
            s[3][c] = __builtin_amdgcn_alignbyte(d[c + 1], d[c], 3);
        }
#pragma unroll
        for (int j = 0; j < DY; j++) {
            const int cr = r - j;
            if (cr < 0 || cr > 15) continue;
            const int h = cr >> 3;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                uint32_t a = acc[j][k][h];
                a = __builtin_amdgcn_sad_u8(s[k][0], cw[cr][0], a);
                a = __builtin_amdgcn_sad_u8(s[k][1], cw[cr][1], a);
                a = __builtin_amdgcn_sad_hi_u8(s[k][2], cw[cr][2], a);
                a = __builtin_amdgcn_sad_hi_u8(s[k][3], cw[cr][3], a);
                acc[j][k][h] = a;
            }
        }
    }
    // per-lane argmin with a 6-bit local index (j-major = raster order)
    uint32_t l16 = 0xffffffffu, lq[4] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu};
#pragma unroll
    for (int j = 0; j < DY; j++)
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t li = (uint32_t)(j * 4 + k);
            const uint32_t top = acc[j][k][0], bot = acc[j][k][1];
            // the four halves summed: |x.lo - 0| + |x.hi - 0| + c
            const uint32_t s16 = __builtin_amdgcn_sad_u16(bot, 0u, __builtin_amdgcn_sad_u16(top, 0u, 0u));
            l16 = min(l16, (s16 << 6) | li);
            lq[0] = min(lq[0], mad_lo16(top, li));
            lq[1] = min(lq[1], mad_hi16(top, li));
            lq[2] = min(lq[2], mad_lo16(bot, li));
            lq[3] = min(lq[3], mad_hi16(bot, li));
        }
    // promote to global keys: cost*2^15 + raster index (every tile holds a
    // candidate inside the range, so its minimum is one)
    auto gkey = [&](uint32_t lk) -> uint32_t {
        uint32_t li = lk & 63u, cost = lk >> 6;
        int dx = dx0 + (int)(li & 3u), dy = dy0 + (int)(li >> 2);
        uint32_t idx = (uint32_t)((dy + R) * Gm::NC + (dx + R));
        return (cost << 15) | idx;
    };
    best16 = min(best16, gkey(l16));
#pragma unroll
    for (int q = 0; q < 4; q++) bq[q] = min(bq[q], gkey(lq[q]));
}

// The full search of one MB (mb of a frame mbw MBs wide) by one workgroup:
// out16[mb] = {dx, dy, cost, sad}; out8[4 mb + q] likewise for quadrant q.
template <int R>
__device__ __forceinline__ void fullsearch_mb(const uint8_t *__restrict__ cur, int cpitch,
                                              const uint8_t *__restrict__ ref, int rpitch, int wc, int hc, int mbw,
                                              int mb, int lambda8, uint4 *__restrict__ out16,
                                              uint4 *__restrict__ out8, uint32_t *win, uint32_t (*red)[5])
{
    using Gm = FsGeom<R>;
    const int mbx = mb % mbw, mby = mb / mbw;
    const int x0 = mbx * 16, y0 = mby * 16;
    const int tid = threadIdx.x;

    // ---- stage the edge-replicated reference window in LDS
    const int wx0 = x0 - R, wy0 = y0 - R;
    for (int i = tid; i < Gm::NR * Gm::WD; i += ME_THREADS) {
        int r = i / Gm::WD, d = i - r * Gm::WD;
        int gy = clampi(wy0 + r, 0, hc - 1);
        int gx = wx0 + 4 * d;
        const uint8_t *row = ref + (size_t)gy * rpitch;
        uint32_t v;
        if (gx >= 0 && gx + 3 < wc) {
            v = *(const uint32_t *)(row + gx);
        } else {
            v = (uint32_t)row[clampi(gx, 0, wc - 1)] | ((uint32_t)row[clampi(gx + 1, 0, wc - 1)] << 8) |
                ((uint32_t)row[clampi(gx + 2, 0, wc - 1)] << 16) |
                ((uint32_t)row[clampi(gx + 3, 0, wc - 1)] << 24);
        }
        win[r * Gm::PITCH + d] = v;
    }

    // ---- current block: uniform loads (scalar unit)
    uint32_t cw[16][4];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const uint32_t *p = (const uint32_t *)(cur + (size_t)(y0 + r) * cpitch + x0);
#pragma unroll
        for (int c = 0; c < 4; c++) cw[r][c] = __builtin_amdgcn_readfirstlane(p[c]);
    }
    __syncthreads();

    uint32_t best16 = 0xffffffffu, bq[4] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu};

#pragma unroll 1
    for (int slot = 0; slot < Gm::SLOTS; slot++) {
        const int t = tid + slot * ME_THREADS;
        if (t < Gm::NT) {
            const int dxg = t % Gm::NG, dyg = t / Gm::NG;
            const int dx0 = -R + 4 * dxg, dy0 = -R + DY * dyg;
            // Two 8x8-quadrant costs per register: the left quadrant in bits 0..15
            // (v_sad_u8), the right one in bits 16..31 (v_sad_hi_u8); acc[j][k][0]
            // rows 0-7, [1] rows 8-15.  Both start at lambda8*|dx| + lambda8*|dy|,
            // plus 20480 per axis outside the range: a valid quadrant costs at most
            // 16320 + 15*128 = 18240 < 20480, and no half exceeds 2*20480 + 2055 +
            // 16320 < 2^16, so the halves never carry into each other.
            uint32_t ax[4], by[DY];
#pragma unroll
            for (int k = 0; k < 4; k++)
                ax[k] = (uint32_t)(lambda8 * iabs(dx0 + k)) + (dx0 + k > R ? 20480u : 0u);
#pragma unroll
            for (int j = 0; j < DY; j++)
                by[j] = (uint32_t)(lambda8 * iabs(dy0 + j)) + (dy0 + j > R ? 20480u : 0u);
            uint32_t acc[DY][4][2];
#pragma unroll
            for (int j = 0; j < DY; j++)
#pragma unroll
                for (int k = 0; k < 4; k++) acc[j][k][0] = acc[j][k][1] = (ax[k] + by[j]) * 0x10001u;
            const uint32_t *wrow = win + (dyg * DY) * Gm::PITCH + dxg;
#pragma unroll
            for (int r = 0; r < DY + 15; r++) {
                uint32_t d[5];
#pragma unroll
                for (int i = 0; i < 5; i++) d[i] = wrow[r * Gm::PITCH + i];
                uint32_t s[4][4];
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    s[0][c] = d[c];
                    s[1][c] = __builtin_amdgcn_alignbyte(d[c + 1], d[c], 1);
                    s[2][c] = __builtin_amdgcn_alignbyte(d[c + 1], d[c], 2);
                    s[3][c] = __builtin_amdgcn_alignbyte(d[c + 1], d[c], 3);
                }
#pragma unroll
                for (int j = 0; j < DY; j++) {
                    const int cr = r - j;
                    if (cr < 0 || cr > 15) continue;
                    const int h = cr >> 3;
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        uint32_t a = acc[j][k][h];
                        a = __builtin_amdgcn_sad_u8(s[k][0], cw[cr][0], a);
                        a = __builtin_amdgcn_sad_u8(s[k][1], cw[cr][1], a);
                        a = __builtin_amdgcn_sad_hi_u8(s[k][2], cw[cr][2], a);
                        a = __builtin_amdgcn_sad_hi_u8(s[k][3], cw[cr][3], a);
                        acc[j][k][h] = a;
                    }
                }
            }
            // per-lane argmin with a 6-bit local index (j-major = raster order)
            uint32_t l16 = 0xffffffffu, lq[4] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu};
#pragma unroll
            for (int j = 0; j < DY; j++)
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const uint32_t li = (uint32_t)(j * 4 + k);
                    const uint32_t top = acc[j][k][0], bot = acc[j][k][1];
                    // the four halves summed: |x.lo - 0| + |x.hi - 0| + c
                    const uint32_t s16 = __builtin_amdgcn_sad_u16(bot, 0u, __builtin_amdgcn_sad_u16(top, 0u, 0u));
                    l16 = min(l16, (s16 << 6) | li);
                    lq[0] = min(lq[0], mad_lo16(top, li));
                    lq[1] = min(lq[1], mad_hi16(top, li));
                    lq[2] = min(lq[2], mad_lo16(bot, li));
                    lq[3] = min(lq[3], mad_hi16(bot, li));
                }
            // promote to global keys: cost*2^15 + raster index (every tile holds a
            // candidate inside the range, so its minimum is one)
            auto gkey = [&](uint32_t lk) -> uint32_t {
                uint32_t li = lk & 63u, cost = lk >> 6;
                int dx = dx0 + (int)(li & 3u), dy = dy0 + (int)(li >> 2);
                uint32_t idx = (uint32_t)((dy + R) * Gm::NC + (dx + R));
                return (cost << 15) | idx;
            };
            best16 = min(best16, gkey(l16));
#pragma unroll
            for (int q = 0; q < 4; q++) bq[q] = min(bq[q], gkey(lq[q]));
        }
    }

    // ---- reduce across the wave, then across the 3 waves
    best16 = wave_min_u32(best16);
#pragma unroll
    for (int q = 0; q < 4; q++) bq[q] = wave_min_u32(bq[q]);
    const int wave = tid >> 6, lane = tid & 63;
    if (lane == 0) {
        red[wave][0] = best16;
#pragma unroll
        for (int q = 0; q < 4; q++) red[wave][1 + q] = bq[q];
    }
    __syncthreads();
    if (tid < 5) {
        uint32_t k = red[0][tid];
#pragma unroll
        for (int w = 1; w < ME_THREADS / 64; w++) k = min(k, red[w][tid]);
        uint32_t idx = k & 0x7fffu, cost = k >> 15;
        int dx = (int)(idx % Gm::NC) - R, dy = (int)(idx / Gm::NC) - R;
        uint32_t l1 = (uint32_t)(iabs(dx) + iabs(dy));
        uint32_t sad = cost - (tid == 0 ? 4u : 1u) * (uint32_t)lambda8 * l1;
        uint4 o = make_uint4((uint32_t)dx, (uint32_t)dy, cost, sad);
        if (tid == 0) out16[mb] = o;
        else out8[4 * mb + tid - 1] = o;
    }
}

// With `list`, only the MBs list[0 .. *count) are searched (count is read on
// the device, so the grid is sized for the largest list).
template <int R>
__global__ __launch_bounds__(ME_THREADS, BC250_ME_WAVES_PER_SIMD) void me_fullsearch_kernel(
    const uint8_t *__restrict__ cur, int cpitch, const uint8_t *__restrict__ ref, int rpitch,
    int wc, int hc, int mbw, int nmb, int lambda8, uint4 *__restrict__ out16,
    uint4 *__restrict__ out8, const int *__restrict__ list, const int *__restrict__ count)
{
    using Gm = FsGeom<R>;
    __shared__ uint32_t win[Gm::NR * Gm::PITCH];
    __shared__ uint32_t red[ME_THREADS / 64][5];
    int mb;
    if (list) {
        const int n = min(*count, nmb), per = (n + 7) >> 3;
        if ((int)(blockIdx.x >> 3) >= per) return;
        const int i = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
        if (i >= n) return;
        mb = list[i];
    } else {
        mb = xcd_item(blockIdx.x, nmb);
    }
    if (mb >= nmb) return;
    fullsearch_mb<R>(cur, cpitch, ref, rpitch, wc, hc, mbw, mb, lambda8, out16, out8, win, red);
}

// The same search over the MBs of a batch of pictures (one workgroup per MB
// of every picture, or per listed MB: list holds batch-wide MB indices).
template <int R>
__global__ __launch_bounds__(ME_THREADS, BC250_ME_WAVES_PER_SIMD) void me_fullsearch_batch_kernel(
    const bc250_me_batch_t b, const int *__restrict__ list, const int *__restrict__ count)
{
    using Gm = FsGeom<R>;
    __shared__ uint32_t win[Gm::NR * Gm::PITCH];
    __shared__ uint32_t red[ME_THREADS / 64][5];
    int g;
    if (list) {
        const int n = min(*count, b.nmb_total), per = (n + 7) >> 3;
        if ((int)(blockIdx.x >> 3) >= per) return;
        const int i = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
        if (i >= n) return;
        g = list[i];
    } else {
        g = xcd_item(blockIdx.x, b.nmb_total);
    }
    if (g >= b.nmb_total) return;
    const int f = batch_frame_of_mb(b, g);
    const bc250_batch_frame_t &fr = b.f[f];
    fullsearch_mb<R>(fr.cur, fr.cpitch, fr.ref, fr.rpitch, fr.wc, fr.hc, fr.wc / 16, g - fr.mb_base, b.lambda8,
                     (uint4 *)fr.out16, (uint4 *)fr.out8, win, red);
}

// ---- The full search of listed MBs with each MB split over NPART
// workgroups by dy groups (GPP groups of DY candidate rows each), so that a
// short list (the learned search's fallback) still fills the chip.  Every
// part's workgroup minimum of the five keys is folded into keys[5 i ..
// 5 i + 4] (i = list position) with atomicMin, which gives the unsplit
// search's keys whatever the order; fs_finish turns keys into records.
template <int R>
struct FsSplit {
    using Gm = FsGeom<R>;
    static constexpr int GPP = 3;
    static constexpr int NPART = (Gm::NGY + GPP - 1) / GPP;
    static constexpr int NTH = (Gm::NG * GPP + 63) / 64 * 64;
    static constexpr int NR = GPP * DY + 15;
};

template <int R>
__device__ __forceinline__ void fullsearch_part(const uint8_t *__restrict__ cur, int cpitch,
                                                const uint8_t *__restrict__ ref, int rpitch, int wc, int hc, int mbw,
                                                int mb, int lambda8, int part, uint32_t *win, uint32_t (*red)[5],
                                                uint32_t *__restrict__ keys)
{
    using Gm = FsGeom<R>;
    using Sp = FsSplit<R>;
    const int x0 = (mb % mbw) * 16, y0 = (mb / mbw) * 16, tid = threadIdx.x;
    const int dyg0 = part * Sp::GPP, ng = min(Sp::GPP, Gm::NGY - dyg0);
    // the window rows of dy groups dyg0 .. dyg0 + ng - 1, edge-replicated
    const int wx0 = x0 - R, wy0 = y0 - R + dyg0 * DY, nrows = ng * DY + 15;
    for (int i = tid; i < nrows * Gm::WD; i += Sp::NTH) {
        const int r = i / Gm::WD, d = i - r * Gm::WD;
        const uint8_t *row = ref + (size_t)clampi(wy0 + r, 0, hc - 1) * rpitch;
        const int gx = wx0 + 4 * d;
        uint32_t v;
        if (gx >= 0 && gx + 3 < wc) {
            v = *(const uint32_t *)(row + gx);
        } else {
            v = (uint32_t)row[clampi(gx, 0, wc - 1)] | ((uint32_t)row[clampi(gx + 1, 0, wc - 1)] << 8) |
                ((uint32_t)row[clampi(gx + 2, 0, wc - 1)] << 16) | ((uint32_t)row[clampi(gx + 3, 0, wc - 1)] << 24);
        }
        win[r * Gm::PITCH + d] = v;
    }
    uint32_t cw[16][4];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const uint32_t *p = (const uint32_t *)(cur + (size_t)(y0 + r) * cpitch + x0);
#pragma unroll
        for (int c = 0; c < 4; c++) cw[r][c] = __builtin_amdgcn_readfirstlane(p[c]);
    }
    __syncthreads();
    uint32_t best16 = 0xffffffffu, bq[4] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu};
    if (tid < Gm::NG * ng) {
        const int dxg = tid % Gm::NG, dyl = tid / Gm::NG;
        fs_task<R>(win + (dyl * DY) * Gm::PITCH + dxg, cw, lambda8, -R + 4 * dxg, -R + DY * (dyg0 + dyl), best16, bq);
    }
    best16 = wave_min_u32(best16);
#pragma unroll
    for (int q = 0; q < 4; q++) bq[q] = wave_min_u32(bq[q]);
    const int wave = tid >> 6, lane = tid & 63;
    if (lane == 0) {
        red[wave][0] = best16;
#pragma unroll
        for (int q = 0; q < 4; q++) red[wave][1 + q] = bq[q];
    }
    __syncthreads();
    if (tid < 5) {
        uint32_t k = red[0][tid];
#pragma unroll
        for (int w = 1; w < Sp::NTH / 64; w++) k = min(k, red[w][tid]);
        atomicMin(keys + tid, k);
    }
}

template <int R>
__global__ __launch_bounds__(FsSplit<R>::NTH) void me_fullsearch_split_kernel(
    const uint8_t *__restrict__ cur, int cpitch, const uint8_t *__restrict__ ref, int rpitch, int wc, int hc,
    int mbw, int nmb, int lambda8, const int *__restrict__ list, const int *__restrict__ count,
    uint32_t *__restrict__ keys)
{
    __shared__ uint32_t win[FsSplit<R>::NR * FsGeom<R>::PITCH];
    __shared__ uint32_t red[FsSplit<R>::NTH / 64][5];
    const int n = min(*count, nmb);
    const int i = (int)blockIdx.x / FsSplit<R>::NPART, part = (int)blockIdx.x % FsSplit<R>::NPART;
    if (i >= n) return;
    const int mb = list[i];
    fullsearch_part<R>(cur, cpitch, ref, rpitch, wc, hc, mbw, mb, lambda8, part, win, red, keys + 5 * mb);
}

template <int R>
__global__ __launch_bounds__(FsSplit<R>::NTH) void me_fullsearch_split_batch_kernel(
    const bc250_me_batch_t b, const int *__restrict__ list, const int *__restrict__ count,
    uint32_t *__restrict__ keys)
{
    __shared__ uint32_t win[FsSplit<R>::NR * FsGeom<R>::PITCH];
    __shared__ uint32_t red[FsSplit<R>::NTH / 64][5];
    const int n = min(*count, b.nmb_total);
    const int i = (int)blockIdx.x / FsSplit<R>::NPART, part = (int)blockIdx.x % FsSplit<R>::NPART;
    if (i >= n) return;
    const int g = list[i];
    if (g < 0 || g >= b.nmb_total) return;
    const bc250_batch_frame_t &fr = b.f[batch_frame_of_mb(b, g)];
    fullsearch_part<R>(fr.cur, fr.cpitch, fr.ref, fr.rpitch, fr.wc, fr.hc, fr.wc / 16, g - fr.mb_base, b.lambda8,
                       part, win, red, keys + 5 * g);
}

__global__ void fs_keys_init_kernel(const int *__restrict__ list, const int *__restrict__ count, int cap,
                                    uint32_t *__restrict__ keys)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 5 * min(*count, cap)) keys[5 * list[i / 5] + i % 5] = 0xffffffffu;
}


__global__ void fs_finish_kernel(const int *__restrict__ list, const int *__restrict__ count, int cap,
                                 const uint32_t *__restrict__ keys, int range, int lambda8, uint4 *__restrict__ out16,
                                 uint4 *__restrict__ out8)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 5 * min(*count, cap)) return;
    const int e = i / 5, q = i % 5, mb = list[e];
    const uint4 o = fs_record(keys[5 * mb + q], q, range, lambda8);
    if (q == 0) out16[mb] = o;
    else out8[4 * mb + q - 1] = o;
}

__global__ void fs_finish_batch_kernel(const bc250_me_batch_t b, const int *__restrict__ list,
                                       const int *__restrict__ count, const uint32_t *__restrict__ keys)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 5 * min(*count, b.nmb_total)) return;
    const int e = i / 5, q = i % 5, g = list[e];
    if (g < 0 || g >= b.nmb_total) return;
    const bc250_batch_frame_t &fr = b.f[batch_frame_of_mb(b, g)];
    const int mb = g - fr.mb_base;
    const uint4 o = fs_record(keys[5 * g + q], q, b.range, b.lambda8);
    if (q == 0) ((uint4 *)fr.out16)[mb] = o;
    else ((uint4 *)fr.out8)[4 * mb + q - 1] = o;
}

// ---------------------------------------------------------------------------
// Sub-pel refinement: 16x16 then each 8x8 quadrant, one wave per MB.  For a
// block with integer winner (dx,dy): SAD at the integer position, then the 8
// half-pel candidates (2,0),(-2,0),(0,2),(0,-2),(2,2),(2,-2),(-2,2),(-2,-2)
// and the 8 quarter-pel ones around the half winner, strict < in that order.
// Per block the reference window (block + 4 on every side, replicated
// edges) is staged in LDS and its H.264 half-sample planes are built once:
// b (horizontal), h (vertical) and j (centre, from the unrounded horizontal
// intermediates), each over the (bw+2)^2 positions a +-3/4-pel candidate can
// touch.  A candidate's sample is then one or two LDS reads and an average.
namespace {
constexpr int RW = 24;       // window pitch (16 + 8)
constexpr int RP = 18;       // plane pitch (16 + 2)

struct RefineLds {
    uint8_t win[RW * RW];
    uint8_t cb[16 * 16];
    int16_t hr[23 * RP];     // unrounded horizontal 6-tap at window rows 1..23, cols 3..20
    uint8_t B[RP * RP], Hh[RP * RP], J[RP * RP];
};

// Sample at window integer position (xi, yi) and fraction (fx, fy), from
// the planes (whose origin is window position (3,3)).
__device__ __forceinline__ int sample_at(const RefineLds &L, int xi, int yi, int fx, int fy)
{
    const int px = xi - 3, py_ = yi - 3;
    const int c = fy * 4 + fx;
    const int g = L.win[yi * RW + xi];
    switch (c) {
    case 0: return g;
    case 1: return (g + L.B[py_ * RP + px] + 1) >> 1;
    case 2: return L.B[py_ * RP + px];
    case 3: return (L.win[yi * RW + xi + 1] + L.B[py_ * RP + px] + 1) >> 1;
    case 4: return (g + L.Hh[py_ * RP + px] + 1) >> 1;
    case 5: return (L.B[py_ * RP + px] + L.Hh[py_ * RP + px] + 1) >> 1;
    case 6: return (L.B[py_ * RP + px] + L.J[py_ * RP + px] + 1) >> 1;
    case 7: return (L.B[py_ * RP + px] + L.Hh[py_ * RP + px + 1] + 1) >> 1;
    case 8: return L.Hh[py_ * RP + px];
    case 9: return (L.Hh[py_ * RP + px] + L.J[py_ * RP + px] + 1) >> 1;
    case 10: return L.J[py_ * RP + px];
    case 11: return (L.J[py_ * RP + px] + L.Hh[py_ * RP + px + 1] + 1) >> 1;
    case 12: return (L.win[(yi + 1) * RW + xi] + L.Hh[py_ * RP + px] + 1) >> 1;
    case 13: return (L.Hh[py_ * RP + px] + L.B[(py_ + 1) * RP + px] + 1) >> 1;
    case 14: return (L.J[py_ * RP + px] + L.B[(py_ + 1) * RP + px] + 1) >> 1;
    default: return (L.Hh[py_ * RP + px + 1] + L.B[(py_ + 1) * RP + px] + 1) >> 1;
    }
}
}  // namespace

// One block of the refinement (BW = 16: the MB, 8: quadrant job - 1), its
// block size a constant so that the sample loops index by shifts.
template <int BW>
__device__ __forceinline__ void refine_job(RefineLds &L, const uint8_t *__restrict__ ref, int rpitch, int wc,
                                           int hc, int mb, int X0, int Y0, int job, uint4 iv,
                                           uint4 *__restrict__ mv16, uint4 *__restrict__ mv8, int lane)
{
    const int HO[8][2] = {{2, 0}, {-2, 0}, {0, 2}, {0, -2}, {2, 2}, {2, -2}, {-2, 2}, {-2, -2}};
    const int QO[8][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}, {1, 1}, {1, -1}, {-1, 1}, {-1, -1}};
    constexpr int LB = BW == 16 ? 4 : 3;
    const int bx0 = job == 0 ? 0 : ((job - 1) & 1) * 8, by0 = job == 0 ? 0 : ((job - 1) >> 1) * 8;
    const int dx = __builtin_amdgcn_readfirstlane((int)iv.x), dy = __builtin_amdgcn_readfirstlane((int)iv.y);
    constexpr int ww = BW + 8, np = BW + 2;
    wave_lds_sync();
    // window with replicated edges
    for (int i = lane; i < ww * ww; i += 64) {
        const int yy = i / ww, xx = i - yy * ww;
        const int gx = clampi(X0 + bx0 + dx - 4 + xx, 0, wc - 1);
        const int gy = clampi(Y0 + by0 + dy - 4 + yy, 0, hc - 1);
        L.win[yy * RW + xx] = ref[(size_t)gy * rpitch + gx];
    }
    wave_lds_sync();
    // unrounded horizontal intermediates at rows 1..ww-1, cols 3..3+np-1
    for (int i = lane; i < (ww - 1) * np; i += 64) {
        const int r = i / np, k = i - r * np;
        const uint8_t *p = &L.win[(1 + r) * RW + 3 + k];
        L.hr[r * RP + k] = (int16_t)tap6(p[-2], p[-1], p[0], p[1], p[2], p[3]);
    }
    wave_lds_sync();
    for (int i = lane; i < np * np; i += 64) {
        const int r = i / np, k = i - r * np;
        // b at window (3+k, 3+r) is hr row (3+r)-1
        L.B[r * RP + k] = (uint8_t)clip255((L.hr[(2 + r) * RP + k] + 16) >> 5);
        const uint8_t *q = &L.win[(3 + r) * RW + 3 + k];
        L.Hh[r * RP + k] = (uint8_t)clip255((tap6(q[-2 * RW], q[-RW], q[0], q[RW], q[2 * RW], q[3 * RW]) + 16) >> 5);
        const int16_t *t = &L.hr[r * RP + k];  // rows (3+r)-2 .. (3+r)+3 are hr rows r .. r+5
        const int j1 = tap6(t[0], t[RP], t[2 * RP], t[3 * RP], t[4 * RP], t[5 * RP]);
        L.J[r * RP + k] = (uint8_t)clip255((j1 + 512) >> 10);
    }
    wave_lds_sync();
    // (ox, oy) is the same in every lane: sample_at's switch is uniform
    auto sad_at = [&](int ox, int oy) -> uint32_t {
        uint32_t s = 0;
        const int fx = ox & 3, fy = oy & 3;
#pragma unroll
        for (int p = lane; p < BW * BW; p += 64) {
            const int px = p & (BW - 1), pyy = p >> LB;
            const int xi = 4 + px + (ox >> 2), yi = 4 + pyy + (oy >> 2);
            s += (uint32_t)iabs((int)L.cb[(by0 + pyy) * 16 + bx0 + px] - sample_at(L, xi, yi, fx, fy));
        }
        return s;
    };
    uint32_t best = wave_sum_u32(sad_at(0, 0));
    uint32_t hs[8];
#pragma unroll
    for (int c = 0; c < 8; c++) hs[c] = sad_at(HO[c][0], HO[c][1]);
#pragma unroll
    for (int c = 0; c < 8; c++) hs[c] = wave_sum_u32(hs[c]);
    int hx = 0, hy = 0;
#pragma unroll
    for (int c = 0; c < 8; c++)
        if (hs[c] < best) { best = hs[c]; hx = HO[c][0]; hy = HO[c][1]; }
    uint32_t qs[8];
#pragma unroll
    for (int c = 0; c < 8; c++) qs[c] = sad_at(hx + QO[c][0], hy + QO[c][1]);
#pragma unroll
    for (int c = 0; c < 8; c++) qs[c] = wave_sum_u32(qs[c]);
    int qx = 0, qy = 0;
#pragma unroll
    for (int c = 0; c < 8; c++)
        if (qs[c] < best) { best = qs[c]; qx = QO[c][0]; qy = QO[c][1]; }
    if (lane == 0) {
        uint4 o = make_uint4((uint32_t)(4 * dx + hx + qx), (uint32_t)(4 * dy + hy + qy), best, 0u);
        if (job == 0) mv16[mb] = o;
        else mv8[4 * mb + job - 1] = o;
    }
}

__global__ __launch_bounds__(64) void me_refine_kernel(
    const uint8_t *__restrict__ cur, int cpitch, const uint8_t *__restrict__ ref, int rpitch,
    int wc, int hc, int mbw, int nmb, const uint4 *__restrict__ in16, const uint4 *__restrict__ in8,
    uint4 *__restrict__ mv16, uint4 *__restrict__ mv8, int do8)
{
    __shared__ RefineLds L;
    const int mb = xcd_item(blockIdx.x, nmb);
    if (mb >= nmb) return;
    const int lane = threadIdx.x;
    const int mbx = mb % mbw, mby = mb / mbw;
    const int X0 = mbx * 16, Y0 = mby * 16;
    for (int i = lane; i < 64; i += 64) {
        const uint32_t v = *(const uint32_t *)(cur + (size_t)(Y0 + (i >> 2)) * cpitch + X0 + (i & 3) * 4);
        *(uint32_t *)&L.cb[(i >> 2) * 16 + (i & 3) * 4] = v;
    }
    refine_job<16>(L, ref, rpitch, wc, hc, mb, X0, Y0, 0, in16[mb], mv16, mv8, lane);
    if (do8)
        for (int job = 1; job < 5; job++)
            refine_job<8>(L, ref, rpitch, wc, hc, mb, X0, Y0, job, in8[4 * mb + job - 1], mv16, mv8, lane);
}

// Host-side launchers (declared in bc250_stages.h).
extern "C" int bc250_launch_me_fullsearch_list(const uint8_t *cur, int cpitch, const uint8_t *ref, int rpitch,
                                               int wc, int hc, int range, int lambda8, uint32_t *out16_,
                                               uint32_t *out8_, const int *list, const int *count, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    uint4 *out16 = (uint4 *)out16_, *out8 = (uint4 *)out8_;
    if (wc <= 0 || hc <= 0 || (wc & 15) || (hc & 15) || (cpitch & 3) || lambda8 < 0 || lambda8 > 15) return -1;
    if ((list == nullptr) != (count == nullptr)) return -1;
    const int mbw = wc / 16, nmb = mbw * (hc / 16);
    const int grid = ((nmb + 7) / 8) * 8;
    switch (range) {
    case 16:
        me_fullsearch_kernel<16><<<grid, ME_THREADS, 0, st>>>(cur, cpitch, ref, rpitch, wc, hc, mbw, nmb,
                                                             lambda8, out16, out8, list, count);
        break;
    case 32:
        me_fullsearch_kernel<32><<<grid, ME_THREADS, 0, st>>>(cur, cpitch, ref, rpitch, wc, hc, mbw, nmb,
                                                             lambda8, out16, out8, list, count);
        break;
    case 64:
        me_fullsearch_kernel<64><<<grid, ME_THREADS, 0, st>>>(cur, cpitch, ref, rpitch, wc, hc, mbw, nmb,
                                                             lambda8, out16, out8, list, count);
        break;
    default:
        return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int bc250_launch_me_fullsearch_list_split(const uint8_t *cur, int cpitch, const uint8_t *ref, int rpitch,
                                                     int wc, int hc, int range, int lambda8, uint32_t *out16,
                                                     uint32_t *out8, const int *list, const int *count,
                                                     uint32_t *keys, int keys_ready, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    if (wc <= 0 || hc <= 0 || (wc & 15) || (hc & 15) || (cpitch & 3) || lambda8 < 0 || lambda8 > 15 || !list ||
        !count || !keys || !out16 || !out8)
        return -1;
    const int mbw = wc / 16, nmb = mbw * (hc / 16);
    const int kb = (5 * nmb + 255) / 256;
    // keys_ready bit 0: the list's keys are set already; bit 1: no records
    // (the caller turns the keys into records itself)
    if (!(keys_ready & 1)) fs_keys_init_kernel<<<kb, 256, 0, st>>>(list, count, nmb, keys);
    switch (range) {
#define BC250_SPLIT(RR)                                                                                               \
    case RR:                                                                                                          \
        me_fullsearch_split_kernel<RR><<<FsSplit<RR>::NPART * nmb, FsSplit<RR>::NTH, 0, st>>>(                        \
            cur, cpitch, ref, rpitch, wc, hc, mbw, nmb, lambda8, list, count, keys);                                 \
        break;
        BC250_SPLIT(16)
        BC250_SPLIT(32)
        BC250_SPLIT(64)
#undef BC250_SPLIT
    default:
        return -1;
    }
    if (!(keys_ready & 2))
        fs_finish_kernel<<<kb, 256, 0, st>>>(list, count, nmb, keys, range, lambda8, (uint4 *)out16, (uint4 *)out8);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int bc250_launch_me_fullsearch_batch_split(const bc250_me_batch_t *b, const int *list, const int *count,
                                                      uint32_t *keys, int keys_ready, void *stream)
{
    if (!b || b->n < 1 || b->n > BC250_BATCH_MAX || b->nmb_total <= 0 || b->lambda8 < 0 || b->lambda8 > 15 ||
        !list || !count || !keys)
        return -1;
    hipStream_t st = (hipStream_t)stream;
    const int kb = (5 * b->nmb_total + 255) / 256;
    if (!keys_ready) fs_keys_init_kernel<<<kb, 256, 0, st>>>(list, count, b->nmb_total, keys);
    switch (b->range) {
#define BC250_SPLIT(RR)                                                                                               \
    case RR:                                                                                                          \
        me_fullsearch_split_batch_kernel<RR><<<FsSplit<RR>::NPART * b->nmb_total, FsSplit<RR>::NTH, 0, st>>>(        \
            *b, list, count, keys);                                                                                   \
        break;
        BC250_SPLIT(16)
        BC250_SPLIT(32)
        BC250_SPLIT(64)
#undef BC250_SPLIT
    default:
        return -1;
    }
    fs_finish_batch_kernel<<<kb, 256, 0, st>>>(*b, list, count, keys);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int bc250_launch_me_fullsearch(const uint8_t *cur, int cpitch, const uint8_t *ref, int rpitch,
                                          int wc, int hc, int range, int lambda8, uint32_t *out16,
                                          uint32_t *out8, void *stream)
{
    return bc250_launch_me_fullsearch_list(cur, cpitch, ref, rpitch, wc, hc, range, lambda8, out16, out8, nullptr,
                                           nullptr, stream);
}

extern "C" int bc250_launch_me_fullsearch_batch(const bc250_me_batch_t *b, const int *list, const int *count,
                                                void *stream)
{
    if (!b || b->n < 1 || b->n > BC250_BATCH_MAX || b->nmb_total <= 0 || b->lambda8 < 0 || b->lambda8 > 15 ||
        (list == nullptr) != (count == nullptr))
        return -1;
    const int grid = ((b->nmb_total + 7) / 8) * 8;
    hipStream_t st = (hipStream_t)stream;
    switch (b->range) {
    case 16: me_fullsearch_batch_kernel<16><<<grid, ME_THREADS, 0, st>>>(*b, list, count); break;
    case 32: me_fullsearch_batch_kernel<32><<<grid, ME_THREADS, 0, st>>>(*b, list, count); break;
    case 64: me_fullsearch_batch_kernel<64><<<grid, ME_THREADS, 0, st>>>(*b, list, count); break;
    default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int bc250_launch_me_refine(const uint8_t *cur, int cpitch, const uint8_t *ref, int rpitch, int wc,
                                      int hc, const uint32_t *in16, const uint32_t *in8, uint32_t *mv16,
                                      uint32_t *mv8, int do8, void *stream)
{
    if (wc <= 0 || hc <= 0 || (wc & 15) || (hc & 15)) return -1;
    const int mbw = wc / 16, nmb = mbw * (hc / 16);
    const int grid = ((nmb + 7) / 8) * 8;
    me_refine_kernel<<<grid, 64, 0, (hipStream_t)stream>>>(cur, cpitch, ref, rpitch, wc, hc, mbw, nmb,
                                                            (const uint4 *)in16, (const uint4 *)in8,
                                                            (uint4 *)mv16, (uint4 *)mv8, do8);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
