// h264_sqme.hip - the reference's own two motion searches as gfx950 kernels,
// beside the full search of me_kernels.hip:
//   * the Tier-2 square search (the CPU SIMD route of src/cpu_simd_me.c:
//     275-497, whose vectors the H.264 encoder hands to dispatch_encode_ext
//     as cpu_mvs), bit-exact;
//   * the diamond search of shaders/motion_estimation.comp:278-469 with exact
//     integer SADs in place of the shader's float-normalised ones.
//
// Square search:
// The search is serial along an MB row (each MB starts from its left
// neighbour's vector), so one wave runs one MB row and rows run in parallel.
// Per MB the wave stages the reference window the search can touch
// ((16 + 2R + 2)^2 bytes, edge-clamped) in LDS; the current MB lives in
// registers, two rows per lane of each 8-lane group.  A SAD is one 8-lane
// group's work (2 rows x 4 dwords of v_sad_u8 per lane, then a DPP sum), so
// the 8 candidates of a square step, or the 4 half-sample candidates, are
// evaluated at once by the 8 groups and chosen in the reference's order with
// strict '<' by wave-uniform code.
#include "bc250_dev.h"
#include "bc250_stages.h"

namespace {
constexpr int RMAX = 16;
constexpr int SMAX = 16 + 2 * RMAX + 2;  // window side
constexpr int SP = 56;                   // window pitch in bytes (>= side + 4, a multiple of 4)

// Sum over the 8 lanes of each lane group: DPP quad_perm [1,0,3,2], [2,3,0,1],
// then the half-row mirror (lane i <-> 7 - i within 8).
BC_DEV uint32_t group8_sum(uint32_t v)
{
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false);
    return v;
}

// 16 window bytes at byte offset o as 4 dwords.
BC_DEV void row16(const uint32_t *win, int o, uint32_t s[4])
{
    const uint32_t *p = win + (o >> 2);
    const uint32_t sh = (uint32_t)(o & 3);
    uint32_t d[5];
#pragma unroll
    for (int i = 0; i < 5; i++) d[i] = p[i];
#pragma unroll
    for (int c = 0; c < 4; c++) s[c] = __builtin_amdgcn_alignbyte(d[c + 1], d[c], sh);
}

// rounded-up average of packed bytes: (a + b + 1) >> 1 per byte
BC_DEV uint32_t avg_up(uint32_t a, uint32_t b) { return (a | b) - (((a ^ b) >> 1) & 0x7F7F7F7Fu); }

// the square pattern (0,1) (0,-1) (1,0) (-1,0) (1,1) (-1,-1) (1,-1) (-1,1)
// and the half-sample offsets (2,0) (-2,0) (0,2) (0,-2), as functions of a
// possibly lane-dependent index (no private arrays)
BC_DEV int pat_x(int k) { return k < 2 ? 0 : ((k & 1) ? -1 : 1); }
BC_DEV int pat_y(int k) { return k < 2 ? ((k & 1) ? -1 : 1) : (k < 4 ? 0 : ((k == 4 || k == 7) ? 1 : -1)); }
BC_DEV int hp_x(int h) { return h == 0 ? 2 : (h == 1 ? -2 : 0); }
BC_DEV int hp_y(int h) { return h == 2 ? 2 : (h == 3 ? -2 : 0); }
}  // namespace

__global__ __launch_bounds__(64) void h264_square_me_kernel(const uint8_t *__restrict__ cur, int cpitch,
                                                            const uint8_t *__restrict__ ref, int rpitch, int wmb,
                                                            int hmb, int R, uint4 *__restrict__ out)
{
    __shared__ uint32_t win[SMAX * SP / 4 + 2];
    uint8_t *wb = (uint8_t *)win;
    const int mby = blockIdx.x, lane = threadIdx.x;
    if (mby >= hmb) return;
    const int grp = lane >> 3, gi = lane & 7;
    const int S = 16 + 2 * R + 2, max_x = wmb * 16, max_y = hmb * 16, py = mby * 16;
    int lmx = 0, lmy = 0;  // the left neighbour's result (quarter samples)
    for (int mbx = 0; mbx < wmb; mbx++) {
        const int px = mbx * 16;
        // ---- window at (px - R - 1, py - R - 1), and this lane's two current rows
        wave_lds_sync();
        for (int t = lane; t < S * S; t += 64) {
            const int wy = t / S, wx = t - wy * S;
            wb[wy * SP + wx] = ref[(size_t)clampi(py - R - 1 + wy, 0, max_y - 1) * rpitch +
                                   clampi(px - R - 1 + wx, 0, max_x - 1)];
        }
        uint32_t cw[2][4];
#pragma unroll
        for (int rr = 0; rr < 2; rr++)
#pragma unroll
            for (int c = 0; c < 4; c++)
                cw[rr][c] = *(const uint32_t *)(cur + (size_t)(py + 2 * gi + rr) * cpitch + px + 4 * c);
        wave_lds_sync();
        // SAD of the block at vector (x, y): this group's 8 lanes, every lane of the group
        auto sad_at = [&](int x, int y) -> uint32_t {
            uint32_t a = 0;
#pragma unroll
            for (int rr = 0; rr < 2; rr++) {
                uint32_t s[4];
                row16(win, (y + R + 1 + 2 * gi + rr) * SP + x + R + 1, s);
#pragma unroll
                for (int c = 0; c < 4; c++) a = __builtin_amdgcn_sad_u8(s[c], cw[rr][c], a);
            }
            return group8_sum(a);
        };
        auto sad_half = [&](int x, int y, int hx, int hy) -> uint32_t {
            uint32_t a = 0;
#pragma unroll
            for (int rr = 0; rr < 2; rr++) {
                uint32_t s0[4], s1[4];
                const int o = (y + R + 1 + 2 * gi + rr) * SP + x + R + 1;
                row16(win, o, s0);
                row16(win, o + hy * SP + hx, s1);
#pragma unroll
                for (int c = 0; c < 4; c++) a = __builtin_amdgcn_sad_u8(avg_up(s0[c], s1[c]), cw[rr][c], a);
            }
            return group8_sum(a);
        };
        auto inside = [&](int x, int y) {
            return px + x >= 0 && px + x + 16 <= max_x && py + y >= 0 && py + y + 16 <= max_y;
        };

        const uint32_t zero = (uint32_t)__builtin_amdgcn_readfirstlane((int)sad_at(0, 0));
        int rmx = 0, rmy = 0;
        uint32_t rcost = zero;
        if (zero > 512) {
            int bx = 0, by = 0;
            uint32_t best = zero;
            bool done = false;
            if (mbx > 0) {
                const int lx = lmx / 4, ly = lmy / 4;
                if ((lx || ly) && iabs(lx) <= R && iabs(ly) <= R && inside(lx, ly)) {
                    const uint32_t c = (uint32_t)__builtin_amdgcn_readfirstlane((int)sad_at(lx, ly)) +
                                       5u * (uint32_t)(iabs(lx) + iabs(ly));
                    if (c < best) {
                        best = c;
                        bx = lx;
                        by = ly;
                    }
                }
            }
            if ((bx || by) && best < 768) {
                rmx = bx * 4;
                rmy = by * 4;
                rcost = best;
                done = true;
            }
            if (!done) {
                for (int step = R / 2; step >= 1; step /= 2) {
                    const int cx = bx, cy = by;
                    // group g evaluates candidate g
                    const int gx = cx + pat_x(grp) * step, gy = cy + pat_y(grp) * step;
                    const bool ok = iabs(gx) <= R && iabs(gy) <= R && inside(gx, gy);
                    const uint32_t sv = sad_at(ok ? gx : 0, ok ? gy : 0);  // every lane, whole groups
                    const uint32_t mine = ok ? sv + 5u * (uint32_t)(iabs(gx) + iabs(gy)) : 0xffffffffu;
#pragma unroll
                    for (int k = 0; k < 8; k++) {
                        const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)mine, 8 * k);
                        if (c < best) {
                            best = c;
                            bx = cx + pat_x(k) * step;
                            by = cy + pat_y(k) * step;
                        }
                    }
                    if (best == 0) break;
                }
                int qx = bx * 4, qy = by * 4;
                if (best > 128 && best < 3000) {
                    const int ox = px + bx, oy = py + by;
                    if (ox > 0 && ox + 17 < max_x && oy > 0 && oy + 17 < max_y) {
                        const int h = grp & 3;
                        const int hx = hp_x(h) / 2, hy = hp_y(h) / 2;
                        const uint32_t mine = sad_half(bx, by, hx, hy);
                        int sx = 0, sy = 0;
                        uint32_t low = best;
#pragma unroll
                        for (int k = 0; k < 4; k++) {
                            const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)mine, 8 * k);
                            if (c < low) {
                                low = c;
                                sx = hp_x(k);
                                sy = hp_y(k);
                            }
                        }
                        qx += sx;
                        qy += sy;
                        best = low;
                    }
                }
                rmx = qx;
                rmy = qy;
                rcost = best;
            }
        }
        if (lane == 0) out[mby * wmb + mbx] = make_uint4((uint32_t)rmx, (uint32_t)rmy, rcost, 0u);
        lmx = rmx;
        lmy = rmy;
    }
}

extern "C" int bc250_launch_square_me(const uint8_t *cur, int cpitch, const uint8_t *ref, int rpitch, int width,
                                      int height, int search_radius, uint32_t *out_mvs, void *stream)
{
    if (width <= 0 || height <= 0 || (cpitch & 3)) return -1;
    const int wmb = (width + 15) / 16, hmb = (height + 15) / 16;
    if (cpitch < wmb * 16 || rpitch < wmb * 16) return -1;
    const int R = search_radius < 2 ? 2 : (search_radius > RMAX ? RMAX : search_radius);
    h264_square_me_kernel<<<hmb, 64, 0, (hipStream_t)stream>>>(cur, cpitch, ref, rpitch, wmb, hmb, R,
                                                               (uint4 *)out_mvs);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ---------------------------------------------------------------------------
// Diamond search (motion_estimation.comp:278-469): one wave per MB, 4 MBs
// per workgroup, every MB independent.  The window the search and its
// sub-sample refinement can touch, (x0 - 20 .. x0 + 35)^2 with replicated
// edges, is staged in LDS; lane l owns pixels (l / 4, 4 (l % 4) .. + 3).
// Integer candidates: 2 x the SAD over pixels with even x + y (a byte mask
// per row parity) + lambda |mv|; zero-vector check on the full SAD; strict
// '<' throughout, in the shader's order.
namespace {
constexpr int DW = 56;  // diamond window side and pitch
}

__global__ __launch_bounds__(256) void h264_diamond_me_kernel(const uint8_t *__restrict__ cur, int cpitch,
                                                              const uint8_t *__restrict__ ref, int rpitch, int wc,
                                                              int hc, int nmb, int lambda, uint4 *__restrict__ out)
{
    __shared__ uint32_t wins[4][DW * DW / 4 + 2];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int mb = blockIdx.x * 4 + wv;
    if (mb >= nmb) return;
    const int mbw = wc / 16, x0 = (mb % mbw) * 16, y0 = (mb / mbw) * 16;
    uint32_t *win = wins[wv];
    uint8_t *wb = (uint8_t *)win;
    for (int t = lane; t < DW * DW; t += 64) {
        const int wy = t / DW, wx = t - wy * DW;
        wb[t] = ref[(size_t)clampi(y0 - 20 + wy, 0, hc - 1) * rpitch + clampi(x0 - 20 + wx, 0, wc - 1)];
    }
    const int r = lane >> 2, c0 = (lane & 3) * 4;
    const uint32_t cw = *(const uint32_t *)(cur + (size_t)(y0 + r) * cpitch + x0 + c0);
    const uint32_t even = (r & 1) ? 0xFF00FF00u : 0x00FF00FFu;  // bytes with even x + y
    wave_lds_sync();
    auto ref4 = [&](int x, int y) -> uint32_t {
        const int o = (r + 20 + y) * DW + c0 + 20 + x;
        return __builtin_amdgcn_alignbyte(win[(o >> 2) + 1], win[o >> 2], (uint32_t)(o & 3));
    };
    const uint32_t zero = wave_sum_u32(__builtin_amdgcn_sad_u8(ref4(0, 0), cw, 0));
    uint32_t best = zero;
    int qx = 0, qy = 0;
    if (zero > 512) {
        int bx = 0, by = 0;
        for (int step = 8; step >= 1; step /= 2) {
            const int cx = bx, cy = by;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int x = cx + (k < 2 ? 0 : (k == 2 ? step : -step));
                const int y = cy + (k == 0 ? step : (k == 1 ? -step : 0));
                if (iabs(x) > 16 || iabs(y) > 16) continue;  // wave-uniform
                const uint32_t s = 2u * wave_sum_u32(__builtin_amdgcn_sad_u8(ref4(x, y) & even, cw & even, 0));
                const uint32_t cost = s + (uint32_t)lambda * (uint32_t)(iabs(x) + iabs(y));
                if (cost < best) {
                    best = cost;
                    bx = x;
                    by = y;
                }
            }
            if (best < 768) break;
        }
        qx = bx * 4;
        qy = by * 4;
        // half, then quarter samples around the running winner
        for (int stage = 0; stage < 2; stage++) {
            const int ax = qx, ay = qy;
            int sx = 0, sy = 0;
#pragma unroll 1
            for (int k = 0; k < 8; k++) {
                const int hx = k < 4 ? (k == 0 ? 2 : (k == 1 ? -2 : 0)) : (k < 6 ? 2 : -2);
                const int hy = k < 4 ? (k == 2 ? 2 : (k == 3 ? -2 : 0)) : ((k & 1) ? -2 : 2);
                const int ox = stage ? hx / 2 : hx, oy = stage ? hy / 2 : hy;
                const int vx = ax + ox, vy = ay + oy;
                uint32_t d = 0;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int p = qpel_win(wb, DW, c0 + i + 20 + (vx >> 2), r + 20 + (vy >> 2), vx & 3, vy & 3);
                    d += (uint32_t)iabs((int)((cw >> (8 * i)) & 255u) - p);
                }
                const uint32_t s = wave_sum_u32(d);
                if (s < best) {
                    best = s;
                    sx = ox;
                    sy = oy;
                }
            }
            qx = ax + sx;
            qy = ay + sy;
        }
    }
    if (lane == 0) out[mb] = make_uint4((uint32_t)qx, (uint32_t)qy, best, 0u);
}

extern "C" int bc250_launch_diamond_me(const uint8_t *cur, int cpitch, const uint8_t *ref, int rpitch, int wc,
                                       int hc, int lambda, uint32_t *out_mvs, void *stream)
{
    if (wc <= 0 || hc <= 0 || (wc & 15) || (hc & 15) || (cpitch & 3) || cpitch < wc || rpitch < wc || lambda < 0 ||
        lambda > 255)
        return -1;
    const int nmb = (wc / 16) * (hc / 16);
    h264_diamond_me_kernel<<<(nmb + 3) / 4, 256, 0, (hipStream_t)stream>>>(cur, cpitch, ref, rpitch, wc, hc, nmb, lambda,
                                                                    (uint4 *)out_mvs);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ---------------------------------------------------------------------------
// The reference-named Tier-2 interface (bc250_cpu_me.h).
#include <stdlib.h>
#include <string.h>

#include "bc250_cpu_me.h"

extern "C" uint32_t cpu_simd_sad_16x16(const uint8_t *src, int src_stride, const uint8_t *ref, int ref_stride)
{
    uint32_t s = 0;
    for (int r = 0; r < 16; r++)
        for (int c = 0; c < 16; c++) s += (uint32_t)abs((int)src[r * src_stride + c] - (int)ref[r * ref_stride + c]);
    return s;
}

extern "C" void cpu_simd_me_config_init(cpu_simd_me_config_t *cfg, uint32_t width, uint32_t height)
{
    if (!cfg) return;
    cfg->width = width;
    cfg->height = height;
    cfg->width_in_mbs = (width + 15) / 16;
    cfg->height_in_mbs = (height + 15) / 16;
    cfg->num_threads = 4;
    const char *names[3] = {"BC250_MAX_CPU_THREADS", "BC250_THREADS", "BC250_CPU_THREADS"};
    for (const char *n : names) {
        const char *e = getenv(n);
        if (e && *e) {
            const int t = atoi(e);
            if (t > 0 && t <= 16) cfg->num_threads = t;
            break;
        }
    }
    cfg->core_ids[0] = cfg->core_ids[1] = -1;
}

extern "C" int cpu_simd_me_search_frame(const uint8_t *src_y, int src_pitch, const uint8_t *ref_y, int ref_pitch,
                                        uint32_t width, uint32_t height, gpu_mv_t *out_mvs,
                                        const cpu_simd_me_config_t *cfg)
{
    if (!src_y || !ref_y || !out_mvs || width == 0 || height == 0) return -1;
    const int wmb = (int)((width + 15) / 16), hmb = (int)((height + 15) / 16);
    const int wp = wmb * 16, hp = hmb * 16;
    if (src_pitch < wp || ref_pitch < wp) return -1;
    const int radius = cfg ? (int)cfg->search_radius : 8;
    uint8_t *d = nullptr;
    const size_t plane = (size_t)wp * hp, mvb = (size_t)wmb * hmb * sizeof(gpu_mv_t);
    if (hipMalloc(&d, 2 * plane + mvb) != hipSuccess) return -1;
    hipStream_t st = nullptr;
    bool ok = hipStreamCreate(&st) == hipSuccess &&
              hipMemcpy2DAsync(d, wp, src_y, src_pitch, wp, hp, hipMemcpyHostToDevice, st) == hipSuccess &&
              hipMemcpy2DAsync(d + plane, wp, ref_y, ref_pitch, wp, hp, hipMemcpyHostToDevice, st) == hipSuccess &&
              bc250_launch_square_me(d, wp, d + plane, wp, (int)width, (int)height, radius,
                                     (uint32_t *)(d + 2 * plane), st) == 0 &&
              hipMemcpyAsync(out_mvs, d + 2 * plane, mvb, hipMemcpyDeviceToHost, st) == hipSuccess &&
              hipStreamSynchronize(st) == hipSuccess;
    if (st) hipStreamDestroy(st);
    hipFree(d);
    return ok ? 0 : -1;
}
