// bc250_dev.h - device-side helpers shared by the gfx950 kernels.
//
// Frame layout in HBM (coded size, 16-aligned):
//   luma   : wc x hc bytes, row pitch `py` (multiple of 256)
//   chroma : NV12 interleaved Cb/Cr, wc bytes x hc/2 rows, pitch `puv`
//   levels : int16 [nmb*24*16], block b of MB m at (m*24+b)*16; blocks
//            0..15 luma raster (by*4+bx), 16..19 Cb TL/TR/BL/BR, 20..23 Cr
//   dc     : int32 [nmb*24]; nz : uint32 [nmb*24]; modes : uint32 [nmb]
//   mvs    : {int32 mvx, mvy; uint32 sad, pad} [nmb], quarter-pel
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define BC_DEV static __device__ __forceinline__

BC_DEV int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
BC_DEV int clip255(int v) { return v < 0 ? 0 : (v > 255 ? 255 : v); }
BC_DEV int iabs(int v) { return v < 0 ? -v : v; }

// H.264 Table 8-15 (QPc from qPI).
BC_DEV int chroma_qp(int qpy)
{
    const int t[22] = {29, 30, 31, 32, 32, 33, 34, 34, 35, 35, 36,
                       36, 37, 37, 37, 38, 38, 38, 39, 39, 39, 39};
    qpy = clampi(qpy, 0, 51);
    return qpy < 30 ? qpy : t[qpy - 30];
}

// Raster position class: 0 both even, 1 both odd, 2 mixed.
BC_DEV int pos_class(int p)
{
    int r = (p >> 2) & 1, c = p & 1;
    return (r ^ c) ? 2 : r;
}

BC_DEV int mf_of(int rem, int cls)
{
    // H.264 forward multipliers MF[qp%6][class]
    const int t[18] = {13107, 5243, 8066, 11916, 4660, 7490, 10082, 4194, 6554,
                       9362,  3647, 5825, 8192,  3355, 5243, 7282,  2893, 4559};
    return t[rem * 3 + cls];
}
BC_DEV int v_of(int rem, int cls)
{
    // H.264 normAdjust4x4 V[qp%6][class]
    const int t[18] = {10, 16, 13, 11, 18, 14, 13, 20, 16, 14, 23, 18, 16, 25, 20, 18, 29, 23};
    return t[rem * 3 + cls];
}
BC_DEV int dc_ls0(int rem)
{
    const int t[6] = {160, 176, 208, 224, 256, 288};
    return t[rem];
}

BC_DEV void fwd4x4(const int in[16], int out[16])
{
    int m[16];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        int a0 = in[4 * i] + in[4 * i + 3], a1 = in[4 * i + 1] + in[4 * i + 2];
        int a2 = in[4 * i + 1] - in[4 * i + 2], a3 = in[4 * i] - in[4 * i + 3];
        m[4 * i] = a0 + a1;
        m[4 * i + 1] = 2 * a3 + a2;
        m[4 * i + 2] = a0 - a1;
        m[4 * i + 3] = a3 - 2 * a2;
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        int a0 = m[j] + m[12 + j], a1 = m[4 + j] + m[8 + j];
        int a2 = m[4 + j] - m[8 + j], a3 = m[j] - m[12 + j];
        out[j] = a0 + a1;
        out[4 + j] = 2 * a3 + a2;
        out[8 + j] = a0 - a1;
        out[12 + j] = a3 - 2 * a2;
    }
}

// Inverse core transform incl. (x+32)>>6.
BC_DEV void inv4x4(const int in[16], int out[16])
{
    int t[16];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        int z0 = in[4 * i] + in[4 * i + 2], z1 = in[4 * i] - in[4 * i + 2];
        int z2 = (in[4 * i + 1] >> 1) - in[4 * i + 3], z3 = in[4 * i + 1] + (in[4 * i + 3] >> 1);
        t[4 * i] = z0 + z3;
        t[4 * i + 1] = z1 + z2;
        t[4 * i + 2] = z1 - z2;
        t[4 * i + 3] = z0 - z3;
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        int z0 = t[j] + t[8 + j], z1 = t[j] - t[8 + j];
        int z2 = (t[4 + j] >> 1) - t[12 + j], z3 = t[4 + j] + (t[12 + j] >> 1);
        out[j] = (z0 + z3 + 32) >> 6;
        out[4 + j] = (z1 + z2 + 32) >> 6;
        out[8 + j] = (z1 - z2 + 32) >> 6;
        out[12 + j] = (z0 - z3 + 32) >> 6;
    }
}

BC_DEV int quant_ac(int c, int pos, int per, int rem, int f)
{
    int a = iabs(c);
    int l = (a * mf_of(rem, pos_class(pos)) + f) >> (15 + per);
    l = c < 0 ? -l : l;
    return clampi(l, -32768, 32767);
}

BC_DEV int dequant_ac(int level, int pos, int per, int rem)
{
    int s = 16 * v_of(rem, pos_class(pos));
    if (per >= 4) return (level * s) << (per - 4);
    return (level * s + (1 << (3 - per))) >> (4 - per);
}

// Round-to-nearest DC quantiser with threshold 5 (luma and chroma).
BC_DEV int quant_dc_k5(int v, int per, int rem)
{
    int ls = dc_ls0(rem);
    int av = iabs(v), num, den;
    if (per >= 5) { num = av; den = ls << (per - 5); }
    else          { num = av << (5 - per); den = ls; }
    int l = (num + den / 2) / den;
    return v < 0 ? -l : l;
}
BC_DEV int dequant_dc_luma(int f, int per, int rem)
{
    int ls = dc_ls0(rem);
    if (per >= 6) return (f * ls) << (per - 6);
    return (f * ls + (1 << (5 - per))) >> (6 - per);
}
BC_DEV int dequant_dc_chroma(int f, int per, int rem)
{
    int ls = dc_ls0(rem);
    if (per >= 5) return (f * ls) << (per - 5);
    return (f * ls) >> (5 - per);
}

BC_DEV void hadamard2x2(const int c[4], int o[4])
{
    o[0] = c[0] + c[1] + c[2] + c[3];
    o[1] = c[0] - c[1] + c[2] - c[3];
    o[2] = c[0] + c[1] - c[2] - c[3];
    o[3] = c[0] - c[1] - c[2] + c[3];
}

// Two-pass luma DC butterfly; `norm` selects the encoder-side (x+1)>>1.
BC_DEV void hadamard4x4(const int in[16], int out[16], bool norm)
{
    int t[16];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        int s01 = in[4 * i] + in[4 * i + 1], d01 = in[4 * i] - in[4 * i + 1];
        int s23 = in[4 * i + 2] + in[4 * i + 3], d23 = in[4 * i + 2] - in[4 * i + 3];
        t[i] = s01 + s23;
        t[4 + i] = s01 - s23;
        t[8 + i] = d01 - d23;
        t[12 + i] = d01 + d23;
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        int s01 = t[4 * i] + t[4 * i + 1], d01 = t[4 * i] - t[4 * i + 1];
        int s23 = t[4 * i + 2] + t[4 * i + 3], d23 = t[4 * i + 2] - t[4 * i + 3];
        int r = norm ? 1 : 0;
        out[4 * i] = (s01 + s23 + r) >> r;
        out[4 * i + 1] = (s01 - s23 + r) >> r;
        out[4 * i + 2] = (d01 - d23 + r) >> r;
        out[4 * i + 3] = (d01 + d23 + r) >> r;
    }
}

// H.264 8.4.2.2.1 luma sample at quarter-pel offset (fx, fy) from integer
// position (x, y) of a window `w` (row pitch `wp`) whose edges were already
// replicated from the coded frame, i.e. reads need no clamping.
BC_DEV int tap6(int a, int b, int c, int d, int e, int f) { return a - 5 * b + 20 * c + 20 * d - 5 * e + f; }

BC_DEV int qpel_win(const uint8_t *w, int wp, int x, int y, int fx, int fy)
{
#define GW(xx, yy) ((int)w[(yy) * wp + (xx)])
#define HR(xx, yy) tap6(GW(xx - 2, yy), GW(xx - 1, yy), GW(xx, yy), GW(xx + 1, yy), GW(xx + 2, yy), GW(xx + 3, yy))
#define VR(xx, yy) tap6(GW(xx, yy - 2), GW(xx, yy - 1), GW(xx, yy), GW(xx, yy + 1), GW(xx, yy + 2), GW(xx, yy + 3))
    int g = GW(x, y);
    int c = fy * 4 + fx;
    if (c == 0) return g;
    if (c == 2) return clip255((HR(x, y) + 16) >> 5);
    if (c == 8) return clip255((VR(x, y) + 16) >> 5);
    if (c == 1 || c == 3) {
        int b = clip255((HR(x, y) + 16) >> 5);
        return ((c == 1 ? g : GW(x + 1, y)) + b + 1) >> 1;
    }
    if (c == 4 || c == 12) {
        int h = clip255((VR(x, y) + 16) >> 5);
        return ((c == 4 ? g : GW(x, y + 1)) + h + 1) >> 1;
    }
    if (c == 5 || c == 7 || c == 13 || c == 15) {
        // diagonal: average of a horizontal and a vertical half sample
        int hy = (c >= 13) ? y + 1 : y;
        int vx = (c == 7 || c == 15) ? x + 1 : x;
        int b = clip255((HR(x, hy) + 16) >> 5);
        int v = clip255((VR(vx, y) + 16) >> 5);
        return (b + v + 1) >> 1;
    }
    // positions using j: f (6), i (9), j (10), k (11), q (14)
    int j1 = tap6(HR(x, y - 2), HR(x, y - 1), HR(x, y), HR(x, y + 1), HR(x, y + 2), HR(x, y + 3));
    int j = clip255((j1 + 512) >> 10);
    if (c == 10) return j;
    int o;
    if (c == 6) o = clip255((HR(x, y) + 16) >> 5);            // b
    else if (c == 14) o = clip255((HR(x, y + 1) + 16) >> 5);  // s
    else if (c == 9) o = clip255((VR(x, y) + 16) >> 5);       // h
    else o = clip255((VR(x + 1, y) + 16) >> 5);               // m (11)
    return (j + o + 1) >> 1;
#undef GW
#undef HR
#undef VR
}

// Wave-wide reductions (64 lanes, all active): DPP steps inside every row of
// 16 lanes (quad_perm [1,0,3,2] and [2,3,0,1], row_half_mirror, row_mirror
// leave the row's result in each of its lanes), then the four row results
// read out as scalars.  No LDS round trips.
BC_DEV uint32_t wave_sum_u32(uint32_t v)
{
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, false);
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 0) + (uint32_t)__builtin_amdgcn_readlane((int)v, 16) +
           (uint32_t)__builtin_amdgcn_readlane((int)v, 32) + (uint32_t)__builtin_amdgcn_readlane((int)v, 48);
}
BC_DEV uint32_t wave_min_u32(uint32_t v)
{
    uint32_t w = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);
    v = w < v ? w : v;
    w = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);
    v = w < v ? w : v;
    w = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false);
    v = w < v ? w : v;
    w = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, false);
    v = w < v ? w : v;
    const uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)v, 0), b = (uint32_t)__builtin_amdgcn_readlane((int)v, 16);
    const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)v, 32), d = (uint32_t)__builtin_amdgcn_readlane((int)v, 48);
    const uint32_t ab = a < b ? a : b, cd = c < d ? c : d;
    return ab < cd ? ab : cd;
}

// Device-scope progress flags for the wavefront kernels.  The producer
// publishes with a release store after its pixel stores; the consumer
// acquires before reading them.  Waits are bounded: past `limit` polls
// the kernel sets *err and proceeds, so a broken dependency can never
// hang the device.
BC_DEV int flag_load_acquire(const int *p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
BC_DEV void flag_store_release(int *p, int v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
BC_DEV void wait_at_least(const int *p, int want, int *err, unsigned limit)
{
    unsigned n = 0;
    while (flag_load_acquire(p) < want) {
        __builtin_amdgcn_s_sleep(1);
        if (++n > limit) {
            atomicOr(err, 1);
            break;
        }
    }
}

// Orders one wave's LDS traffic around a hand-off: LDS executes a wave's
// instructions in order, the clobber stops the compiler from moving them.
BC_DEV void wave_lds_sync()
{
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_wave_barrier();
}

// System-coherent (sc0 sc1) accesses for hand-offs between workgroups that
// may run on different XCDs: they bypass the non-coherent caches, so no L2
// writeback / invalidate is needed around them.
BC_DEV uint32_t ld_sys(const uint32_t *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM); }
BC_DEV void st_sys(uint32_t *p, uint32_t v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM); }
// 64-bit forms: an aligned 8-byte access is single-copy atomic, so a record
// {tag, value} written by one store is seen whole or not at all.
BC_DEV uint64_t ld_sys64(const uint64_t *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM); }
BC_DEV void st_sys64(uint64_t *p, uint64_t v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM); }

// Waits until every store this wave issued has completed.
BC_DEV void wait_stores() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// Launch epochs of the wavefront kernels.  Every hand-off record and
// progress word a wavefront kernel publishes carries the 16-bit epoch of its
// launch in its high half, (epoch << 16) | step, so a word left in memory by
// an earlier launch -- the previous frame's hand-off at the same position --
// never matches what a reader of this launch waits for, whatever the order
// in which earlier clears and stores reach memory.
BC_DEV uint32_t wf_tag(uint32_t epoch, int step) { return (epoch << 16) | (uint32_t)step; }

namespace {
// Runs on the launch stream just before a wavefront kernel: zeroes its
// ticket, error flag and progress words (`n` ints at `zero`) and advances
// the launch epoch at *epoch (1..65535, never 0), all with system-scope
// stores, so that no plain store of these words is left in an L2 of one XCD
// while workgroups on another XCD poll them.
__global__ void wavefront_begin_kernel(int *zero, int n, uint32_t *epoch)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        __hip_atomic_store(zero + i, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    if (threadIdx.x == 0) {
        uint32_t e = (__hip_atomic_load(epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) + 1) & 0xffffu;
        __hip_atomic_store(epoch, e ? e : 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}
}  // namespace
// The epoch a wavefront kernel tags its words with (written by
// wavefront_begin_kernel, which completed before the kernel started).
BC_DEV uint32_t wf_epoch(const uint32_t *epoch) { return ld_sys(epoch); }

// Slice of MB `mb` for start(s) = floor(s*total/ns).
BC_DEV int slice_of(int mb, int total, int ns)
{
    if (ns <= 1) return 0;
    int s = (int)(((long long)mb * ns) / total);
    if (s >= ns) s = ns - 1;
    while (s + 1 < ns && ((long long)(s + 1) * total) / ns <= mb) s++;
    while (s > 0 && ((long long)s * total) / ns > mb) s--;
    return s;
}

// XCD-aware mapping: consecutive workgroup ids land on different XCDs
// (round robin over 8), so give each XCD a contiguous run of items.
BC_DEV int xcd_item(int bid, int nitems)
{
    int per = (nitems + 7) >> 3;
    return (bid & 7) * per + (bid >> 3);
}
