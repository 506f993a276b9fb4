// h264_intra.hip - I-slice encode as ONE persistent wavefront kernel for
// gfx950 (replaces the reference's 187 per-diagonal dispatches at 1080p).
//
// A workgroup is 4 waves; wave w encodes MB row 4*g + w of the row group g
// it took from an atomic ticket (so a row's producer always started before
// its consumer), left to right.  MB (x,y) needs the reconstruction of MB
// (x,y-1) (top row), of MB (x-1,y-1) (corner) and of MB (x-1,y) (left
// column, produced by the same wave and kept in LDS).  When a wave has
// reconstructed MB x it hands that MB's bottom luma row and bottom chroma
// row (32 bytes) to the wave below: through an LDS ring inside the
// workgroup, or through a system-coherent record in global memory to the
// next workgroup, published by a progress counter after an s_waitcnt.  The
// counter carries the launch epoch in its high half, so a counter left at
// "row done" by the previous I picture never releases a reader early.  No
// cache writeback or invalidate is needed and the frame itself is only
// written, never read back, during the kernel.
// Per MB: I16x16 SAD mode decision (DC, V, H, Plane, strict <), residual,
// forward 4x4, intra AC quantisation of positions 1..15, luma DC Hadamard /
// quant / inverse / dequant, reconstruction; chroma DC prediction per
// quadrant with the same chain.  Waits are bounded and report on *err.
#include "bc250_dev.h"
#include "bc250_stages.h"

namespace {
constexpr int WAVES = 4;
constexpr int RING = 16;
constexpr int HAND = 8;  // dwords per MB hand-off: luma row 15 (16 B), chroma row 7 (16 B, Cb/Cr interleaved)

struct IntraLds {
    int res[16][16];  // luma residual [block][pos]
    int pred[16][16];
    int acl[16][16];  // luma AC levels
    int cres[8][16], cacl[8][16];  // chroma residual / AC levels [comp*4+quadrant][pos]
    int dcin[16], dcl[16], dcv[16];  // luma DC: coefficients, levels, dequantised values
    int cdin[8], cdl[8], cdcv[8];    // chroma DC, [comp*4+quadrant]
    int cpred[2][4];
    uint32_t top[HAND];   // hand-off received for the current MB
    uint32_t out[HAND];   // hand-off produced by the current MB
    uint8_t leftY[16];
    uint8_t leftC[2][8];
};

// Sign of the Hadamard basis used by the luma DC butterflies:
// rows ++++, ++--, +--+, +-+-.
__device__ __forceinline__ int hsign(int a, int b)
{
    const int t = (0xF | (0x3 << 4) | (0x9 << 8) | (0x5 << 12)) >> (4 * a);
    return ((t >> b) & 1) ? 1 : -1;
}
// Coefficient L = 4r + k of the encoder's two-pass luma DC butterfly:
// sum over (i, j) of H[k][i] * H[r][j] * in[4i + j].
__device__ __forceinline__ int hadamard_coef(const int *in, int L)
{
    const int r = L >> 2, k = L & 3;
    int s = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        int row = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) row += hsign(r, j) * in[4 * i + j];
        s += hsign(k, i) * row;
    }
    return s;
}
// Coefficient q of the 2x2 chroma DC Hadamard: rows ++++, +-+-, ++--, +--+.
__device__ __forceinline__ int hadamard2_coef(const int *c, int q)
{
    const int s1 = (q & 1) ? -1 : 1, s2 = (q & 2) ? -1 : 1;
    return c[0] + s1 * c[1] + s2 * c[2] + s1 * s2 * c[3];
}
}  // namespace

__global__ __launch_bounds__(256) void h264_i_encode_kernel(
    const uint8_t *__restrict__ cy, const uint8_t *__restrict__ cuv, int cpy, int cpuv,
    uint8_t *__restrict__ oy, uint8_t *__restrict__ ouv, int opy, int opuv, int mbw, int mbh, int qp,
    int num_slices, int16_t *__restrict__ levels, int32_t *__restrict__ dc, uint32_t *__restrict__ nz,
    uint32_t *__restrict__ modes, int *sync /* [0] ticket, [1] err, [2..2+mbh) progress, [2+mbh] epoch */,
    uint32_t *ghand /* [mbh][mbw][HAND] */, unsigned spin_limit)
{
    __shared__ IntraLds wl[WAVES];
    __shared__ uint32_t ring[WAVES - 1][RING][HAND];
    __shared__ int lprog[WAVES], lcons[WAVES];
    __shared__ int s_group;

    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    int *err = sync + 1, *gprog = sync + 2;
    const uint32_t ep = wf_epoch((const uint32_t *)sync + 2 + mbh);
    if (tid == 0) s_group = atomicAdd(sync, 1);
    if (tid < WAVES) { lprog[tid] = 0; lcons[tid] = 0; }
    __syncthreads();
    const int y = s_group * WAVES + w;
    if (y >= mbh) return;
    IntraLds &S = wl[w];
    volatile int *vlprog = lprog, *vlcons = lcons;
    const bool last_row = y == mbh - 1;
    const bool lds_below = w < WAVES - 1 && !last_row;
    const bool lds_above = w > 0;
    const int total = mbw * mbh;
    const int qpc = chroma_qp(qp);
    const int per = qp / 6, rem = qp % 6, fi = (1 << (15 + per)) / 3;
    const int perc = qpc / 6, remc = qpc % 6, fic = (1 << (15 + perc)) / 3;

    // lane -> one 4-pixel row of one luma 4x4 block
    const int lr = lane >> 2, lc0 = (lane & 3) * 4;
    const int lb = (lr >> 2) * 4 + (lane & 3), lk = (lr & 3) * 4;
    // lanes 0..31 -> one 4-sample row of one chroma 4x4 block
    const int ccomp = (lane >> 4) & 1, csq = (lane >> 2) & 3, crow = lane & 3;
    int corner_next = 128;  // bottom-right luma sample of MB (x-1, y-1)

    // one-MB-ahead prefetch of the source samples
    uint32_t pre_y = 0;
    uint2 pre_c = make_uint2(0, 0);
    auto prefetch = [&](int x) {
        pre_y = *(const uint32_t *)(cy + (size_t)(y * 16 + lr) * cpy + x * 16 + lc0);
        if (lane < 32)
            pre_c = *(const uint2 *)(cuv + (size_t)(y * 8 + (csq >> 1) * 4 + crow) * cpuv + 2 * (x * 8 + (csq & 1) * 4));
    };
    prefetch(0);

    for (int x = 0; x < mbw; x++) {
        const int m = y * mbw + x;
        const int X0 = x * 16, Y0 = y * 16, U0 = x * 8, V0 = y * 8;
        const uint32_t cw = pre_y;
        const uint2 cw2 = pre_c;
        if (x + 1 < mbw) prefetch(x + 1);
        // ---- hand-off from the row above: MB (x, y-1) must be reconstructed
        if (y > 0) {
            uint32_t h = 0;
            if (lds_above) {
                unsigned n = 0;
                while (vlprog[w - 1] < x + 1) {
                    __builtin_amdgcn_s_sleep(1);
                    if (++n > spin_limit) { atomicOr(err, 1); break; }
                }
                wave_lds_sync();
                if (lane < HAND) h = ring[w - 1][x % RING][lane];
                wave_lds_sync();
                if (lane == 0) vlcons[w] = x + 1;
            } else {
                if (lane == 0) {
                    // progress words of this launch are (ep << 16) | (MBs done)
                    unsigned n = 0;
                    for (;;) {
                        const uint32_t v = ld_sys((const uint32_t *)&gprog[y - 1]);
                        if ((v >> 16) == ep && (int)(v & 0xffffu) >= x + 1) break;
                        __builtin_amdgcn_s_sleep(1);
                        if (++n > spin_limit) { atomicOr(err, 1); break; }
                    }
                }
                __builtin_amdgcn_wave_barrier();
                asm volatile("" ::: "memory");
                if (lane < HAND) h = ld_sys(ghand + ((size_t)(y - 1) * mbw + x) * HAND + lane);
            }
            if (lane < HAND) S.top[lane] = h;
            wave_lds_sync();
        }
        const int sl = slice_of(m, total, num_slices);
        const bool te = y > 0, le = x > 0;
        const bool ta = te && slice_of(m - mbw, total, num_slices) == sl;
        const bool la = le && slice_of(m - 1, total, num_slices) == sl;
        const bool ca = te && le && slice_of(m - mbw - 1, total, num_slices) == sl;
        const bool plane_ok = ta && la && ca;

        // ---- stage A: luma mode decision + residual; chroma DC prediction + residual
        int top = 128, left = 128, ctop = 128, cleft = 128;
        const int corner = (te && le) ? corner_next : 128;
        if (lane < 16) {
            if (te) top = ((const uint8_t *)S.top)[lane];
            if (le) left = S.leftY[lane];
            const int comp = lane >> 3, i = lane & 7;
            if (ta) ctop = ((const uint8_t *)S.top)[16 + 2 * i + comp];
            if (la) cleft = S.leftC[comp][i];
        }
        if (te) corner_next = ((const uint8_t *)S.top)[15];
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
        // chroma DC prediction of this lane's quadrant (H.264 8.3.4.1-3)
        int t0 = 0, t1 = 0, l0 = 0, l1 = 0;
        for (int i = 0; i < 4; i++) {
            t0 += __shfl(ctop, ccomp * 8 + i, 64);
            t1 += __shfl(ctop, ccomp * 8 + 4 + i, 64);
            l0 += __shfl(cleft, ccomp * 8 + i, 64);
            l1 += __shfl(cleft, ccomp * 8 + 4 + i, 64);
        }
        int cp;
        if (csq == 0) cp = (ta && la) ? (t0 + l0 + 4) >> 3 : ta ? (t0 + 2) >> 2 : la ? (l0 + 2) >> 2 : 128;
        else if (csq == 1) cp = ta ? (t1 + 2) >> 2 : la ? (l0 + 2) >> 2 : 128;
        else if (csq == 2) cp = la ? (l1 + 2) >> 2 : ta ? (t0 + 2) >> 2 : 128;
        else cp = (ta && la) ? (t1 + l1 + 4) >> 3 : ta ? (t1 + 2) >> 2 : la ? (l1 + 2) >> 2 : 128;
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
            S.pred[lb][lk + i] = p;
            S.res[lb][lk + i] = s[i] - p;
        }
        if (lane < 32) {
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const uint32_t word = c < 2 ? cw2.x : cw2.y;
                const int sv2 = (word >> (16 * (c & 1) + 8 * ccomp)) & 255;
                S.cres[ccomp * 4 + csq][crow * 4 + c] = sv2 - cp;
            }
            if (crow == 0) S.cpred[ccomp][csq] = cp;
        }
        if (lane == 0) modes[m] = mode;
        wave_lds_sync();

        // ---- stage B: forward 4x4 + intra AC quantisation (luma lanes 0..15, chroma 16..23)
        if (lane < 24) {
            const bool lu = lane < 16;
            const int cb = lane - 16;
            const int b = lu ? lane : lane;  // block index in the 24-block layout
            int r[16], co[16];
#pragma unroll
            for (int k = 0; k < 16; k++) r[k] = lu ? S.res[lane][k] : S.cres[cb][k];
            fwd4x4(r, co);
            if (lu) S.dcin[lane] = co[0];
            else S.cdin[cb] = co[0];
            dc[m * 24 + b] = co[0];
            uint32_t msk = 0, pk[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int k = 1; k < 16; k++) {
                const int l = lu ? quant_ac(co[k], k, per, rem, fi) : quant_ac(co[k], k, perc, remc, fic);
                if (lu) S.acl[lane][k] = l;
                else S.cacl[cb][k] = l;
                pk[k >> 1] |= (uint32_t)(uint16_t)(int16_t)l << (16 * (k & 1));
                msk |= (l != 0 ? 1u : 0u) << k;
            }
            uint4 *dst = (uint4 *)(levels + ((size_t)m * 24 + b) * 16);
            dst[0] = make_uint4(pk[0], pk[1], pk[2], pk[3]);
            dst[1] = make_uint4(pk[4], pk[5], pk[6], pk[7]);
            nz[m * 24 + b] = msk;
        }
        wave_lds_sync();
        // ---- stage C: DC transforms + quantisers, one coefficient per lane
        if (lane < 16) {
            const int raw = (hadamard_coef(S.dcin, lane) + 1) >> 1;
            S.dcl[lane] = quant_dc_k5(raw, per, rem);
        } else if (lane < 24) {
            const int k = lane - 16, comp = k >> 2, q = k & 3;
            S.cdl[k] = quant_dc_k5(hadamard2_coef(&S.cdin[comp * 4], q), perc, remc);
        }
        wave_lds_sync();
        if (lane < 16) {
            S.dcv[lane] = dequant_dc_luma(hadamard_coef(S.dcl, lane), per, rem);
        } else if (lane < 24) {
            const int k = lane - 16, comp = k >> 2, q = k & 3;
            S.cdcv[k] = dequant_dc_chroma(hadamard2_coef(&S.cdl[comp * 4], q), perc, remc);
        }
        wave_lds_sync();
        // ---- stage D: dequantise, inverse transform, reconstruct
        if (lane < 16) {
            int d[16], r[16];
            d[0] = S.dcv[lane];
#pragma unroll
            for (int k = 1; k < 16; k++) d[k] = dequant_ac(S.acl[lane][k], k, per, rem);
            inv4x4(d, r);
            const int bx = lane & 3, by = lane >> 2;
#pragma unroll
            for (int row = 0; row < 4; row++) {
                uint32_t word = 0;
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    int v = clip255(S.pred[lane][row * 4 + c] + r[row * 4 + c]);
                    word |= (uint32_t)v << (8 * c);
                    if (bx == 3 && c == 3) S.leftY[by * 4 + row] = (uint8_t)v;
                }
                *(uint32_t *)(oy + (size_t)(Y0 + by * 4 + row) * opy + X0 + bx * 4) = word;
                if (by == 3 && row == 3) S.out[bx] = word;  // bottom luma row -> hand-off
            }
        } else if (lane < 20) {
            const int sq = lane - 16;
            int rr[2][16];
#pragma unroll
            for (int comp = 0; comp < 2; comp++) {
                int d[16];
                d[0] = S.cdcv[comp * 4 + sq];
#pragma unroll
                for (int k = 1; k < 16; k++) d[k] = dequant_ac(S.cacl[comp * 4 + sq][k], k, perc, remc);
                inv4x4(d, rr[comp]);
            }
            const int ux = U0 + (sq & 1) * 4, vy = V0 + (sq >> 1) * 4;
#pragma unroll
            for (int row = 0; row < 4; row++) {
                uint32_t lo = 0, hi = 0;
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    const int k = row * 4 + c;
                    const uint32_t cbv = (uint32_t)clip255(S.cpred[0][sq] + rr[0][k]);
                    const uint32_t crv = (uint32_t)clip255(S.cpred[1][sq] + rr[1][k]);
                    if (c < 2) lo |= (cbv | (crv << 8)) << (16 * c);
                    else hi |= (cbv | (crv << 8)) << (16 * (c - 2));
                    if ((sq & 1) && c == 3) {
                        S.leftC[0][(sq >> 1) * 4 + row] = (uint8_t)cbv;
                        S.leftC[1][(sq >> 1) * 4 + row] = (uint8_t)crv;
                    }
                }
                *(uint2 *)(ouv + (size_t)(vy + row) * opuv + 2 * ux) = make_uint2(lo, hi);
                if ((sq >> 1) && row == 3) {  // bottom chroma row -> hand-off
                    S.out[4 + (sq & 1) * 2] = lo;
                    S.out[5 + (sq & 1) * 2] = hi;
                }
            }
        }
        wave_lds_sync();
        // ---- hand MB x's bottom rows to the row below and publish
        if (!last_row) {
            const uint32_t h = lane < HAND ? S.out[lane] : 0u;
            if (lds_below) {
                unsigned n = 0;
                while (vlcons[w + 1] < x + 1 - RING) {
                    __builtin_amdgcn_s_sleep(1);
                    if (++n > spin_limit) { atomicOr(err, 2); break; }
                }
                if (lane < HAND) ring[w][x % RING][lane] = h;
                wave_lds_sync();
                if (lane == 0) vlprog[w] = x + 1;
            } else {
                if (lane < HAND) st_sys(ghand + ((size_t)y * mbw + x) * HAND + lane, h);
                wait_stores();
                __builtin_amdgcn_wave_barrier();
                if (lane == 0) st_sys((uint32_t *)&gprog[y], wf_tag(ep, x + 1));
            }
        }
    }
}

// Workspace: a 256-byte-aligned header of mbh + 3 ints (ticket, error flag,
// one progress word per MB row, launch epoch), then the hand-off records.
static size_t intra_header_bytes(int mbh) { return ((size_t)(mbh + 3) * sizeof(int) + 255) & ~(size_t)255; }

extern "C" size_t bc250_intra_work_bytes(int wc, int hc)
{
    const int mbw = wc / 16, mbh = hc / 16;
    return intra_header_bytes(mbh) + (size_t)mbh * mbw * HAND * 4;
}

extern "C" int bc250_launch_i_encode(const uint8_t *cy, const uint8_t *cuv, int cpy, int cpuv, uint8_t *oy,
                                     uint8_t *ouv, int opy, int opuv, int wc, int hc, int qp, int num_slices,
                                     int16_t *levels, int32_t *dc, uint32_t *nz, uint32_t *modes, int *sync,
                                     unsigned spin_limit, void *stream)
{
    if (wc <= 0 || hc <= 0 || (wc & 15) || (hc & 15) || qp < 0 || qp > 51 || (opy & 3) || (opuv & 7) ||
        (cpy & 3) || (cpuv & 7) || num_slices < 1 || wc / 16 >= 65535)
        return -1;
    const int mbw = wc / 16, mbh = hc / 16;
    uint32_t *ghand = (uint32_t *)((uint8_t *)sync + intra_header_bytes(mbh));
    const int groups = (mbh + WAVES - 1) / WAVES;
    wavefront_begin_kernel<<<1, 256, 0, (hipStream_t)stream>>>(sync, mbh + 2, (uint32_t *)sync + 2 + mbh);
    h264_i_encode_kernel<<<groups, 64 * WAVES, 0, (hipStream_t)stream>>>(cy, cuv, cpy, cpuv, oy, ouv, opy, opuv,
                                                                          mbw, mbh, qp, num_slices, levels, dc, nz,
                                                                          modes, sync, ghand, spin_limit);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
