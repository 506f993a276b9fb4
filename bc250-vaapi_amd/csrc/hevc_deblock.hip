// hevc_deblock.hip - H.265 8.7.2 deblocking for gfx950, for the pipeline's
// coding shape (8x8 CUs, one PU each, one reference picture), 8-bit NV12 or
// 10-bit P010 (the sample in the high 10 bits of each 16-bit word).
//
// HEVC filters edges on the 8x8 grid and changes at most three samples on
// either side, so within one direction every edge segment is independent:
// two flat launches, all vertical edges, then all horizontal edges on that
// result.  One thread per 4-line segment: a luma segment of one 8x8-grid
// edge, or a chroma segment (Cb and Cr together, NV12 interleaved) of one
// edge of the 8x8 chroma grid (chroma is filtered at bS 2 only).  bS: 2
// intra; 1 if either CU has coded luma or the vectors differ by >= 4
// quarter samples; else 0 (the reference decoder's strength(),
// enc/src/hevc_filter_template.c:221-290).  Picture borders and slice
// borders (CTU-row slices, no filtering across slices) are left alone.
#include "bc250_hevc.h"
#include "hevc_dev.h"

using namespace hevc;

namespace {
__constant__ uint8_t kBetaT[52] = {0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  6,  7,
                                   8,  9,  10, 11, 12, 13, 14, 15, 16, 17, 18, 20, 22, 24, 26, 28, 30, 32,
                                   34, 36, 38, 40, 42, 44, 46, 48, 50, 52, 54, 56, 58, 60, 62, 64};
__constant__ uint8_t kTcT[54] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,  0,  1,  1,  1,  1,  1,  1,  1,  1, 1,
                                 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 5, 5, 6, 6, 7, 8, 9, 10, 11, 13, 14, 16, 18, 20, 22, 24};

BC_DEV int cu_at(int x, int y, int wctu)
{
    return ((y >> 4) * wctu + (x >> 4)) * 4 + ((y >> 3) & 1) * 2 + ((x >> 3) & 1);
}

// Whether the 8x8-grid edge at coordinate e (x of a vertical edge, y of a
// horizontal one) between CU8 p (before it) and q lies inside one CU16 /
// CU32, i.e. is neither a TU nor a PU edge: both sides carry the same CU size
// 2^L >= 16 and e is not a multiple of 2^L.
BC_DEV bool inside_cu(const uint8_t *culog2, int cp, int cq, int e)
{
    if (!culog2) return false;
    const int L = culog2[cq];
    return L >= 4 && culog2[cp] == L && (e & ((1 << L) - 1)) != 0;
}

BC_DEV int strength(int intra, const uint8_t *cbf, const uint4 *mv, int cp, int cq)
{
    if (intra) return 2;
    if (cbf[cp] | cbf[cq]) return 1;
    const uint4 a = mv[cp], b = mv[cq];
    return (iabs((int)a.x - (int)b.x) >= 4 || iabs((int)a.y - (int)b.y) >= 4) ? 1 : 0;
}

// One luma segment: v[k][0..7] = p3 p2 p1 p0 q0 q1 q2 q3 of line k (8.7.2.5.3,
// 8.7.2.5.6, 8.7.2.5.7); beta and tc already scaled to the bit depth.
BC_DEV void luma_segment(int v[4][8], int beta, int tc, int maxv)
{
    const int dp0 = iabs(v[0][1] - 2 * v[0][2] + v[0][3]), dp3 = iabs(v[3][1] - 2 * v[3][2] + v[3][3]);
    const int dq0 = iabs(v[0][6] - 2 * v[0][5] + v[0][4]), dq3 = iabs(v[3][6] - 2 * v[3][5] + v[3][4]);
    if (dp0 + dq0 + dp3 + dq3 >= beta) return;
    bool strong = true;
#pragma unroll
    for (int k = 0; k < 4; k += 3) {
        const int dd = k ? dp3 + dq3 : dp0 + dq0;
        strong = strong && 2 * dd < (beta >> 2) && iabs(v[k][0] - v[k][3]) + iabs(v[k][4] - v[k][7]) < (beta >> 3) &&
                 iabs(v[k][3] - v[k][4]) < ((5 * tc + 1) >> 1);
    }
    const bool dEp = (dp0 + dp3) < ((beta + (beta >> 1)) >> 3), dEq = (dq0 + dq3) < ((beta + (beta >> 1)) >> 3);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int p3 = v[k][0], p2 = v[k][1], p1 = v[k][2], p0 = v[k][3];
        const int q0 = v[k][4], q1 = v[k][5], q2 = v[k][6], q3 = v[k][7];
        if (strong) {
            const int t2 = 2 * tc;
            v[k][3] = clampi((p2 + 2 * p1 + 2 * p0 + 2 * q0 + q1 + 4) >> 3, p0 - t2, p0 + t2);
            v[k][2] = clampi((p2 + p1 + p0 + q0 + 2) >> 2, p1 - t2, p1 + t2);
            v[k][1] = clampi((2 * p3 + 3 * p2 + p1 + p0 + q0 + 4) >> 3, p2 - t2, p2 + t2);
            v[k][4] = clampi((p1 + 2 * p0 + 2 * q0 + 2 * q1 + q2 + 4) >> 3, q0 - t2, q0 + t2);
            v[k][5] = clampi((p0 + q0 + q1 + q2 + 2) >> 2, q1 - t2, q1 + t2);
            v[k][6] = clampi((p0 + q0 + q1 + 3 * q2 + 2 * q3 + 4) >> 3, q2 - t2, q2 + t2);
        } else {
            int delta = (9 * (q0 - p0) - 3 * (q1 - p1) + 8) >> 4;
            if (iabs(delta) >= tc * 10) continue;
            delta = clampi(delta, -tc, tc);
            v[k][3] = clampi(p0 + delta, 0, maxv);
            v[k][4] = clampi(q0 - delta, 0, maxv);
            const int th = tc >> 1;
            if (dEp) v[k][2] = clampi(p1 + clampi((((p2 + p0 + 1) >> 1) - p1 + delta) >> 1, -th, th), 0, maxv);
            if (dEq) v[k][5] = clampi(q1 + clampi((((q2 + q0 + 1) >> 1) - q1 - delta) >> 1, -th, th), 0, maxv);
        }
    }
}

BC_DEV void chroma_line(int p1, int &p0, int &q0, int q1, int tc, int maxv)
{
    const int delta = clampi((((q0 - p0) * 4) + p1 - q1 + 4) >> 3, -tc, tc);
    p0 = clampi(p0 + delta, 0, maxv);
    q0 = clampi(q0 - delta, 0, maxv);
}

}  // namespace

// VERT: vertical edges (filtered along x); otherwise horizontal edges.
// Pitches in bytes; Pix uint8_t (8-bit) or uint16_t (P010, BD 10).
template <typename Pix, int BD, bool VERT>
__global__ __launch_bounds__(256) void hevc_deblock_kernel(uint8_t *__restrict__ y, uint8_t *__restrict__ uv, int py,
                                                           int puv, int wc, int hc, int qp, int intra,
                                                           const uint8_t *__restrict__ cbf,
                                                           const uint4 *__restrict__ mv,
                                                           const uint8_t *__restrict__ culog2, int num_slices,
                                                           int ctu_log2, int nl, int nc)
{
    constexpr int MAXV = (1 << BD) - 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int wctu = wc >> 4, hctu = hc >> 4;
    if (t < nl) {
        // luma: edges at 8, 16, .. (VERT: columns), segments of 4 lines along the edge
        const int nseg = (VERT ? hc : wc) >> 2;
        const int e = 8 * (t / nseg + 1), s = 4 * (t % nseg);
        if (!VERT && (e & 15) == 0 && slice_first_row16(e >> 4, hctu, num_slices, ctu_log2) == (e >> 4)) return;
        const int x = VERT ? e : s, yy = VERT ? s : e;
        const int cq = cu_at(x, yy, wctu), cp = cu_at(VERT ? x - 1 : x, VERT ? yy : yy - 1, wctu);
        const int bs = inside_cu(culog2, cp, cq, e) ? 0 : strength(intra, cbf, mv, cp, cq);
        if (!bs) return;
        const int beta = kBetaT[clampi(qp, 0, 51)] << (BD - 8), tc = kTcT[clampi(qp + 2 * (bs - 1), 0, 53)] << (BD - 8);
        int v[4][8];
        if (VERT) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint8_t *row = y + (size_t)(yy + k) * py;
                ld4<Pix>(row, x - 4, &v[k][0]);
                ld4<Pix>(row, x, &v[k][4]);
            }
        } else {
#pragma unroll
            for (int i = 0; i < 8; i++) {
                int r[4];
                ld4<Pix>(y + (size_t)(yy - 4 + i) * py, x, r);
#pragma unroll
                for (int k = 0; k < 4; k++) v[k][i] = r[k];
            }
        }
        luma_segment(v, beta, tc, MAXV);
        if (VERT) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                uint8_t *row = y + (size_t)(yy + k) * py;
                st4<Pix>(row, x - 4, &v[k][0]);
                st4<Pix>(row, x, &v[k][4]);
            }
        } else {
#pragma unroll
            for (int i = 1; i < 7; i++) {
                const int r[4] = {v[0][i], v[1][i], v[2][i], v[3][i]};
                st4<Pix>(y + (size_t)(yy - 4 + i) * py, x, r);
            }
        }
        return;
    }
    const int c = t - nl;
    if (c >= nc) return;
    // chroma: edges at chroma 8, 16, .. (luma multiples of 16), 4-line segments, bS 2 only
    const int nseg = (VERT ? hc : wc) >> 3;
    const int e = 8 * (c / nseg + 1), s = 4 * (c % nseg);
    if (!VERT && (e & 7) == 0 && slice_first_row16(e >> 3, hctu, num_slices, ctu_log2) == (e >> 3)) return;
    const int cx = VERT ? e : s, cy = VERT ? s : e;
    const int lx = 2 * cx, ly = 2 * cy;
    const int cq = cu_at(lx, ly, wctu), cp = cu_at(VERT ? lx - 1 : lx, VERT ? ly : ly - 1, wctu);
    if (inside_cu(culog2, cp, cq, 2 * e) || strength(intra, cbf, mv, cp, cq) != 2) return;
    const int tc = kTcT[clampi(hevc::chroma_qp(qp) + 2, 0, 53)] << (BD - 8);
    if (VERT) {
        // per line: interleaved Cb/Cr of chroma columns cx-2 .. cx+1 (8 samples)
#pragma unroll
        for (int k = 0; k < 4; k++) {
            uint8_t *row = uv + (size_t)(cy + k) * puv;
            int a[4], b[4];
            ld4<Pix>(row, 2 * cx - 4, a);
            ld4<Pix>(row, 2 * cx, b);
#pragma unroll
            for (int comp = 0; comp < 2; comp++) chroma_line(a[comp], a[2 + comp], b[comp], b[2 + comp], tc, MAXV);
            st4<Pix>(row, 2 * cx - 4, a);
            st4<Pix>(row, 2 * cx, b);
        }
    } else {
        // 4 chroma columns (8 interleaved samples) across rows cy-2 .. cy+1
        int r[4][8];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const uint8_t *row = uv + (size_t)(cy - 2 + i) * puv;
            ld4<Pix>(row, 2 * cx, &r[i][0]);
            ld4<Pix>(row, 2 * cx + 4, &r[i][4]);
        }
#pragma unroll
        for (int k = 0; k < 8; k++) chroma_line(r[0][k], r[1][k], r[2][k], r[3][k], tc, MAXV);
#pragma unroll
        for (int i = 1; i < 3; i++) {
            uint8_t *row = uv + (size_t)(cy - 2 + i) * puv;
            st4<Pix>(row, 2 * cx, &r[i][0]);
            st4<Pix>(row, 2 * cx + 4, &r[i][4]);
        }
    }
}

// cbf[cu] = any of the CU's 64 luma levels (P layout: CTU t at coef[384 t],
// CU z at +64 z) is non-zero; one 64-lane wave per CU, one level per lane.
__global__ __launch_bounds__(256) void hevc_cu_cbf_kernel(const int16_t *__restrict__ coef, int ncu,
                                                          uint8_t *__restrict__ cbf)
{
    const int cu = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (cu >= ncu) return;
    const int v = coef[(size_t)(cu >> 2) * 384 + (cu & 3) * 64 + lane];
    const bool any = __any(v != 0);
    if (lane == 0) cbf[cu] = any ? 1 : 0;
}

// The same with CU16 / CU32 (culog2 per CU8, the P layout of
// bc250_hevc_launch_p_encode_ctu): the flag of a CU8 is that of the
// transform block holding it, 64, 256 or 4 x 256 luma levels.
__global__ __launch_bounds__(256) void hevc_cu_cbf_ctu_kernel(const int16_t *__restrict__ coef, int ncu, int wtu,
                                                              const uint8_t *__restrict__ culog2,
                                                              uint8_t *__restrict__ cbf)
{
    const int cu = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (cu >= ncu) return;
    const int t = cu >> 2, L = culog2[cu];
    bool nz = false;
    if (L == 3) {
        nz = coef[(size_t)t * 384 + (cu & 3) * 64 + lane] != 0;
    } else {
        const int tx = (t % wtu) & ~1, ty = (t / wtu) & ~1;
        for (int a = 0; a < (L == 5 ? 4 : 1); a++) {
            const int ta = L == 5 ? (ty + (a >> 1)) * wtu + tx + (a & 1) : t;
#pragma unroll
            for (int i = 0; i < 4; i++) nz = nz || coef[(size_t)ta * 384 + lane + 64 * i] != 0;
        }
    }
    const bool any = __any(nz);
    if (lane == 0) cbf[cu] = any ? 1 : 0;
}

extern "C" int bc250_hevc_cu_cbf_ctu(const int16_t *coef, int wc, int hc, const uint8_t *culog2, uint8_t *cbf,
                                     void *stream)
{
    if (wc <= 0 || hc <= 0 || (wc & 15) || (hc & 15)) return -1;
    const int wtu = wc / 16, nctu = wtu * (hc / 16);
    if (!culog2) return bc250_hevc_cu_cbf(coef, nctu, cbf, stream);
    hevc_cu_cbf_ctu_kernel<<<nctu, 256, 0, (hipStream_t)stream>>>(coef, nctu * 4, wtu, culog2, cbf);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int bc250_hevc_cu_cbf(const int16_t *coef, int nctu, uint8_t *cbf, void *stream)
{
    if (nctu < 0) return -1;
    if (nctu == 0) return 0;
    hevc_cu_cbf_kernel<<<nctu, 256, 0, (hipStream_t)stream>>>(coef, nctu * 4, cbf);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename Pix, int BD>
static void launch_deblock(uint8_t *y, uint8_t *uv, int py, int puv, int wc, int hc, int qp, int is_intra,
                           const uint8_t *cbf, const uint4 *mvs, const uint8_t *culog2, int num_slices, int ctu_log2,
                           hipStream_t st)
{
    // vertical pass
    int nl = (wc / 8 - 1) * (hc / 4), nc = (wc / 16 - 1) * (hc / 8);
    int n = nl + nc;
    if (n > 0)
        hevc_deblock_kernel<Pix, BD, true>
            <<<(n + 255) / 256, 256, 0, st>>>(y, uv, py, puv, wc, hc, qp, is_intra, cbf, mvs, culog2, num_slices,
                                              ctu_log2, nl, nc);
    // horizontal pass
    nl = (hc / 8 - 1) * (wc / 4);
    nc = (hc / 16 - 1) * (wc / 8);
    n = nl + nc;
    if (n > 0)
        hevc_deblock_kernel<Pix, BD, false>
            <<<(n + 255) / 256, 256, 0, st>>>(y, uv, py, puv, wc, hc, qp, is_intra, cbf, mvs, culog2, num_slices,
                                              ctu_log2, nl, nc);
}

extern "C" int bc250_hevc_launch_deblock_ctu(void *y, void *uv, int py, int puv, int wc, int hc, int qp,
                                             int is_intra, const uint8_t *cbf, const uint32_t *mvs,
                                             const uint8_t *culog2, int num_slices, int ctu_log2, int bd,
                                             void *stream)
{
    const int bps = bd == 8 ? 1 : 2;
    if ((bd != 8 && bd != 10) || wc <= 0 || hc <= 0 || (wc & 15) || (hc & 15) || qp < 0 || qp > 51 || (py & 7) ||
        (puv & 7) || py < wc * bps || puv < wc * bps || num_slices < 1 || (ctu_log2 != 4 && ctu_log2 != 5))
        return -1;
    hipStream_t st = (hipStream_t)stream;
    if (bd == 8)
        launch_deblock<uint8_t, 8>((uint8_t *)y, (uint8_t *)uv, py, puv, wc, hc, qp, is_intra, cbf,
                                   (const uint4 *)mvs, culog2, num_slices, ctu_log2, st);
    else
        launch_deblock<uint16_t, 10>((uint8_t *)y, (uint8_t *)uv, py, puv, wc, hc, qp, is_intra, cbf,
                                     (const uint4 *)mvs, culog2, num_slices, ctu_log2, st);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int bc250_hevc_launch_deblock(void *y, void *uv, int py, int puv, int wc, int hc, int qp, int is_intra,
                                         const uint8_t *cbf, const uint32_t *mvs, int num_slices, int bd,
                                         void *stream)
{
    return bc250_hevc_launch_deblock_ctu(y, uv, py, puv, wc, hc, qp, is_intra, cbf, mvs, nullptr, num_slices, 4, bd,
                                         stream);
}
