// hevc_inter.hip - HEVC transform / quant and motion-compensation stages for
// gfx950: batched forms of the reference's per-block functions, the
// quarter-sample refinement of the full-search 8x8 vectors, and the fused
// P-picture CU coding (MC, 8x8 luma / 4x4 chroma DCT, quantisation,
// dequantisation, inverse, reconstruction), one wave per 8x8 CU.
//
// Transforms run as one lane per coefficient: each pass is a dot product of a
// matrix row with a column held in LDS, then the rounding shift and int16
// saturation the reference applies after every pass.
#include "bc250_hevc.h"
#include "hevc_dev.h"
#include "learned_dev.h"

using namespace hevc;

namespace {

// Bit-depth dependent shifts of the reference's transforms.
struct Depth {
    int s1_4, qsh4, s1_8, qsh8, s2inv, maxv, qmax4;
};
BC_DEV Depth depth_of(int bd)
{
    Depth d;
    d.s1_4 = bd == 8 ? 1 : 3;    // 4x4 forward first pass
    d.qsh4 = bd + 2 - 5;         // 4x4 bdShift = BitDepth + log2(4) - 5
    d.s1_8 = bd - 6;             // 8x8 forward first pass
    d.qsh8 = bd + 3 - 5;         // 8x8 bdShift
    d.s2inv = 20 - bd;           // inverse second pass
    d.maxv = (1 << bd) - 1;
    d.qmax4 = bd == 8 ? 51 : 63;  // forward 4x4 QP clamp
    return d;
}

// 24-bit multiply(-add)s, always the full-rate v_mul_i32_i24 / v_mad_i32_i24
// (a plain or __mul24 product may become the quarter-rate v_mul_lo_u32).
// Every operand here is a sample, a tap or an int16-range intermediate, so
// the 24-bit products are exact.
template <int C>
BC_DEV int mad24c(int v, int acc)
{
    int d;
    asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(d) : "v"(v), "i"(C), "v"(acc));
    return d;
}
BC_DEV int mad24(int a, int b, int acc)
{
    int d;
    asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(acc));
    return d;
}

// One n x n forward or inverse pass for the coefficient this lane owns:
// lane (i, j) -> sat16((sum_k M[i][k] * X[k][j] + rnd) >> s)  (forward rows of
// M; `transposed` uses M[k][i] instead).  X row-major in LDS.
// Every product in this file has an int8 matrix entry or filter tap and an
// int16-range operand (samples, residuals, saturated intermediates), so the
// full-rate 24-bit multiply (__mul24) is exact.
template <int N>
BC_DEV int pass(const int8_t *M, bool transposed, const int *X, int i, int j, int s)
{
    int acc = 0;
#pragma unroll
    for (int k = 0; k < N; k++) acc += __mul24((int)(transposed ? M[k * N + i] : M[i * N + k]), X[k * N + j]);
    return sat16((acc + (1 << (s - 1))) >> s);
}

// Second pass of a separable transform applied along rows: lane (i, j) ->
// sat16((sum_k M[j][k] * T[i][k] + rnd) >> s), transposed as above.
template <int N>
BC_DEV int pass_rows(const int8_t *M, bool transposed, const int *T, int i, int j, int s)
{
    int acc = 0;
#pragma unroll
    for (int k = 0; k < N; k++) acc += __mul24((int)(transposed ? M[k * N + j] : M[j * N + k]), T[i * N + k]);
    return sat16((acc + (1 << (s - 1))) >> s);
}

// HEVC 8-tap luma / 4-tap chroma uni-prediction of one sample from an int
// window W (pitch P) whose (wx, wy) is the sample's integer position, with
// the reference's intermediate int16 stores; returns the clipped sample.
template <bool CHROMA>
BC_DEV int mc_sample(const int *W, int P, int wx, int wy, int fx, int fy, int bd)
{
    constexpr int NT = CHROMA ? 4 : 8, BACK = CHROMA ? 1 : 3;
    const int8_t *fh = CHROMA ? kEpel[fx] : kQpel[fx];
    const int8_t *fv = CHROMA ? kEpel[fy] : kQpel[fy];
    const int down = bd - 8;
    if (!fx && !fy) return W[wy * P + wx];
    int p14;
    if (!fy) {
        int s = 0;
#pragma unroll
        for (int k = 0; k < NT; k++) s = mad24((int)fh[k], W[wy * P + wx + k - BACK], s);
        p14 = (int16_t)(s >> down);
    } else if (!fx) {
        int s = 0;
#pragma unroll
        for (int k = 0; k < NT; k++) s = mad24((int)fv[k], W[(wy + k - BACK) * P + wx], s);
        p14 = (int16_t)(s >> down);
    } else {
        int s = 0;
#pragma unroll
        for (int k = 0; k < NT; k++) {
            int t = 0;
#pragma unroll
            for (int j = 0; j < NT; j++) t = mad24((int)fh[j], W[(wy + k - BACK) * P + wx + j - BACK], t);
            s = mad24((int)fv[k], (int)(int16_t)(t >> down), s);
        }
        p14 = (int16_t)(s >> 6);
    }
    return clampi((p14 + (1 << (13 - bd))) >> (14 - bd), 0, (1 << bd) - 1);
}

}  // namespace

// ---------------------------------------------------------------------------
// Batched 4x4 / 8x8 transform + quant and dequant + inverse: one wave per
// four 4x4 blocks (lane = block*16 + coefficient) or per 8x8 block.
__global__ __launch_bounds__(256) void hevc_tq4_kernel(const int16_t *__restrict__ res, int n, int q, int use_dst,
                                                       int round_q12, int bd, int16_t *__restrict__ out)
{
    __shared__ int X[4][64], T[4][64];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int b = (blockIdx.x * 4 + wv) * 4 + (lane >> 4), k = lane & 15;
    const int i = k >> 2, j = k & 3, base = (lane >> 4) * 16;
    const Depth D = depth_of(bd);
    const int8_t *M = use_dst ? kDst4 : kDct4;
    X[wv][lane] = b < n ? res[(size_t)b * 16 + k] : 0;
    __syncthreads();
    T[wv][lane] = pass<4>(M, false, &X[wv][base], i, j, D.s1_4);
    __syncthreads();
    const int v = pass_rows<4>(M, false, &T[wv][base], i, j, 8);
    const Quant Q = make_quant(clampi(q, 0, D.qmax4), D.qsh4, round_q12);
    if (b < n) out[(size_t)b * 16 + k] = (int16_t)quant(v, Q);
}

__global__ __launch_bounds__(256) void hevc_dqit4_kernel(const int16_t *__restrict__ coef, int n, int q, int use_dst,
                                                         int bd, int16_t *__restrict__ out)
{
    __shared__ int X[4][64], T[4][64];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int b = (blockIdx.x * 4 + wv) * 4 + (lane >> 4), k = lane & 15;
    const int i = k >> 2, j = k & 3, base = (lane >> 4) * 16;
    const Depth D = depth_of(bd);
    const int8_t *M = use_dst ? kDst4 : kDct4;
    const int qq = bd == 8 ? q : clampi(q, 0, 63);
    X[wv][lane] = b < n ? dequant(coef[(size_t)b * 16 + k], qq, D.qsh4) : 0;
    __syncthreads();
    T[wv][lane] = pass<4>(M, true, &X[wv][base], i, j, 7);
    __syncthreads();
    const int v = pass_rows<4>(M, true, &T[wv][base], i, j, D.s2inv);
    if (b < n) out[(size_t)b * 16 + k] = (int16_t)v;
}

__global__ __launch_bounds__(256) void hevc_tq8_kernel(const int16_t *__restrict__ res, int n, int q, int bd,
                                                       int round_q12, int16_t *__restrict__ out)
{
    __shared__ int X[4][64], T[4][64];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int b = blockIdx.x * 4 + wv, i = lane >> 3, j = lane & 7;
    const Depth D = depth_of(bd);
    X[wv][lane] = b < n ? res[(size_t)b * 64 + lane] : 0;
    __syncthreads();
    T[wv][lane] = pass<8>(kDct8, false, X[wv], i, j, D.s1_8);
    __syncthreads();
    const int v = pass_rows<8>(kDct8, false, T[wv], i, j, 9);
    const Quant Q = make_quant(clampi(q, 0, 63), D.qsh8, round_q12);
    if (b < n) out[(size_t)b * 64 + lane] = (int16_t)quant(v, Q);
}

__global__ __launch_bounds__(256) void hevc_dqit8_kernel(const int16_t *__restrict__ coef, int n, int q, int bd,
                                                         int16_t *__restrict__ out)
{
    __shared__ int X[4][64], T[4][64];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int b = blockIdx.x * 4 + wv, i = lane >> 3, j = lane & 7;
    const Depth D = depth_of(bd);
    X[wv][lane] = b < n ? dequant(coef[(size_t)b * 64 + lane], clampi(q, 0, 63), D.qsh8) : 0;
    __syncthreads();
    T[wv][lane] = pass<8>(kDct8, true, X[wv], i, j, 7);
    __syncthreads();
    const int v = pass_rows<8>(kDct8, true, T[wv], i, j, D.s2inv);
    if (b < n) out[(size_t)b * 64 + lane] = (int16_t)v;
}

// N x N DCT-II (N = 16 or 32) transform + quant and dequant + inverse, one
// 256-thread workgroup per block.  The N-point matrix is built in LDS: row k
// is row k*32/N of the 32-point one, entries are the standard's magnitudes of
// 64*sqrt(2)*cos(pi*b/64) (b = (2n+1)k mod 128 folded into 0..32) with the
// cosine's sign.  Shifts generalise the 4x4 / 8x8 ones: forward (log2n + bd -
// 9, log2n + 6), quant / dequant bdShift bd + log2n - 5, inverse (7, 20 - bd).
BC_DEV void build_dct_matrix(int *M, int N)
{
    for (int e = threadIdx.x; e < N * N; e += blockDim.x) {
        const int k = e / N;
        M[e] = dct_coef(k, e - k * N, N);
    }
}

template <bool INVERSE>
__global__ __launch_bounds__(256) void hevc_nxn_kernel(const int16_t *__restrict__ in, int n, int log2n, int q, int bd,
                                                       int round_q12, int16_t *__restrict__ out)
{
    __shared__ int M[1024], X[1024], T[1024];
    const int N = 1 << log2n, N2 = N * N, b = blockIdx.x;
    build_dct_matrix(M, N);
    const int bds = bd + log2n - 5;
    for (int e = threadIdx.x; e < N2; e += 256) {
        const int v = in[(size_t)b * N2 + e];
        X[e] = INVERSE ? dequant(v, clampi(q, 0, 63), bds) : v;
    }
    __syncthreads();
    const int s1 = INVERSE ? 7 : log2n + bd - 9, s2 = INVERSE ? 20 - bd : log2n + 6;
    for (int e = threadIdx.x; e < N2; e += 256) {  // columns: T[i][c] = sum_r M'[i][r] X[r][c]
        const int i = e >> log2n, c = e & (N - 1);
        int acc = 0;
        for (int r = 0; r < N; r++) acc += __mul24(INVERSE ? M[r * N + i] : M[i * N + r], X[r * N + c]);
        T[e] = sat16((acc + (1 << (s1 - 1))) >> s1);
    }
    __syncthreads();
    const Quant Q = make_quant(clampi(q, 0, bd == 8 ? 51 : 63), bds, round_q12);
    for (int e = threadIdx.x; e < N2; e += 256) {  // rows: Y[i][j] = sum_c M'[j][c] T[i][c]
        const int i = e >> log2n, j = e & (N - 1);
        int acc = 0;
        for (int c = 0; c < N; c++) acc += __mul24(INVERSE ? M[c * N + j] : M[j * N + c], T[i * N + c]);
        const int v = sat16((acc + (1 << (s2 - 1))) >> s2);
        out[(size_t)b * N2 + e] = (int16_t)(INVERSE ? v : quant(v, Q));
    }
}

// Batched uni-prediction: one wave per block (w*h <= 256 samples, each lane
// up to 4 of them); the block's reference window (w+7) x (h+7) in LDS.
template <typename Pix>
__global__ __launch_bounds__(64) void hevc_mc_batch_kernel(const Pix *__restrict__ ref, int stride, int w_pic,
                                                          int h_pic, const int4 *__restrict__ blk, int n, int w,
                                                          int h, int chroma, int bd, Pix *__restrict__ dst)
{
    __shared__ int W[23 * 23];
    const int b = blockIdx.x, lane = threadIdx.x;
    if (b >= n) return;
    const int4 d = blk[b];
    const int steps = chroma ? 3 : 2, back = chroma ? 1 : 3;
    const int ix = d.x + (d.z >> steps), iy = d.y + (d.w >> steps);
    const int fx = d.z & ((1 << steps) - 1), fy = d.w & ((1 << steps) - 1);
    const int P = w + 7;
    for (int t = lane; t < P * (h + 7); t += 64) {
        const int r = t / P, c = t - r * P;
        W[t] = ref[(size_t)clampi(iy - back + r, 0, h_pic - 1) * stride + clampi(ix - back + c, 0, w_pic - 1)];
    }
    __syncthreads();
    for (int t = lane; t < w * h; t += 64) {
        const int r = t / w, c = t - r * w;
        const int v = chroma ? mc_sample<true>(W, P, c + back, r + back, fx, fy, bd)
                             : mc_sample<false>(W, P, c + back, r + back, fx, fy, bd);
        dst[(size_t)b * w * h + t] = (Pix)v;
    }
}

// ---------------------------------------------------------------------------
// Quarter-sample refinement, one wave per 8x8 CU, 4 CUs (one CTU) per
// workgroup.  Every candidate lies within 3 quarter samples of the integer
// vector (dx, dy), so all taps fall in a 16x16 window W at (x0+dx-4, y0+dy-4)
// and a sample's integer position is window column pc + 3 or pc + 4 (row
// likewise).  The horizontal 8-tap results for fx = 1..3 at window columns
// 3..11 of every row are computed once (H, the int16 intermediates of
// 8.5.3.3.3.1); a candidate then costs 8 vertical taps per sample.  Pitches
// in bytes; Pix uint8_t (8-bit) or uint16_t (P010, BD 10).
struct RefineLds8 {
    int w[16 * 16];      // window samples
    int h[3][16 * 16];   // fx = 1..3: horizontal sums >> (BD - 8) at columns 3..11 (pitch 16)
};

// The luma interpolation sum of fraction F (1..3) over the eight samples
// v(0..7), taps as immediates (8.5.3.3.3.1; the same sums as kQpel[F]).
template <int F, typename G>
BC_DEV int qpel_sum(G v)
{
    if constexpr (F == 1)
        return mad24c<-5>(v(5), mad24c<17>(v(4), mad24c<58>(v(3), mad24c<-10>(v(2), 4 * v(1) - v(0) + v(6)))));
    else if constexpr (F == 2)
        return mad24c<40>(v(3) + v(4), mad24c<-11>(v(2) + v(5), 4 * (v(1) + v(6)) - v(0) - v(7)));
    else
        return mad24c<-10>(v(5), mad24c<58>(v(4), mad24c<17>(v(3), mad24c<-5>(v(2), v(1) + 4 * v(6) - v(7)))));
}

// qpel_sum of a fraction f known to be the same in every lane of the wave.
template <typename G>
BC_DEV int qpel_sum_u(int f, G v)
{
    if (f == 1) return qpel_sum<1>(v);
    if (f == 2) return qpel_sum<2>(v);
    return qpel_sum<3>(v);
}

// Refinement of the CU8 at (x0, y0) from its integer vector iv (the same in
// every lane) by one wave: returns {mvx, mvy (quarter samples), sad, 0}, the
// same in every lane; with pred non-null, also this lane's luma prediction
// sample at the returned vector (the value the CU coding's motion
// compensation gives, from the same window).
template <typename Pix, int BD>
BC_DEV uint4 refine_cu(RefineLds8 &L, const uint8_t *__restrict__ cur, int cpitch, const uint8_t *__restrict__ ref,
                       int rpitch, int wc, int hc, int x0, int y0, uint4 iv, int lane, int *pred = nullptr)
{
    constexpr int DOWN = BD - 8, MAXV = (1 << BD) - 1, SH = 14 - BD, OFF = 1 << (13 - BD);
    const int dx = __builtin_amdgcn_readfirstlane((int)iv.x), dy = __builtin_amdgcn_readfirstlane((int)iv.y);
    const int ox = x0 + dx - 4, oy = y0 + dy - 4;
    for (int t = lane; t < 256; t += 64)
        L.w[t] = ldpix<Pix>(ref + (size_t)clampi(oy + (t >> 4), 0, hc - 1) * rpitch, clampi(ox + (t & 15), 0, wc - 1));
    const int pr = lane >> 3, pc = lane & 7;
    const int cv = ldpix<Pix>(cur + (size_t)(y0 + pr) * cpitch, x0 + pc);
    wave_lds_sync();
    // horizontal intermediates of the three fractions at 16 rows x columns
    // 3..11: a lane per position, its eight samples read once
    for (int t = lane; t < 144; t += 64) {
        const int row = t / 9, col = t - row * 9 + 3;
        const int *w = &L.w[row * 16 + col - 3];
        int sm[8];
#pragma unroll
        for (int k = 0; k < 8; k++) sm[k] = w[k];
        auto v = [&](int k) { return sm[k]; };
        const int e = row * 16 + col - 3;
        L.h[0][e] = (int16_t)(qpel_sum<1>(v) >> DOWN);
        L.h[1][e] = (int16_t)(qpel_sum<2>(v) >> DOWN);
        L.h[2][e] = (int16_t)(qpel_sum<3>(v) >> DOWN);
    }
    wave_lds_sync();
    // (mx, my) is the same in every lane, so each branch below is uniform
    auto pred_of = [&](int mx, int my) {
        const int fx = mx & 3, fy = my & 3;
        const int wx = pc + (mx >> 2) - dx + 4, wy = pr + (my >> 2) - dy + 4;  // integer position in W
        if (!fx && !fy) return L.w[wy * 16 + wx];
        int p14;
        if (!fy) {
            p14 = L.h[fx - 1][wy * 16 + wx - 3];
        } else if (!fx) {
            const int *c = &L.w[(wy - 3) * 16 + wx];
            p14 = (int16_t)(qpel_sum_u(fy, [&](int k) { return c[k * 16]; }) >> DOWN);
        } else {
            const int *c = &L.h[fx - 1][(wy - 3) * 16 + wx - 3];
            p14 = (int16_t)(qpel_sum_u(fy, [&](int k) { return c[k * 16]; }) >> 6);
        }
        return clampi((p14 + OFF) >> SH, 0, MAXV);
    };
    auto diff_of = [&](int mx, int my) { return (uint32_t)iabs(cv - pred_of(mx, my)); };
    const int hx[8] = {2, -2, 0, 0, 2, 2, -2, -2}, hy[8] = {0, 0, 2, -2, 2, -2, 2, -2};
    int bx = dx * 4, by = dy * 4;
    uint32_t best = wave_sum_u32(diff_of(bx, by));
#pragma unroll 1
    for (int stage = 1; stage <= 2; stage++) {
        const int cx = bx, cy = by;
        // the stage's eight differences first, two to a word (64 lanes of at
        // most 1023 sum below 2^16), then their sums, then the choice in
        // candidate order (strictly lower wins)
        uint32_t d[8];
#pragma unroll
        for (int k = 0; k < 8; k += 2) {
            const int e = stage == 1 ? 1 : 2;
            const uint32_t lo = diff_of(cx + hx[k] / e, cy + hy[k] / e);
            const uint32_t hi = diff_of(cx + hx[k + 1] / e, cy + hy[k + 1] / e);
            const uint32_t sum = wave_sum_u32(lo | hi << 16);
            d[k] = sum & 0xffffu;
            d[k + 1] = sum >> 16;
        }
#pragma unroll
        for (int k = 0; k < 8; k++) {
            if (d[k] < best) {
                best = d[k];
                bx = cx + (stage == 1 ? hx[k] : hx[k] / 2);
                by = cy + (stage == 1 ? hy[k] : hy[k] / 2);
            }
        }
    }
    if (pred) *pred = pred_of(bx, by);
    return make_uint4((uint32_t)bx, (uint32_t)by, best, 0u);
}

template <typename Pix, int BD>
__global__ __launch_bounds__(256) void hevc_refine8_kernel(const uint8_t *__restrict__ cur, int cpitch,
                                                           const uint8_t *__restrict__ ref, int rpitch, int wc, int hc,
                                                           int nctu, const uint4 *__restrict__ imv8,
                                                           uint4 *__restrict__ out)
{
    __shared__ RefineLds8 Ls[4];
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int ctu = xcd_item(blockIdx.x, nctu);
    if (ctu >= nctu) return;
    const int cu = ctu * 4 + wv, mbw = wc / 16;
    const int x0 = (ctu % mbw) * 16 + (wv & 1) * 8, y0 = (ctu / mbw) * 16 + (wv >> 1) * 8;
    const uint4 v = refine_cu<Pix, BD>(Ls[wv], cur, cpitch, ref, rpitch, wc, hc, x0, y0, imv8[cu], lane);
    if (lane == 0) out[cu] = v;
}

// ---------------------------------------------------------------------------
// P-picture CU coding, one wave per 8x8 CU (lane = luma sample / coefficient;
// lanes 0..31 also carry the Cb and Cr 4x4 blocks), 4 CUs per workgroup.
struct PcuLds {
    int wl[15 * 15];    // luma reference window at (x0 + mvx>>2 - 3, y0 + mvy>>2 - 3)
    int wcr[2][7 * 7];  // Cb / Cr windows at (x0/2 + mvx>>3 - 1, ...)
    int a[64], b[64];   // luma residual -> pass 1 -> dequantised -> pass 1 (inverse)
    int ca[32], cb[32];
};

// Packed levels for the readback (CTU16 P pictures): every 16-level group of
// the levels array that holds a nonzero level, with its group index (the
// 24 groups of a CTU16: 16 luma, 4 Cb, 4 Cr), at slots taken with one atomic
// add per CU; *count starts at zero.  idx null: no packing.
struct PackOut {
    uint32_t *idx;
    int16_t *data;
    int *count;
};

// The learned search's full-search keys, 5 per MB (bc250_learned_search_
// picture_keys): an MB whose 16x16 key is set was searched in full, and its
// records are made here from the keys (its 8x8 one for the CU8, the 16x16
// one to out16 for the next picture's proposer).  keys null: none.
struct FsKeys {
    const uint32_t *keys;
    uint4 *out16;
    int range, lambda8;
};

// Coding of CU8 wv of CTU ctu (origin (x0, y0)) with vector (mvx, mvy) by one
// wave.  HAVE_PRED: the lane's luma prediction sample is `pred_in` (the
// refinement's), so no luma window is staged.
template <typename Pix, int BD, bool HAVE_PRED = false>
BC_DEV void encode_cu(PcuLds &L, const uint8_t *__restrict__ cy, const uint8_t *__restrict__ cuv, int cpy, int cpuv,
                      const uint8_t *__restrict__ ry, const uint8_t *__restrict__ ruv, int rpy, int rpuv,
                      uint8_t *__restrict__ oy, uint8_t *__restrict__ ouv, int opy, int opuv, int wc, int hc, int qp,
                      int round_q12, int ctu, int wv, int lane, int x0, int y0, int mvx, int mvy,
                      int16_t *__restrict__ coef, uint8_t *__restrict__ cbf, uint8_t *__restrict__ modes,
                      int pred_in = 0, const PackOut pk = PackOut{nullptr, nullptr, nullptr})
{
    constexpr int MAXV = (1 << BD) - 1, QOFF = 6 * (BD - 8);
    constexpr int S1_8 = BD - 6, QSH8 = BD - 2, S1_4 = BD == 8 ? 1 : 3, QSH4 = BD - 3, S2INV = 20 - BD;
    const int cu = ctu * 4 + wv;
    const int ux0 = x0 >> 1, uy0 = y0 >> 1, wcc = wc >> 1, hcc = hc >> 1;
    const int lx = x0 + (mvx >> 2) - 3, ly = y0 + (mvy >> 2) - 3;
    const int cx = ux0 + (mvx >> 3) - 1, cyy = uy0 + (mvy >> 3) - 1;
    if (!HAVE_PRED)
        for (int t = lane; t < 225; t += 64)
            L.wl[t] = ldpix<Pix>(ry + (size_t)clampi(ly + t / 15, 0, hc - 1) * rpy, clampi(lx + t % 15, 0, wc - 1));
    for (int t = lane; t < 98; t += 64) {
        const int comp = t >= 49, k = t - comp * 49;
        L.wcr[comp][k] =
            ldpix<Pix>(ruv + (size_t)clampi(cyy + k / 7, 0, hcc - 1) * rpuv, 2 * clampi(cx + k % 7, 0, wcc - 1) + comp);
    }
    const int pr = lane >> 3, pc = lane & 7;
    const int cur = ldpix<Pix>(cy + (size_t)(y0 + pr) * cpy, x0 + pc);
    const bool cl = lane < 32;
    const int comp = (lane >> 4) & 1, ck = lane & 15, cr_ = ck >> 2, cc_ = ck & 3;
    const int ccur = cl ? ldpix<Pix>(cuv + (size_t)(uy0 + cr_) * cpuv, 2 * (ux0 + cc_) + comp) : 0;
    wave_lds_sync();

    // prediction and residual
    const int pred = HAVE_PRED ? pred_in : mc_sample<false>(L.wl, 15, pc + 3, pr + 3, mvx & 3, mvy & 3, BD);
    const int cpred = cl ? mc_sample<true>(L.wcr[comp], 7, cc_ + 1, cr_ + 1, mvx & 7, mvy & 7, BD) : 0;
    L.a[lane] = cur - pred;
    if (cl) L.ca[lane] = ccur - cpred;
    wave_lds_sync();

    // forward: columns, then rows
    L.b[lane] = pass<8>(kDct8, false, L.a, pr, pc, S1_8);
    if (cl) L.cb[lane] = pass<4>(kDct4, false, &L.ca[comp * 16], cr_, cc_, S1_4);
    wave_lds_sync();
    const int q = clampi(qp + QOFF, 0, 63), cq = hevc::chroma_qp(qp) + QOFF;
    const int lv = quant(pass_rows<8>(kDct8, false, L.b, pr, pc, 9), make_quant(q, QSH8, round_q12));
    const int clv = cl ? quant(pass_rows<4>(kDct4, false, &L.cb[comp * 16], cr_, cc_, 8),
                               make_quant(cq, QSH4, round_q12))
                       : 0;
    int16_t *co = coef + (size_t)ctu * 384;
    co[wv * 64 + lane] = (int16_t)lv;
    // optional side outputs: the CU's coded-luma flag, and zero P modes
    if (cbf) {
        const bool any = __any(lv != 0);
        if (lane == 0) cbf[cu] = any ? 1 : 0;
    }
    if (modes && lane < 4) modes[(size_t)ctu * 16 + wv * 4 + lane] = 0;
    if (cl) co[256 + comp * 64 + wv * 16 + ck] = (int16_t)clv;
    if (pk.idx) {
        // the CU's six groups: luma rows 2i, 2i+1 (lanes 16i..), Cb, Cr
        const uint64_t ml = __ballot(lv != 0), mc = __ballot(cl && clv != 0);
        uint32_t gm = 0;
#pragma unroll
        for (int i = 0; i < 4; i++) gm |= ((ml >> (16 * i)) & 0xffffull) ? 1u << i : 0u;
        gm |= (mc & 0xffffull) ? 16u : 0u;
        gm |= ((mc >> 16) & 0xffffull) ? 32u : 0u;
        if (gm) {
            int base = 0;
            if (lane == 0) base = atomicAdd(pk.count, __popc(gm));
            base = __shfl(base, 0);
            const int gi = lane >> 4, j = lane & 15;
            if ((gm >> gi) & 1u) {
                const int k = base + __popc(gm & ((1u << gi) - 1u));
                pk.data[(size_t)k * 16 + j] = (int16_t)lv;
                if (j == 0) pk.idx[k] = (uint32_t)(ctu * 24 + wv * 4 + gi);
            }
            if (cl && ((gm >> (4 + comp)) & 1u)) {
                const int k = base + __popc(gm & ((1u << (4 + comp)) - 1u));
                pk.data[(size_t)k * 16 + j] = (int16_t)clv;
                if (j == 0) pk.idx[k] = (uint32_t)(ctu * 24 + 16 + comp * 4 + wv);
            }
        }
    }

    // dequantise, inverse: columns, then rows
    L.a[lane] = dequant(lv, q, QSH8);
    if (cl) L.ca[lane] = dequant(clv, cq, QSH4);
    wave_lds_sync();
    L.b[lane] = pass<8>(kDct8, true, L.a, pr, pc, 7);
    if (cl) L.cb[lane] = pass<4>(kDct4, true, &L.ca[comp * 16], cr_, cc_, 7);
    wave_lds_sync();
    const int r = pass_rows<8>(kDct8, true, L.b, pr, pc, S2INV);
    stpix<Pix>(oy + (size_t)(y0 + pr) * opy, x0 + pc, clampi(pred + r, 0, MAXV));
    if (cl) {
        const int crr = pass_rows<4>(kDct4, true, &L.cb[comp * 16], cr_, cc_, S2INV);
        stpix<Pix>(ouv + (size_t)(uy0 + cr_) * opuv, 2 * (ux0 + cc_) + comp, clampi(cpred + crr, 0, MAXV));
    }
}

template <typename Pix, int BD>
__global__ __launch_bounds__(256) void hevc_p_encode_kernel(
    const uint8_t *__restrict__ cy, const uint8_t *__restrict__ cuv, int cpy, int cpuv, const uint8_t *__restrict__ ry,
    const uint8_t *__restrict__ ruv, int rpy, int rpuv, uint8_t *__restrict__ oy, uint8_t *__restrict__ ouv, int opy,
    int opuv, int wc, int hc, int nctu, int qp, int round_q12, const uint4 *__restrict__ mvs,
    int16_t *__restrict__ coef, uint8_t *__restrict__ cbf, uint8_t *__restrict__ modes)
{
    __shared__ PcuLds Ls[4];
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int ctu = xcd_item(blockIdx.x, nctu);
    if (ctu >= nctu) return;
    const int mbw = wc / 16;
    const int x0 = (ctu % mbw) * 16 + (wv & 1) * 8, y0 = (ctu / mbw) * 16 + (wv >> 1) * 8;
    const uint4 mv = mvs[ctu * 4 + wv];
    encode_cu<Pix, BD>(Ls[wv], cy, cuv, cpy, cpuv, ry, ruv, rpy, rpuv, oy, ouv, opy, opuv, wc, hc, qp, round_q12, ctu,
                       wv, lane, x0, y0, (int)mv.x, (int)mv.y, coef, cbf, modes);
}

// Refinement and coding of each CU8 in one launch: the wave that refines a
// CU's vector codes the CU with it (the same work and results as
// hevc_refine8_kernel followed by hevc_p_encode_kernel).
// At least 6 waves per SIMD: the refinement half needs the occupancy that
// the coding half's registers would otherwise take (measured: 1080p P coding
// 0.137 -> 0.128 ms per frame; 8 waves spill more and lose it again).
#ifndef BC250_RPE_WAVES_PER_EU
#define BC250_RPE_WAVES_PER_EU 6
#endif
// One wave per workgroup (a CU8 each, the CTU's four on one XCD): nothing is
// shared between a CTU's CU8s, and single-wave workgroups measured faster
// than four-wave ones.
template <typename Pix, int BD>
__global__ __launch_bounds__(64, BC250_RPE_WAVES_PER_EU) void hevc_refine_p_encode_kernel(
    const uint8_t *__restrict__ cy, const uint8_t *__restrict__ cuv, int cpy, int cpuv, const uint8_t *__restrict__ ry,
    const uint8_t *__restrict__ ruv, int rpy, int rpuv, uint8_t *__restrict__ oy, uint8_t *__restrict__ ouv, int opy,
    int opuv, int wc, int hc, int nctu, int qp, int round_q12, const uint4 *__restrict__ imv8,
    uint4 *__restrict__ mv8, int16_t *__restrict__ coef, uint8_t *__restrict__ cbf, uint8_t *__restrict__ modes,
    const PackOut pk, const FsKeys fk)
{
    union CuLds {
        RefineLds8 r;
        PcuLds p;
    };
    __shared__ CuLds Ls;
    const int lane = threadIdx.x;
    const int cu = xcd_item(blockIdx.x, 4 * nctu);
    if (cu >= 4 * nctu) return;
    const int ctu = cu >> 2, wv = cu & 3, mbw = wc / 16;
    const int x0 = (ctu % mbw) * 16 + (wv & 1) * 8, y0 = (ctu / mbw) * 16 + (wv >> 1) * 8;
    int pred;
    uint4 iv = imv8[cu];
    if (fk.keys) {
        const uint32_t k16 = fk.keys[5 * ctu], k8 = fk.keys[5 * ctu + 1 + wv];
        if (k16 != 0xffffffffu) {
            iv = fs_record(k8, 1 + wv, fk.range, fk.lambda8);
            if (wv == 0 && lane == 0) fk.out16[ctu] = fs_record(k16, 0, fk.range, fk.lambda8);
        }
    }
    const uint4 v = refine_cu<Pix, BD>(Ls.r, cy, cpy, ry, rpy, wc, hc, x0, y0, iv, lane, &pred);
    if (lane == 0) mv8[cu] = v;
    wave_lds_sync();  // the refinement's LDS is the coding's from here on
    encode_cu<Pix, BD, true>(Ls.p, cy, cuv, cpy, cpuv, ry, ruv, rpy, rpuv, oy, ouv, opy, opuv, wc, hc, qp,
                             round_q12, ctu, wv, lane, x0, y0, (int)v.x, (int)v.y, coef, cbf, modes, pred, pk);
}

// ---------------------------------------------------------------------------
// P picture at CTU32: CU16 / CU32 choice and coding, one 256-thread workgroup
// per CTU32, after hevc_p_encode_kernel (whose CU8 levels and reconstruction
// it replaces where it chooses a larger CU).  A 16x16 area whose four CU8
// vectors are equal is one CU16; a CTU32 inside the picture whose four areas
// are CU16s with one vector is one CU32.  A CU16 / CU32 codes one luma TU of
// its size and one chroma TU of half its size per component (N-point DCT-II,
// shifts (log2n + bd - 9, log2n + 6), quantiser shift bd + log2n - 5,
// inverse (7, 20 - bd)).  Levels: area q of a CU32 (z-order) holds luma
// levels [256 q, 256 q + 256) of the TU (row-major) and chroma levels
// [64 q, 64 q + 64) per component, in the area's luma / Cb / Cr slots.
// culog2[4 t + z] = log2 of the CU holding CU8 z of area t; modes (may be
// null) gets the same at [16 t + z].
struct LargeLds {
    int w[39 * 39];     // reference window: luma (n + 7)^2, chroma (n/2 + 3)^2
    int x[1024];        // residual, then dequantised levels
    int t[1024];        // first-pass results
    int m[1024];        // N-point matrix
    int mv[16][2];
    int cu16[4], all;
};

template <typename Pix, int BD>
__global__ __launch_bounds__(256) void hevc_p_large_kernel(
    const uint8_t *__restrict__ cy, const uint8_t *__restrict__ cuv, int cpy, int cpuv, const uint8_t *__restrict__ ry,
    const uint8_t *__restrict__ ruv, int rpy, int rpuv, uint8_t *__restrict__ oy, uint8_t *__restrict__ ouv, int opy,
    int opuv, int wc, int hc, int qp, int round_q12, const uint4 *__restrict__ mvs, int16_t *__restrict__ coef,
    uint8_t *__restrict__ culog2, uint8_t *__restrict__ modes)
{
    constexpr int MAXV = (1 << BD) - 1, QOFF = 6 * (BD - 8);
    __shared__ LargeLds L;
    const int tid = threadIdx.x, wtu = wc >> 4, htu = hc >> 4, w32 = (wtu + 1) >> 1;
    const int X = blockIdx.x % w32, Y = blockIdx.x / w32;
    if (tid < 16) {
        const int a = tid >> 2, z = tid & 3, tx = 2 * X + (a & 1), ty = 2 * Y + (a >> 1);
        uint4 v = make_uint4(0, 0, 0, 0);
        if (tx < wtu && ty < htu) v = mvs[(ty * wtu + tx) * 4 + z];
        L.mv[tid][0] = (int)v.x;
        L.mv[tid][1] = (int)v.y;
    }
    __syncthreads();
    if (tid < 4) {
        const int tx = 2 * X + (tid & 1), ty = 2 * Y + (tid >> 1);
        bool eq = tx < wtu && ty < htu;
        for (int z = 1; z < 4; z++)
            eq = eq && L.mv[tid * 4 + z][0] == L.mv[tid * 4][0] && L.mv[tid * 4 + z][1] == L.mv[tid * 4][1];
        L.cu16[tid] = eq;
    }
    __syncthreads();
    if (tid == 0) {
        bool all = L.cu16[0] && L.cu16[1] && L.cu16[2] && L.cu16[3];
        for (int a = 1; a < 4; a++) all = all && L.mv[4 * a][0] == L.mv[0][0] && L.mv[4 * a][1] == L.mv[0][1];
        L.all = all;
    }
    __syncthreads();
    const bool all = L.all;
    if (tid < 16) {
        const int a = tid >> 2, z = tid & 3, tx = 2 * X + (a & 1), ty = 2 * Y + (a >> 1);
        if (tx < wtu && ty < htu) {
            const int t = ty * wtu + tx, lg = all ? 5 : (L.cu16[a] ? 4 : 3);
            culog2[t * 4 + z] = (uint8_t)lg;
            if (modes) modes[t * 16 + z] = (uint8_t)lg;
        }
    }
    const int q = clampi(qp + QOFF, 0, 63), cq = hevc::chroma_qp(qp) + QOFF;
    for (int a = 0; a < (all ? 1 : 4); a++) {
        if (!all && !L.cu16[a]) continue;
        const int log2n = all ? 5 : 4, n = 1 << log2n;
        const int x0 = 32 * X + (all ? 0 : (a & 1) * 16), y0 = 32 * Y + (all ? 0 : (a >> 1) * 16);
        const int mvx = L.mv[4 * a][0], mvy = L.mv[4 * a][1];
        for (int comp = -1; comp < 2; comp++) {
            const bool chroma = comp >= 0;
            const int lg = chroma ? log2n - 1 : log2n, m = 1 << lg, mm = m * m;
            const int P = chroma ? m + 3 : m + 7, back = chroma ? 1 : 3;
            const int bx = chroma ? x0 >> 1 : x0, by = chroma ? y0 >> 1 : y0;
            const int pw = chroma ? wc >> 1 : wc, ph = chroma ? hc >> 1 : hc;
            const int wx0 = bx + (chroma ? mvx >> 3 : mvx >> 2) - back, wy0 = by + (chroma ? mvy >> 3 : mvy >> 2) - back;
            const int fx = chroma ? mvx & 7 : mvx & 3, fy = chroma ? mvy & 7 : mvy & 3;
            const uint8_t *ref = chroma ? ruv : ry;
            const int rp = chroma ? rpuv : rpy;
            for (int e = tid; e < P * P; e += 256) {
                const int r = e / P, c = e - r * P;
                const int sx = clampi(wx0 + c, 0, pw - 1);
                L.w[e] = ldpix<Pix>(ref + (size_t)clampi(wy0 + r, 0, ph - 1) * rp, chroma ? 2 * sx + comp : sx);
            }
            for (int e = tid; e < mm; e += 256) {
                const int k = e >> lg;
                L.m[e] = dct_coef(k, e - (k << lg), m);
            }
            __syncthreads();
            int pred[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int e = tid + 256 * i;
                if (e < mm) {
                    const int r = e >> lg, c = e & (m - 1);
                    pred[i] = chroma ? mc_sample<true>(L.w, P, c + back, r + back, fx, fy, BD)
                                     : mc_sample<false>(L.w, P, c + back, r + back, fx, fy, BD);
                    const int cur = chroma ? ldpix<Pix>(cuv + (size_t)(by + r) * cpuv, 2 * (bx + c) + comp)
                                           : ldpix<Pix>(cy + (size_t)(by + r) * cpy, bx + c);
                    L.x[e] = cur - pred[i];
                }
            }
            __syncthreads();
            const int qq = chroma ? cq : q, bds = BD + lg - 5;
            const int s1 = lg + BD - 9, s2 = lg + 6, s2i = 20 - BD;
            for (int e = tid; e < mm; e += 256) {  // columns: T[i][c] = sum_r M[i][r] X[r][c]
                const int i = e >> lg, c = e & (m - 1);
                int acc = 0;
                for (int r = 0; r < m; r++) acc += __mul24(L.m[(i << lg) + r], L.x[(r << lg) + c]);
                L.t[e] = sat16((acc + (1 << (s1 - 1))) >> s1);
            }
            __syncthreads();
            const Quant Q = make_quant(qq, bds, round_q12);
            const int per = mm >> (log2n == 5 ? 2 : 0), base = chroma ? 256 + comp * 64 : 0;
            int lv[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {  // rows: Y[i][j] = sum_c M[j][c] T[i][c]
                const int e = tid + 256 * i;
                if (e < mm) {
                    const int ii = e >> lg, j = e & (m - 1);
                    int acc = 0;
                    for (int c = 0; c < m; c++) acc += __mul24(L.m[(j << lg) + c], L.t[(ii << lg) + c]);
                    lv[i] = quant(sat16((acc + (1 << (s2 - 1))) >> s2), Q);
                    const int ar = e / per, tx = (x0 >> 4) + (ar & 1), ty = (y0 >> 4) + (ar >> 1);
                    coef[(size_t)(ty * wtu + tx) * 384 + base + (e - ar * per)] = (int16_t)lv[i];
                }
            }
            __syncthreads();  // T and X free
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int e = tid + 256 * i;
                if (e < mm) L.x[e] = dequant(lv[i], qq, bds);
            }
            __syncthreads();
            for (int e = tid; e < mm; e += 256) {  // inverse columns: T[i][c] = sum_r M[r][i] X[r][c]
                const int i = e >> lg, c = e & (m - 1);
                int acc = 0;
                for (int r = 0; r < m; r++) acc += __mul24(L.m[(r << lg) + i], L.x[(r << lg) + c]);
                L.t[e] = sat16((acc + 64) >> 7);
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < 4; i++) {  // inverse rows, reconstruction
                const int e = tid + 256 * i;
                if (e < mm) {
                    const int ii = e >> lg, j = e & (m - 1);
                    int acc = 0;
                    for (int c = 0; c < m; c++) acc += __mul24(L.m[(c << lg) + j], L.t[(ii << lg) + c]);
                    const int rv = clampi(pred[i] + sat16((acc + (1 << (s2i - 1))) >> s2i), 0, MAXV);
                    if (chroma) stpix<Pix>(ouv + (size_t)(by + ii) * opuv, 2 * (bx + j) + comp, rv);
                    else stpix<Pix>(oy + (size_t)(by + ii) * opy, bx + j, rv);
                }
            }
            __syncthreads();  // window, matrix and buffers reused by the next block
        }
    }
}

// Fills n bytes with v.
__global__ void fill_u8_kernel(uint8_t *p, int n, uint8_t v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
static int launch_ok() { return hipGetLastError() == hipSuccess ? 0 : -1; }
static bool bd_ok(int bd) { return bd == 8 || bd == 10; }

extern "C" int bc250_hevc_tq4_batch(const int16_t *res, int n, int qp, int use_dst, int round_q12, int bd,
                                    int16_t *out, void *stream)
{
    if (n < 0 || !bd_ok(bd) || round_q12 < 0 || round_q12 > 12) return -1;
    if (n == 0) return 0;
    hevc_tq4_kernel<<<(n + 15) / 16, 256, 0, (hipStream_t)stream>>>(res, n, qp, use_dst, round_q12, bd, out);
    return launch_ok();
}

extern "C" int bc250_hevc_dqit4_batch(const int16_t *coef, int n, int qp, int use_dst, int bd, int16_t *out,
                                      void *stream)
{
    if (n < 0 || !bd_ok(bd) || qp < 0 || qp > 63) return -1;
    if (n == 0) return 0;
    hevc_dqit4_kernel<<<(n + 15) / 16, 256, 0, (hipStream_t)stream>>>(coef, n, qp, use_dst, bd, out);
    return launch_ok();
}

extern "C" int bc250_hevc_tq8_batch(const int16_t *res, int n, int qp, int bd, int round_q12, int16_t *out,
                                    void *stream)
{
    if (n < 0 || !bd_ok(bd) || round_q12 < 0 || round_q12 > 12) return -1;
    if (n == 0) return 0;
    hevc_tq8_kernel<<<(n + 3) / 4, 256, 0, (hipStream_t)stream>>>(res, n, qp, bd, round_q12, out);
    return launch_ok();
}

extern "C" int bc250_hevc_dqit8_batch(const int16_t *coef, int n, int qp, int bd, int16_t *out, void *stream)
{
    if (n < 0 || !bd_ok(bd)) return -1;
    if (n == 0) return 0;
    hevc_dqit8_kernel<<<(n + 3) / 4, 256, 0, (hipStream_t)stream>>>(coef, n, qp, bd, out);
    return launch_ok();
}

extern "C" int bc250_hevc_mc_uni_batch(const void *ref, int stride, int w_pic, int h_pic, const int32_t *blk, int n,
                                       int w, int h, int chroma, int bd, void *dst, void *stream)
{
    if (n < 0 || w < 1 || h < 1 || w > 16 || h > 16 || w_pic < 1 || h_pic < 1 || stride < w_pic || !bd_ok(bd))
        return -1;
    if (n == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    if (bd == 8)
        hevc_mc_batch_kernel<uint8_t><<<n, 64, 0, st>>>((const uint8_t *)ref, stride, w_pic, h_pic,
                                                        (const int4 *)blk, n, w, h, chroma ? 1 : 0, 8,
                                                        (uint8_t *)dst);
    else
        hevc_mc_batch_kernel<uint16_t><<<n, 64, 0, st>>>((const uint16_t *)ref, stride, w_pic, h_pic,
                                                         (const int4 *)blk, n, w, h, chroma ? 1 : 0, 10,
                                                         (uint16_t *)dst);
    return launch_ok();
}

int hevc::launch_nxn_matrix(const int16_t *in, int n, int log2n, int q, int bd, int round_q12, bool inverse,
                            int16_t *out, hipStream_t stream)
{
    if (inverse) hevc_nxn_kernel<true><<<n, 256, 0, stream>>>(in, n, log2n, q, bd, 0, out);
    else hevc_nxn_kernel<false><<<n, 256, 0, stream>>>(in, n, log2n, q, bd, round_q12, out);
    return launch_ok();
}

extern "C" int bc250_hevc_tqn_batch(const int16_t *res, int n, int log2n, int qp, int bd, int round_q12,
                                    int16_t *out, void *stream)
{
    if (n < 0 || log2n < 2 || log2n > 5 || !bd_ok(bd) || round_q12 < 0 || round_q12 > 12) return -1;
    if (n == 0) return 0;
    if (log2n == 5) return bc250_hevc_tq32_batch(res, n, qp, bd, round_q12, BC250_T32_DEFAULT, out, stream);
    return hevc::launch_nxn_matrix(res, n, log2n, qp, bd, round_q12, false, out, (hipStream_t)stream);
}

extern "C" int bc250_hevc_dqitn_batch(const int16_t *coef, int n, int log2n, int qp, int bd, int16_t *out,
                                      void *stream)
{
    if (n < 0 || log2n < 2 || log2n > 5 || !bd_ok(bd)) return -1;
    if (n == 0) return 0;
    if (log2n == 5) return bc250_hevc_dqit32_batch(coef, n, qp, bd, BC250_T32_DEFAULT, out, stream);
    return hevc::launch_nxn_matrix(coef, n, log2n, qp, bd, 0, true, out, (hipStream_t)stream);
}

static int grid_xcd(int items) { return ((items + 7) / 8) * 8; }

// P010 luma -> 8-bit luma holding the top 8 bits of each 10-bit sample (the
// high byte of the P010 word), four samples per thread.
__global__ __launch_bounds__(256) void hevc_p010_hi8_kernel(const uint8_t *__restrict__ src, int sp,
                                                            uint8_t *__restrict__ dst, int dp, int w4, int h)
{
    const int x4 = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x4 >= w4 || y >= h) return;
    const uint2 v = *(const uint2 *)(src + (size_t)y * sp + (size_t)x4 * 8);
    *(uint32_t *)(dst + (size_t)y * dp + (size_t)x4 * 4) =
        ((v.x >> 8) & 0xffu) | ((v.x >> 24) << 8) | (((v.y >> 8) & 0xffu) << 16) | ((v.y >> 24) << 24);
}

extern "C" int bc250_hevc_p010_to_8bit(const void *src, int spitch, uint8_t *dst, int dpitch, int w, int h,
                                       void *stream)
{
    if (w <= 0 || h <= 0 || (w & 3) || (spitch & 7) || (dpitch & 3) || spitch < 2 * w || dpitch < w) return -1;
    const int w4 = w / 4;
    hevc_p010_hi8_kernel<<<dim3((w4 + 255) / 256, h), 256, 0, (hipStream_t)stream>>>((const uint8_t *)src, spitch, dst,
                                                                                       dpitch, w4, h);
    return launch_ok();
}

extern "C" int bc250_hevc_launch_refine8(const void *cur, int cpitch, const void *ref, int rpitch, int wc, int hc,
                                         const uint32_t *imv8, uint32_t *out, int bd, void *stream)
{
    const int bps = bd == 8 ? 1 : 2;
    if (!bd_ok(bd) || wc <= 0 || hc <= 0 || (wc & 15) || (hc & 15) || cpitch < wc * bps || rpitch < wc * bps ||
        (cpitch % bps) || (rpitch % bps))
        return -1;
    const int nctu = (wc / 16) * (hc / 16);
    hipStream_t st = (hipStream_t)stream;
    if (bd == 8)
        hevc_refine8_kernel<uint8_t, 8><<<grid_xcd(nctu), 256, 0, st>>>(
            (const uint8_t *)cur, cpitch, (const uint8_t *)ref, rpitch, wc, hc, nctu, (const uint4 *)imv8, (uint4 *)out);
    else
        hevc_refine8_kernel<uint16_t, 10><<<grid_xcd(nctu), 256, 0, st>>>(
            (const uint8_t *)cur, cpitch, (const uint8_t *)ref, rpitch, wc, hc, nctu, (const uint4 *)imv8, (uint4 *)out);
    return launch_ok();
}

// The CTU32 large-CU pass (hevc_p_large_kernel) over the refined vectors.
static int launch_large(const void *cy, const void *cuv, int cpy, int cpuv, const void *ry, const void *ruv, int rpy,
                        int rpuv, void *oy, void *ouv, int opy, int opuv, int wc, int hc, int qp, int round_q12,
                        const uint32_t *mvs, int16_t *coef, uint8_t *culog2, uint8_t *modes, int bd, hipStream_t st)
{
    const int wtu = wc / 16, htu = hc / 16, n32 = ((wtu + 1) / 2) * ((htu + 1) / 2);
    auto *C = (const uint8_t *)cy, *CU = (const uint8_t *)cuv, *R = (const uint8_t *)ry, *RU = (const uint8_t *)ruv;
    auto *O = (uint8_t *)oy, *OU = (uint8_t *)ouv;
    if (bd == 8)
        hevc_p_large_kernel<uint8_t, 8><<<n32, 256, 0, st>>>(C, CU, cpy, cpuv, R, RU, rpy, rpuv, O, OU, opy, opuv, wc,
                                                             hc, qp, round_q12, (const uint4 *)mvs, coef, culog2,
                                                             modes);
    else
        hevc_p_large_kernel<uint16_t, 10><<<n32, 256, 0, st>>>(C, CU, cpy, cpuv, R, RU, rpy, rpuv, O, OU, opy, opuv,
                                                               wc, hc, qp, round_q12, (const uint4 *)mvs, coef,
                                                               culog2, modes);
    return launch_ok();
}

static int launch_p_encode(const void *cy, const void *cuv, int cpy, int cpuv, const void *ry, const void *ruv,
                           int rpy, int rpuv, void *oy, void *ouv, int opy, int opuv, int wc, int hc, int qp,
                           int round_q12, const uint32_t *mvs, int16_t *coef, uint8_t *cbf, uint8_t *modes, int bd,
                           void *stream)
{
    const int b = bd == 8 ? 1 : 2, mw = wc * b;
    if (!bd_ok(bd) || wc <= 0 || hc <= 0 || (wc & 15) || (hc & 15) || qp < 0 || qp > 51 || round_q12 < 0 ||
        round_q12 > 12 || cpy < mw || cpuv < mw || rpy < mw || rpuv < mw || opy < mw || opuv < mw ||
        ((cpy | cpuv | rpy | rpuv | opy | opuv) % b))
        return -1;
    const int nctu = (wc / 16) * (hc / 16);
    hipStream_t st = (hipStream_t)stream;
    auto *C = (const uint8_t *)cy, *CU = (const uint8_t *)cuv, *R = (const uint8_t *)ry, *RU = (const uint8_t *)ruv;
    auto *O = (uint8_t *)oy, *OU = (uint8_t *)ouv;
    if (bd == 8)
        hevc_p_encode_kernel<uint8_t, 8><<<grid_xcd(nctu), 256, 0, st>>>(C, CU, cpy, cpuv, R, RU, rpy, rpuv, O, OU, opy,
                                                                         opuv, wc, hc, nctu, qp, round_q12,
                                                                         (const uint4 *)mvs, coef, cbf, modes);
    else
        hevc_p_encode_kernel<uint16_t, 10><<<grid_xcd(nctu), 256, 0, st>>>(C, CU, cpy, cpuv, R, RU, rpy, rpuv, O, OU,
                                                                           opy, opuv, wc, hc, nctu, qp, round_q12,
                                                                           (const uint4 *)mvs, coef, cbf, modes);
    return launch_ok();
}

extern "C" int bc250_hevc_launch_p_encode(const void *cy, const void *cuv, int cpy, int cpuv, const void *ry,
                                          const void *ruv, int rpy, int rpuv, void *oy, void *ouv, int opy, int opuv,
                                          int wc, int hc, int qp, int round_q12, const uint32_t *mvs, int16_t *coef,
                                          int bd, void *stream)
{
    return launch_p_encode(cy, cuv, cpy, cpuv, ry, ruv, rpy, rpuv, oy, ouv, opy, opuv, wc, hc, qp, round_q12, mvs,
                           coef, nullptr, nullptr, bd, stream);
}

extern "C" int bc250_hevc_launch_p_encode_ctu(const void *cy, const void *cuv, int cpy, int cpuv, const void *ry,
                                              const void *ruv, int rpy, int rpuv, void *oy, void *ouv, int opy,
                                              int opuv, int wc, int hc, int qp, int round_q12, const uint32_t *mvs,
                                              int16_t *coef, int ctu_log2, uint8_t *culog2, uint8_t *modes,
                                              uint8_t *cbf, int bd, void *stream)
{
    if ((ctu_log2 != 4 && ctu_log2 != 5) || (ctu_log2 == 5 && !culog2)) return -1;
    const int r = launch_p_encode(cy, cuv, cpy, cpuv, ry, ruv, rpy, rpuv, oy, ouv, opy, opuv, wc, hc, qp, round_q12,
                                  mvs, coef, cbf, modes, bd, stream);
    if (r != 0) return r;
    hipStream_t st = (hipStream_t)stream;
    const int wtu = wc / 16, htu = hc / 16, ncu = wtu * htu * 4;
    if (ctu_log2 == 4) {
        if (culog2) fill_u8_kernel<<<(ncu + 255) / 256, 256, 0, st>>>(culog2, ncu, 3);
        return launch_ok();
    }
    return launch_large(cy, cuv, cpy, cpuv, ry, ruv, rpy, rpuv, oy, ouv, opy, opuv, wc, hc, qp, round_q12, mvs, coef,
                        culog2, modes, bd, st);
}

static int refine_p_encode(const void *cy, const void *cuv, int cpy, int cpuv, const void *ry, const void *ruv, int rpy,
                           int rpuv, void *oy, void *ouv, int opy, int opuv, int wc, int hc, int qp, int round_q12,
                           const uint32_t *imv8, uint32_t *mv8, int16_t *coef, int ctu_log2, uint8_t *culog2,
                           uint8_t *modes, uint8_t *cbf, int bd, const PackOut &pk, void *stream,
                           const FsKeys &fk = FsKeys{nullptr, nullptr, 0, 0})
{
    const int b = bd == 8 ? 1 : 2, mw = wc * b;
    if (!bd_ok(bd) || wc <= 0 || hc <= 0 || (wc & 15) || (hc & 15) || qp < 0 || qp > 51 || round_q12 < 0 ||
        round_q12 > 12 || cpy < mw || cpuv < mw || rpy < mw || rpuv < mw || opy < mw || opuv < mw ||
        ((cpy | cpuv | rpy | rpuv | opy | opuv) % b) || (ctu_log2 != 4 && ctu_log2 != 5) ||
        (ctu_log2 == 5 && !culog2) || !imv8 || !mv8 || (pk.idx && (ctu_log2 != 4 || !pk.data || !pk.count)))
        return -1;
    const int nctu = (wc / 16) * (hc / 16);
    hipStream_t st = (hipStream_t)stream;
    auto *C = (const uint8_t *)cy, *CU = (const uint8_t *)cuv, *R = (const uint8_t *)ry, *RU = (const uint8_t *)ruv;
    auto *O = (uint8_t *)oy, *OU = (uint8_t *)ouv;
    if (bd == 8)
        hevc_refine_p_encode_kernel<uint8_t, 8><<<grid_xcd(4 * nctu), 64, 0, st>>>(
            C, CU, cpy, cpuv, R, RU, rpy, rpuv, O, OU, opy, opuv, wc, hc, nctu, qp, round_q12, (const uint4 *)imv8,
            (uint4 *)mv8, coef, cbf, modes, pk, fk);
    else
        hevc_refine_p_encode_kernel<uint16_t, 10><<<grid_xcd(4 * nctu), 64, 0, st>>>(
            C, CU, cpy, cpuv, R, RU, rpy, rpuv, O, OU, opy, opuv, wc, hc, nctu, qp, round_q12, (const uint4 *)imv8,
            (uint4 *)mv8, coef, cbf, modes, pk, fk);
    if (launch_ok() != 0) return -1;
    const int ncu = nctu * 4;
    if (ctu_log2 == 4) {
        if (culog2) fill_u8_kernel<<<(ncu + 255) / 256, 256, 0, st>>>(culog2, ncu, 3);
        return launch_ok();
    }
    return launch_large(cy, cuv, cpy, cpuv, ry, ruv, rpy, rpuv, oy, ouv, opy, opuv, wc, hc, qp, round_q12, mv8, coef,
                        culog2, modes, bd, st);
}

extern "C" int bc250_hevc_launch_refine_p_encode(const void *cy, const void *cuv, int cpy, int cpuv, const void *ry,
                                                 const void *ruv, int rpy, int rpuv, void *oy, void *ouv, int opy,
                                                 int opuv, int wc, int hc, int qp, int round_q12,
                                                 const uint32_t *imv8, uint32_t *mv8, int16_t *coef, int ctu_log2,
                                                 uint8_t *culog2, uint8_t *modes, uint8_t *cbf, int bd, void *stream)
{
    return refine_p_encode(cy, cuv, cpy, cpuv, ry, ruv, rpy, rpuv, oy, ouv, opy, opuv, wc, hc, qp, round_q12, imv8, mv8,
                           coef, ctu_log2, culog2, modes, cbf, bd, PackOut{nullptr, nullptr, nullptr}, stream);
}

extern "C" int bc250_hevc_launch_refine_p_encode_pack(const void *cy, const void *cuv, int cpy, int cpuv,
                                                      const void *ry, const void *ruv, int rpy, int rpuv, void *oy,
                                                      void *ouv, int opy, int opuv, int wc, int hc, int qp,
                                                      int round_q12, const uint32_t *imv8, uint32_t *mv8,
                                                      int16_t *coef, uint8_t *modes, uint8_t *cbf, int bd,
                                                      uint32_t *pack_idx, int16_t *pack_data, int *pack_count,
                                                      void *stream)
{
    if (!pack_idx || !pack_data || !pack_count) return -1;
    return refine_p_encode(cy, cuv, cpy, cpuv, ry, ruv, rpy, rpuv, oy, ouv, opy, opuv, wc, hc, qp, round_q12, imv8, mv8,
                           coef, 4, nullptr, modes, cbf, bd, PackOut{pack_idx, pack_data, pack_count}, stream);
}

extern "C" int bc250_hevc_launch_refine_p_encode_pack_keys(
    const void *cy, const void *cuv, int cpy, int cpuv, const void *ry, const void *ruv, int rpy, int rpuv, void *oy,
    void *ouv, int opy, int opuv, int wc, int hc, int qp, int round_q12, const uint32_t *imv8, uint32_t *mv8,
    int16_t *coef, uint8_t *modes, uint8_t *cbf, int bd, uint32_t *pack_idx, int16_t *pack_data, int *pack_count,
    const uint32_t *fs_keys, uint32_t *out16, int range, int lambda8, void *stream)
{
    if (!pack_idx || !pack_data || !pack_count || !fs_keys || !out16 || range < 1 || range > 64 || lambda8 < 0 ||
        lambda8 > 15)
        return -1;
    return refine_p_encode(cy, cuv, cpy, cpuv, ry, ruv, rpy, rpuv, oy, ouv, opy, opuv, wc, hc, qp, round_q12, imv8, mv8,
                           coef, 4, nullptr, modes, cbf, bd, PackOut{pack_idx, pack_data, pack_count}, stream,
                           FsKeys{fs_keys, (uint4 *)out16, range, lambda8});
}
