// me_cnn.hip - the learned motion proposer's network on gfx950: a small CNN
// over 64x64 luma tiles (current picture + co-located reference) that
// proposes candidate vectors per 16x16 block, with its training kernels.
//
// Network (per 64x64 tile, 3x3 convolutions with a one-pixel zero border):
//   input  2 ch  64x64 : (cur - 128) / 64, (ref - 128) / 64
//   conv1  2->16,  stride 2 -> 32x32
//   conv2  16->32, stride 2 -> 16x16
//   conv3  32->32, stride 1 -> 16x16
//   conv4  32->32, stride 1 -> 16x16, ReLU after each layer
//   head   per 16x16 block (a 4x4 cell of conv4): the cell's mean (32) and
//          the 16 neighbour vectors (32 numbers, pel / 16) -> 12 outputs: 4
//          hypotheses (dx, dy) in units of 16 pel, then 4 confidence logits
// Output pixel (y, x) of a layer of stride S reads input pixels S*y - 1 ..
// S*y + 1 (rows) and S*x - 1 .. S*x + 1 (columns).  Activations and weights
// are bf16, every product accumulates in fp32.
//
// Layout: activations NHWC per tile with a one-pixel zero border: the input
// [tile][66][66][2], conv1 [tile][34][34][16], conv2..conv4
// [tile][18][18][32] bf16; only the interior is ever written by a layer, so
// the border stays the zero padding.  A convolution is an implicit GEMM,
// out^T (Cout x 16 pixels) = W (Cout x K) * im2col^T (K x 16 pixels), K =
// 9*Cin ordered (ky, kx, ci) and zero-padded to a multiple of 32, one
// v_mfma_f32_16x16x32_bf16 per 16 output channels per 32 K:
//   A (weights)  lane l: row co = l%16, k = 8*(l/16) .. +7
//   B (im2col)   lane l: col px = l%16, k = 8*(l/16) .. +7
//   D            lane l: col px = l%16, rows co = 4*(l/16) .. +3
// so a lane ends with 4 consecutive channels of one pixel (one 8-byte
// store).  A wave keeps its whole weight tile in VGPRs and walks 16-pixel
// row segments.  Input gradients of the stride-1 layers are the same kernel
// with Cin and Cout swapped and the weights flipped, the ReLU mask of the
// layer below fused into the epilogue; conv2's (stride 2) is that
// convolution over the zero-upsampled gradient, read in place.  Weight
// gradients are a second MFMA kernel over pixel chunks staged in LDS,
// reduced per tile into a partial buffer and then across tiles in a fixed
// order, so training is deterministic.
#include "bc250_dev.h"
#include "bc250_learned.h"
#include "learned_dev.h"

#include <hip/hip_bf16.h>

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int T = 64;        // input tile side
constexpr int TP = T + 2;    // padded side
constexpr int F1 = 32, F1P = F1 + 2;   // conv1 output side, padded
constexpr int F2 = 16, F2P = F2 + 2;   // conv2..conv4 output side, padded
constexpr float POOL_MEAN = BC250_CNN_POOL_SCALE;  // 1 / (4 x 4) cell pixels

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// fp32 -> bf16, round to nearest even (v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint32_t pack2bf(float a, float b)
{
    const bf16x2 v = {(__bf16)a, (__bf16)b};
    return __builtin_bit_cast(uint32_t, v);
}
__device__ __forceinline__ uint16_t f2bf(float f) { return (uint16_t)pack2bf(f, 0.f); }
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Sum over each group of 4 lanes of a DPP row, complete in the group's last
// lane (3, 7, 11, 15): (v3 + v2) + (v1 + v0), a row_shr scan.
__device__ __forceinline__ float row4_sum(float v)
{
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x111, 0xf, 0xf, true));
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x112, 0xf, 0xf, true));
    return v;
}

constexpr int kpad(int cin) { return (9 * cin + 31) / 32 * 32; }

enum { MODE_RELU = 0, MODE_MASK = 1, MODE_RELU_POOL = 2 };

// im2col operand of one lane: output pixel (y, x) of a layer of stride S,
// k = k0 .. k0+7, from the padded input `in` of side INP (input element (S*y
// + ky, S*x + kx) in padded coordinates).  UP = 2: `in` is a gradient of
// padded side INP/2 + 1 read as its zero-upsampled view of side INP:
// element (j, i) is in[(j+1)/2][(i+1)/2] when j and i are both odd, else 0.
template <int CIN, int S, int INP, int UP>
__device__ __forceinline__ bf16x8 im2col8(const uint16_t *__restrict__ in, int y, int x, int k0)
{
    constexpr int BP = UP == 2 ? INP / 2 + 1 : INP;
    bf16x8 v;
    if constexpr (CIN % 8 == 0) {
        const int pos = k0 / CIN, ci = k0 % CIN;
        uint4 q = make_uint4(0, 0, 0, 0);
        if (pos < 9) {
            const int jy = S * y + pos / 3, jx = S * x + pos % 3;
            if constexpr (UP == 2) {
                if ((jy & 1) && (jx & 1))
                    q = *(const uint4 *)(in + ((size_t)((jy + 1) >> 1) * BP + ((jx + 1) >> 1)) * CIN + ci);
            } else {
                q = *(const uint4 *)(in + ((size_t)jy * INP + jx) * CIN + ci);
            }
        }
        v = __builtin_bit_cast(bf16x8, q);
    } else {
        static_assert(UP == 1, "only whole-chunk channels are read upsampled");
        uint16_t e[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int k = k0 + j, pos = k / CIN, ci = k % CIN;
            e[j] = pos < 9 ? in[((size_t)(S * y + pos / 3) * INP + (S * x + pos % 3)) * CIN + ci] : (uint16_t)0;
        }
        v = __builtin_bit_cast(bf16x8, make_uint4(e[0] | (uint32_t)e[1] << 16, e[2] | (uint32_t)e[3] << 16,
                                                  e[4] | (uint32_t)e[5] << 16, e[6] | (uint32_t)e[7] << 16));
    }
    return v;
}

// One 3x3 convolution layer over `ntiles` tiles: output side OUT (padded
// OUTP), stride S over an input of side IN (padded INP; UP as im2col8).
// w: [COUT][KPAD] bf16, bias fp32, channel c at bias[c * bstride] (null:
// none).  MODE_RELU: out = relu(conv + b).  MODE_MASK: out = conv * (mask >
// 0) (input gradients through the ReLU below).  MODE_RELU_POOL (conv4):
// relu(conv + b), its bf16 values summed over each 4-pixel row group into
// rows [tile][16 rows][4 groups][COUT] fp32 (`pool`), out optional;
// pool_rows_kernel then sums 4 rows per cell in a fixed order.
template <int CIN, int COUT, int MODE, int S, int IN, int OUT, int UP>
__global__ __launch_bounds__(256) void conv3x3_kernel(const uint16_t *__restrict__ in, const uint16_t *__restrict__ w,
                                                      const float *__restrict__ bias, uint16_t *__restrict__ out,
                                                      const uint16_t *__restrict__ mask, float *__restrict__ pool,
                                                      int ntiles, int bstride)
{
    constexpr int KS = kpad(CIN) / 32, NT = COUT / 16;
    constexpr int INP = IN + 2, OUTP = OUT + 2;
    constexpr size_t IN_TILE = UP == 2 ? (size_t)(INP / 2 + 1) * (INP / 2 + 1) * CIN : (size_t)INP * INP * CIN;
    static_assert(OUT % 16 == 0 && (MODE != MODE_RELU_POOL || OUT == F2), "row segments of 16 pixels");
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6), nwaves = gridDim.x * 4;
    const int lr = lane & 15, lg = lane >> 4;
    bf16x8 wa[NT][KS];
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
        for (int s = 0; s < KS; s++)
            wa[t][s] = __builtin_bit_cast(bf16x8, *(const uint4 *)(w + (size_t)(t * 16 + lr) * kpad(CIN) + s * 32 + 8 * lg));
    float bv[NT][4];
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) bv[t][r] = bias ? bias[(size_t)(t * 16 + 4 * lg + r) * bstride] : 0.f;

    // a task is one output row of one tile
    for (int task = wave; task < ntiles * OUT; task += nwaves) {
        const int tile = task / OUT, y = task % OUT;
        const uint16_t *tin = in + (size_t)tile * IN_TILE;
        for (int x0 = 0; x0 < OUT; x0 += 16) {
            f32x4 acc[NT];
#pragma unroll
            for (int t = 0; t < NT; t++) acc[t] = f32x4{bv[t][0], bv[t][1], bv[t][2], bv[t][3]};
#pragma unroll
            for (int s = 0; s < KS; s++) {
                const bf16x8 b = im2col8<CIN, S, INP, UP>(tin, y, x0 + lr, s * 32 + 8 * lg);
#pragma unroll
                for (int t = 0; t < NT; t++) acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa[t][s], b, acc[t], 0, 0, 0);
            }
            const size_t opix = ((size_t)tile * OUTP + (y + 1)) * OUTP + (x0 + lr + 1);
#pragma unroll
            for (int t = 0; t < NT; t++) {
                float v[4];
#pragma unroll
                for (int r = 0; r < 4; r++) v[r] = MODE != MODE_MASK ? fmaxf(acc[t][r], 0.f) : acc[t][r];
                const int c0 = t * 16 + 4 * lg;
                if (MODE == MODE_MASK) {
                    const uint2 m = *(const uint2 *)(mask + opix * COUT + c0);
                    const uint32_t mm[2] = {m.x, m.y};
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        if (bf2f((uint16_t)(mm[r >> 1] >> (16 * (r & 1)))) <= 0.f) v[r] = 0.f;
                }
                const uint2 o = make_uint2(pack2bf(v[0], v[1]), pack2bf(v[2], v[3]));
                if (out) *(uint2 *)(out + opix * COUT + c0) = o;
                if (MODE == MODE_RELU_POOL) {
                    // each 4 pixels of this row lie in one cell: sum the
                    // bf16-rounded values over the 4 lanes of a group
                    v[0] = row4_sum(__uint_as_float(o.x << 16));
                    v[1] = row4_sum(__uint_as_float(o.x & 0xffff0000u));
                    v[2] = row4_sum(__uint_as_float(o.y << 16));
                    v[3] = row4_sum(__uint_as_float(o.y & 0xffff0000u));
                    if ((lr & 3) == 3)
                        *(float4 *)(pool + (((size_t)tile * F2 + y) * 4 + (lr >> 2)) * COUT + c0) =
                            make_float4(v[0], v[1], v[2], v[3]);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Fused inference: the four convolutions and the cell sums of one tile per
// workgroup, every intermediate activation held in LDS.  Two regions
// alternate:
//   A: input (66 x 66 x 2)  -> B: conv1 (34 x 34 x 16) -> A: conv2 (18 x 18 x 32)
//   -> B: conv3 (18 x 18 x 32) -> A: conv4 row sums (16 x 4 x 32 fp32)
// (57,728 bytes: two workgroups per CU).  A layer writes its output's zero
// border itself (the region held another layout before).  The MFMA
// operands, their K order and the epilogue are the layer kernels'
// (conv3x3_kernel, pool_rows_kernel), so the cell sums are the same bits.
// conv3 and conv4 reuse each loaded input row for R output rows: a task
// walks input rows i = 0 .. R+1 and kx = 0 .. 2, and output row j takes the
// step at ky = i - j, so every accumulator still receives its K steps in the
// order (ky, kx) = (0,0), (0,1), .. (2,2).
constexpr int FA_ELEMS = F2P * F2P * 32;   // input / conv2 output / conv4 row sums
constexpr int FB_ELEMS = F1P * F1P * 16;   // conv1 / conv3 output
static_assert(TP * TP * 2 <= FA_ELEMS, "input fits A");
static_assert(F2 * 4 * 32 * 2 <= FA_ELEMS, "row sums fit A");
static_assert(F2P * F2P * 32 <= FB_ELEMS, "conv3 output fits B");

// LDS placement of the 8-channel chunks of a 16- or 32-channel pixel in a
// region of row pitch PITCH pixels.  The B operand's 16 lanes of one lane
// group read one chunk of 16 pixels S apart; unswizzled, pixels 64 bytes
// apart fall on the same banks.  SW 1 XORs the chunk index with the pixel
// bits above the ones that already spread a 16-lane read over 8 bank groups
// of 16 bytes (32 banks), SW 2 over 16 (64 banks); SW 0 is the plain NHWC
// order.
template <int C, int SW, int PITCH>
__device__ __forceinline__ int chunk_off(int row, int x, int chunk)
{
    constexpr int C8 = C / 8, LC = C8 == 4 ? 2 : 1;
    constexpr int SH = SW == 0 ? 0 : (SW == 1 ? 3 : 4) - LC;
    const int sw = SW == 0 ? 0 : (x >> SH) & (C8 - 1);
    return ((row * PITCH + x) * C8 + (chunk ^ sw)) * 8;
}

// The operands one task reads, for its 16-pixel segment: conv3 / conv4 the
// (R + 2) x 3 input chunks of its rows, conv2 its 5 im2col K steps, conv1
// its one K step.
template <int CIN, int R>
struct TaskB {
    static constexpr int PER = CIN == 32 ? (R + 2) * 3 : kpad(CIN) / 32;
    bf16x8 v[PER];
};

// Operands of the task at output rows r0 .. r0+R-1, pixels x0 .. x0+15 (this
// lane: pixel x0 + lr, lane group lg) from the padded source of side SRCP.
template <int CIN, int S, int SRCP, int R, int SW>
__device__ __forceinline__ void load_task(const uint16_t *__restrict__ src, int r0, int x0, int lr, int lg,
                                          TaskB<CIN, R> &b)
{
    const int x = x0 + lr;
    if constexpr (CIN == 32) {
        static_assert(S == 1, "conv3 / conv4");
#pragma unroll
        for (int i = 0; i < R + 2; i++)
#pragma unroll
            for (int kx = 0; kx < 3; kx++)
                b.v[i * 3 + kx] = __builtin_bit_cast(bf16x8, *(const uint4 *)(src + chunk_off<32, SW, SRCP>(r0 + i, x + kx, lg)));
    } else if constexpr (CIN == 16) {
        static_assert(R == 1 && S == 2, "conv2 walks single rows at stride 2");
        // the last K step holds position 8 in lane groups 0, 1 and zero
        // padding in 2, 3: read position 8 there too and clear it
#pragma unroll
        for (int s = 0; s < TaskB<CIN, R>::PER; s++) {
            const int pos = 2 * s + (lg >> 1), pc = pos < 9 ? pos : 8;
            uint4 q = *(const uint4 *)(src + chunk_off<16, SW, SRCP>(2 * r0 + pc / 3, 2 * x + pc % 3, lg & 1));
            if (pos >= 9) q = make_uint4(0, 0, 0, 0);
            b.v[s] = __builtin_bit_cast(bf16x8, q);
        }
    } else {
        static_assert(CIN == 2 && R == 1 && S == 2, "conv1: one K step at stride 2");
        // k = 8 lg .. 8 lg + 7 are positions 4 lg .. 4 lg + 3, two channels each
        uint32_t e[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int pos = 4 * lg + q, pc = pos < 9 ? pos : 8;
            e[q] = *(const uint32_t *)(src + ((2 * r0 + pc / 3) * SRCP + 2 * x + pc % 3) * 2);
            if (pos >= 9) e[q] = 0u;
        }
        b.v[0] = __builtin_bit_cast(bf16x8, make_uint4(e[0], e[1], e[2], e[3]));
    }
}

template <int CIN, int COUT>
struct LayerW {
    static constexpr int KP = kpad(CIN), KS = KP / 32, NT = COUT / 16;
    bf16x8 a[NT][KS];
    float b[NT][4];
};

// Each lane's share of a layer's weights (the MFMA A operand) and biases.
template <int CIN, int COUT>
__device__ __forceinline__ void load_w(const uint16_t *__restrict__ w, const float *__restrict__ bias, LayerW<CIN, COUT> &lw)
{
    using LW = LayerW<CIN, COUT>;
    const int lane = threadIdx.x & 63, lr = lane & 15, lg = lane >> 4;
#pragma unroll
    for (int t = 0; t < LW::NT; t++)
#pragma unroll
        for (int s = 0; s < LW::KS; s++)
            lw.a[t][s] = __builtin_bit_cast(bf16x8, *(const uint4 *)(w + (size_t)(t * 16 + lr) * LW::KP + s * 32 + 8 * lg));
#pragma unroll
    for (int t = 0; t < LW::NT; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) lw.b[t][r] = bias[(size_t)(t * 16 + 4 * lg + r) * (LW::KP + 1)];
}

// Zero border of a padded C-channel region of side P (rows 0 and P-1, columns
// 0 and P-1), all threads.
template <int C, int SW, int P>
__device__ __forceinline__ void zero_border(uint16_t *__restrict__ dst, int nthreads)
{
    constexpr int C8 = C / 8, NB = 4 * (P - 1);
    for (int i = threadIdx.x; i < NB * C8; i += nthreads) {
        const int b = i / C8, ch = i % C8;
        int row, x;
        if (b < P) { row = 0; x = b; }
        else if (b < 2 * P) { row = P - 1; x = b - P; }
        else if (b < 3 * P - 2) { row = b - 2 * P + 1; x = 0; }
        else { row = b - (3 * P - 2) + 1; x = P - 1; }
        *(uint4 *)(dst + chunk_off<C, SW, P>(row, x, ch)) = make_uint4(0, 0, 0, 0);
    }
}

// One layer of the fused tile: output side OUT (padded OUTP), tasks of R
// rows by one 16-pixel segment over NW waves; the source region is padded
// side SRCP.  rowsum (conv4 only): the bf16 values' 4-pixel group sums,
// [row][group][COUT] fp32.  A wave loads the next task's operands before the
// current task's epilogue, so their LDS latency overlaps the conversions
// and stores.
template <int CIN, int COUT, int S, int SRCP, int OUTP, int R, int NW, int SW>
__device__ __forceinline__ void fused_layer(const uint16_t *__restrict__ src, uint16_t *__restrict__ dst,
                                            const LayerW<CIN, COUT> &lw, float *__restrict__ rowsum)
{
    constexpr int OUT = OUTP - 2, NT = COUT / 16, SPR = OUT / 16;  // segments per row
    constexpr int NTASK = OUT / R * SPR;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, lg = lane >> 4;
    TaskB<CIN, R> b;
    if (wave < NTASK) load_task<CIN, S, SRCP, R, SW>(src, wave / SPR * R, wave % SPR * 16, lr, lg, b);
    for (int task = wave; task < NTASK; task += NW) {
        const int r0 = task / SPR * R, x0 = task % SPR * 16;
        f32x4 acc[R][NT];
#pragma unroll
        for (int j = 0; j < R; j++)
#pragma unroll
            for (int t = 0; t < NT; t++) acc[j][t] = f32x4{lw.b[t][0], lw.b[t][1], lw.b[t][2], lw.b[t][3]};
        if constexpr (CIN == 32) {
#pragma unroll
            for (int i = 0; i < R + 2; i++)
#pragma unroll
                for (int kx = 0; kx < 3; kx++)
#pragma unroll
                    for (int j = 0; j < R; j++) {
                        const int ky = i - j;
                        if (ky >= 0 && ky <= 2)
#pragma unroll
                            for (int t = 0; t < NT; t++)
                                acc[j][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(lw.a[t][3 * ky + kx], b.v[i * 3 + kx],
                                                                                    acc[j][t], 0, 0, 0);
                    }
        } else {
#pragma unroll
            for (int s = 0; s < TaskB<CIN, R>::PER; s++)
#pragma unroll
                for (int t = 0; t < NT; t++)
                    acc[0][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(lw.a[t][s], b.v[s], acc[0][t], 0, 0, 0);
        }
        __builtin_amdgcn_sched_barrier(0);
        const int nx = task + NW;
        if (nx < NTASK) load_task<CIN, S, SRCP, R, SW>(src, nx / SPR * R, nx % SPR * 16, lr, lg, b);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < R; j++) {
            const int row = r0 + j;
#pragma unroll
            for (int t = 0; t < NT; t++) {
                float v[4];
#pragma unroll
                for (int r = 0; r < 4; r++) v[r] = fmaxf(acc[j][t][r], 0.f);
                const int c0 = t * 16 + 4 * lg;
                const uint2 o = make_uint2(pack2bf(v[0], v[1]), pack2bf(v[2], v[3]));
                if (dst) *(uint2 *)(dst + chunk_off<COUT, SW, OUTP>(row + 1, x0 + lr + 1, c0 >> 3) + (c0 & 7)) = o;
                if (rowsum) {
                    v[0] = row4_sum(__uint_as_float(o.x << 16));
                    v[1] = row4_sum(__uint_as_float(o.x & 0xffff0000u));
                    v[2] = row4_sum(__uint_as_float(o.y << 16));
                    v[3] = row4_sum(__uint_as_float(o.y & 0xffff0000u));
                    if ((lr & 3) == 3)
                        *(float4 *)(rowsum + ((row * 4) + ((x0 + lr) >> 2)) * COUT + c0) =
                            make_float4(v[0], v[1], v[2], v[3]);
                }
            }
        }
    }
}

// in: input tiles [tile][66][66][2]; packed: forward weights; params: fp32
// parameters (biases); pool: [tile][16][32] cell sums of conv4.
struct FusedWeights {
    const uint16_t *w[4];   // forward weights [COUT][KPAD] per layer
    const float *b[4];      // biases, channel c at b[c * (KPAD + 1)]
};

// Where a tile's input comes from: the input tiles [tile][66][66][2]
// (FROM_TILES), one picture's 8-bit luma pair (FROM_PICTURE: tile i is
// (i % tw, i / tw) of the picture), or a batch of pictures (FROM_BATCH).
// Staged values are the input tiles' own: (sample - 128) / 64 as bf16,
// samples past the picture repeating its last row / column, zero on the
// border.
enum { FROM_TILES = 0, FROM_PICTURE = 1, FROM_BATCH = 2 };
struct FusedInput {
    const uint16_t *tiles;
    const uint8_t *cur, *ref;
    int cpitch, rpitch, w, h, tw;
};

__device__ __forceinline__ uint32_t input_pair(int c, int r)
{
    return (uint32_t)f2bf((c - 128) * (1.f / 64.f)) | ((uint32_t)f2bf((r - 128) * (1.f / 64.f)) << 16);
}

// The input tile (tx, ty) of a picture pair into q ([66][66] dwords, the
// border zero).  A tile inside the picture with 4-byte aligned rows reads
// both pictures as dwords, every thread's loads issued before their
// conversions; an edge tile repeats the picture's last row / column sample
// by sample.
template <int NTH>
__device__ __forceinline__ void stage_picture(uint32_t *__restrict__ q, const uint8_t *__restrict__ cur, int cpitch,
                                              const uint8_t *__restrict__ ref, int rpitch, int w, int h, int tx, int ty)
{
    const int tid = threadIdx.x;
    const bool inside = tx * T + T <= w && ty * T + T <= h &&
                        !(((uintptr_t)cur | (uintptr_t)ref | (uintptr_t)cpitch | (uintptr_t)rpitch) & 3);
    if (inside) {
        constexpr int PER = T * T / 4 / NTH;
        static_assert(PER * NTH * 4 == T * T, "whole dwords per thread");
        uint32_t a[PER], c[PER];
#pragma unroll
        for (int j = 0; j < PER; j++) {
            const int i = tid + j * NTH, r = i >> 4, d = i & 15;
            a[j] = *(const uint32_t *)(cur + (size_t)(ty * T + r) * cpitch + tx * T + 4 * d);
            c[j] = *(const uint32_t *)(ref + (size_t)(ty * T + r) * rpitch + tx * T + 4 * d);
        }
#pragma unroll
        for (int j = 0; j < PER; j++) {
            const int i = tid + j * NTH, r = i >> 4, d = i & 15;
            uint32_t *o = q + (r + 1) * TP + 1 + 4 * d;
#pragma unroll
            for (int k = 0; k < 4; k++) o[k] = input_pair((a[j] >> (8 * k)) & 255, (c[j] >> (8 * k)) & 255);
        }
        for (int i = tid; i < 4 * (TP - 1); i += NTH) {
            const int row = i < TP ? 0 : i < 2 * TP ? TP - 1 : i < 3 * TP - 2 ? i - 2 * TP + 1 : i - (3 * TP - 2) + 1;
            const int col = i < TP ? i : i < 2 * TP ? i - TP : i < 3 * TP - 2 ? 0 : TP - 1;
            q[row * TP + col] = 0u;
        }
        return;
    }
    for (int i = tid; i < TP * TP; i += NTH) {
        const int tr = i / TP - 1, tc = i % TP - 1;
        uint32_t v = 0u;
        if (tr >= 0 && tr < T && tc >= 0 && tc < T) {
            const int gy = min(ty * T + tr, h - 1), gx = min(tx * T + tc, w - 1);
            v = input_pair(cur[(size_t)gy * cpitch + gx], ref[(size_t)gy * rpitch + gx]);
        }
        q[i] = v;
    }
}

// The proposer run at the end of the tile's launch (PROPOSE): the head's
// weights, and for a single picture its previous 16x16 records, candidate
// array and range (a batch takes them from its descriptor).
struct ProposeArgs {
    const float *wh;
    const uint4 *prev16;
    int2 *cand;
    int range;
};

template <int NW, int R, int SW, int FROM, bool PROPOSE>
__device__ __forceinline__ void cnn_forward_tile(const FusedInput &src, const bc250_me_batch_t *bt,
                                                 const FusedWeights &fw, float *__restrict__ pool, int tile,
                                                 uint16_t *sa, uint16_t *sb, const ProposeArgs &pa)
{
    constexpr int NTH = NW * 64;
    const int tid = threadIdx.x;
    // the weights of conv1..conv3 are requested first (their latency hides
    // under the input staging), conv4's at the end of conv3, into conv3's
    // registers
    LayerW<2, 16> w1;
    LayerW<16, 32> w2;
    LayerW<32, 32> w3;
    load_w(fw.w[0], fw.b[0], w1);
    load_w(fw.w[1], fw.b[1], w2);
    load_w(fw.w[2], fw.b[2], w3);
    if constexpr (FROM == FROM_TILES) {
        const uint32_t *g = (const uint32_t *)(src.tiles + (size_t)tile * TP * TP * 2);
        uint32_t *q = (uint32_t *)sa;
        for (int i = tid; i < TP * TP; i += NTH) q[i] = g[i];
    } else if constexpr (FROM == FROM_PICTURE) {
        stage_picture<NTH>((uint32_t *)sa, src.cur, src.cpitch, src.ref, src.rpitch, src.w, src.h, tile % src.tw,
                           tile / src.tw);
    } else {
        const bc250_batch_frame_t &fr = bt->f[batch_frame_of_tile(*bt, tile)];
        const int tw = (fr.wc / 16 + 3) / 4, lt = tile - fr.tile_base;
        stage_picture<NTH>((uint32_t *)sa, fr.cur, fr.cpitch, fr.ref, fr.rpitch, fr.wc, fr.hc, lt % tw, lt / tw);
    }
    __syncthreads();
    zero_border<16, SW, F1P>(sb, NTH);
    fused_layer<2, 16, 2, TP, F1P, 1, NW, SW>(sa, sb, w1, nullptr);
    __syncthreads();
    zero_border<32, SW, F2P>(sa, NTH);
    fused_layer<16, 32, 2, F1P, F2P, 1, NW, SW>(sb, sa, w2, nullptr);
    __syncthreads();
    zero_border<32, SW, F2P>(sb, NTH);
    fused_layer<32, 32, 1, F2P, F2P, R, NW, SW>(sa, sb, w3, nullptr);
    load_w(fw.w[3], fw.b[3], w3);
    __syncthreads();
    float *rows = (float *)sa;
    fused_layer<32, 32, 1, F2P, F2P, R, NW, SW>(sb, nullptr, w3, rows);
    __syncthreads();
    float *spool = (float *)sb;  // conv3's region is free again
    for (int i = tid; i < 16 * 32; i += NTH) {
        const int cell = i >> 5, c = i & 31, cy = cell >> 2, cx = cell & 3;
        float s = 0.f;
        for (int y = 4 * cy; y < 4 * cy + 4; y++) s += rows[(y * 4 + cx) * 32 + c];
        pool[((size_t)tile * 16 + cell) * 32 + c] = s;
        if constexpr (PROPOSE) spool[i] = s;
    }
    if constexpr (PROPOSE) {
        __syncthreads();
        float(*smv)[32] = (float(*)[32])(spool + 16 * 32);
        float(*sz)[BC250_CNN_NOUT] = (float(*)[BC250_CNN_NOUT])(spool + 2 * 16 * 32);
        if constexpr (FROM == FROM_BATCH) {
            const bc250_batch_frame_t &fr = bt->f[batch_frame_of_tile(*bt, tile)];
            const int mbw = fr.wc / 16, tw = (mbw + 3) / 4, lt = tile - fr.tile_base;
            propose_tile(spool, pa.wh, (const uint4 *)fr.prev16, mbw, fr.hc / 16, lt % tw, lt / tw, bt->range,
                         (int2 *)fr.cand, smv, sz, NTH);
        } else {
            propose_tile(spool, pa.wh, pa.prev16, src.w / 16, src.h / 16, tile % src.tw, tile / src.tw, pa.range,
                         pa.cand, smv, sz, NTH);
        }
    }
}

// The launch shapes of the fused forward (the serving autotuner's choice):
// variant 0: 4 waves, 4-row tasks, plain LDS order; 1: 4 waves, 4-row tasks,
// 32-bank swizzle; 2: 8 waves, 2-row tasks, 64-bank swizzle.
template <int V> struct FwdShape;
template <> struct FwdShape<0> { static constexpr int NW = 4, R = 4, SW = 0; };
template <> struct FwdShape<1> { static constexpr int NW = 4, R = 4, SW = 1; };
template <> struct FwdShape<2> { static constexpr int NW = 8, R = 2, SW = 2; };

template <int V>
__global__ __launch_bounds__(FwdShape<V>::NW * 64) void cnn_forward_fused_kernel(const uint16_t *__restrict__ in,
                                                                                const FusedWeights fw,
                                                                                float *__restrict__ pool, int ntiles)
{
    using FS = FwdShape<V>;
    __shared__ __attribute__((aligned(16))) uint16_t sa[FA_ELEMS];
    __shared__ __attribute__((aligned(16))) uint16_t sb[FB_ELEMS];
    const int tile = blockIdx.x;
    if (tile >= ntiles) return;
    FusedInput src{};
    src.tiles = in;
    cnn_forward_tile<FS::NW, FS::R, FS::SW, FROM_TILES, false>(src, nullptr, fw, pool, tile, sa, sb, ProposeArgs{});
}

template <int V, bool PROPOSE>
__global__ __launch_bounds__(FwdShape<V>::NW * 64) void cnn_forward_picture_kernel(const FusedInput src,
                                                                                  const FusedWeights fw,
                                                                                  float *__restrict__ pool, int ntiles,
                                                                                  const ProposeArgs pa)
{
    using FS = FwdShape<V>;
    __shared__ __attribute__((aligned(16))) uint16_t sa[FA_ELEMS];
    __shared__ __attribute__((aligned(16))) uint16_t sb[FB_ELEMS];
    const int tile = blockIdx.x;
    if (tile >= ntiles) return;
    cnn_forward_tile<FS::NW, FS::R, FS::SW, FROM_PICTURE, PROPOSE>(src, nullptr, fw, pool, tile, sa, sb, pa);
}

template <int V, bool PROPOSE>
__global__ __launch_bounds__(FwdShape<V>::NW * 64) void cnn_forward_batch_kernel(const bc250_me_batch_t b,
                                                                                const FusedWeights fw,
                                                                                float *__restrict__ pool,
                                                                                const ProposeArgs pa)
{
    using FS = FwdShape<V>;
    __shared__ __attribute__((aligned(16))) uint16_t sa[FA_ELEMS];
    __shared__ __attribute__((aligned(16))) uint16_t sb[FB_ELEMS];
    const int tile = blockIdx.x;
    if (tile >= b.ntile_total) return;
    FusedInput src{};
    cnn_forward_tile<FS::NW, FS::R, FS::SW, FROM_BATCH, PROPOSE>(src, &b, fw, pool, tile, sa, sb, pa);
}

// Weight (and bias) gradients of one layer (stride S, input side IN, output
// side OUT): dW[co][k] = sum over output pixels p of dz[p][co] *
// im2col(a)[p][k], db[co] = sum dz[p][co].  One workgroup per tile, chunks of
// 32 output pixels staged in LDS; part[tile] = [COUT][KPAD + 1] fp32 (the
// last column is db).
template <int CIN, int COUT, int S, int IN, int OUT>
__global__ __launch_bounds__(256) void conv_wgrad_kernel(const uint16_t *__restrict__ a, const uint16_t *__restrict__ dz,
                                                         float *__restrict__ part)
{
    constexpr int KP = kpad(CIN), NT = COUT / 16, NK = KP / 16, NOUT = NT * NK;
    constexpr int PER_WAVE = (NOUT + 3) / 4;
    constexpr int INP = IN + 2, OUTP = OUT + 2;
    __shared__ uint16_t sa[32][KP + 8];   // im2col rows of 32 pixels
    __shared__ uint16_t sd[32][COUT + 8]; // dz of the same pixels
    const int tile = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 15, lg = lane >> 4;
    const uint16_t *ta = a + (size_t)tile * INP * INP * CIN;
    const uint16_t *td = dz + (size_t)tile * OUTP * OUTP * COUT;
    f32x4 acc[PER_WAVE];
#pragma unroll
    for (int i = 0; i < PER_WAVE; i++) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
    float dbs = 0.f;
    for (int chunk = 0; chunk < OUT * OUT / 32; chunk++) {
        __syncthreads();
        for (int i = tid; i < 32 * KP; i += 256) {
            const int p = i / KP, k = i % KP, pos = k / CIN, ci = k % CIN;
            const int y = (chunk * 32 + p) / OUT, x = (chunk * 32 + p) % OUT;
            sa[p][k] = pos < 9 ? ta[((size_t)(S * y + pos / 3) * INP + (S * x + pos % 3)) * CIN + ci] : (uint16_t)0;
        }
        for (int i = tid; i < 32 * COUT; i += 256) {
            const int p = i / COUT, c = i % COUT;
            const int y = (chunk * 32 + p) / OUT, x = (chunk * 32 + p) % OUT;
            sd[p][c] = td[((size_t)(y + 1) * OUTP + (x + 1)) * COUT + c];
        }
        __syncthreads();
        if (tid < COUT)
            for (int p = 0; p < 32; p++) dbs += bf2f(sd[p][tid]);
#pragma unroll
        for (int i = 0; i < PER_WAVE; i++) {
            const int o = wave * PER_WAVE + i;
            if (o < NOUT) {
                const int t = o / NK, n = o % NK;
                uint16_t ea[8], eb[8];
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    ea[j] = sd[8 * lg + j][t * 16 + lr];   // A: co x pixel
                    eb[j] = sa[8 * lg + j][n * 16 + lr];   // B: pixel x k
                }
                const bf16x8 va = __builtin_bit_cast(
                    bf16x8, make_uint4(ea[0] | (uint32_t)ea[1] << 16, ea[2] | (uint32_t)ea[3] << 16,
                                       ea[4] | (uint32_t)ea[5] << 16, ea[6] | (uint32_t)ea[7] << 16));
                const bf16x8 vb = __builtin_bit_cast(
                    bf16x8, make_uint4(eb[0] | (uint32_t)eb[1] << 16, eb[2] | (uint32_t)eb[3] << 16,
                                       eb[4] | (uint32_t)eb[5] << 16, eb[6] | (uint32_t)eb[7] << 16));
                acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va, vb, acc[i], 0, 0, 0);
            }
        }
    }
    float *pt = part + (size_t)tile * COUT * (KP + 1);
#pragma unroll
    for (int i = 0; i < PER_WAVE; i++) {
        const int o = wave * PER_WAVE + i;
        if (o < NOUT) {
            const int t = o / NK, n = o % NK;
#pragma unroll
            for (int r = 0; r < 4; r++) pt[(size_t)(t * 16 + 4 * lg + r) * (KP + 1) + n * 16 + lr] = acc[i][r];
        }
    }
    if (tid < COUT) pt[(size_t)tid * (KP + 1) + KP] = dbs;
}

// Cell sums of conv4 from its row-group sums: pool[tile][cy*4+cx][c] =
// sum over y in 4 cy .. 4 cy + 3 of rows[tile][y][cx][c], in order.
__global__ void pool_rows_kernel(const float *__restrict__ rows, float *__restrict__ pool, int ntiles)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;  // (tile, cell, c)
    if (i >= ntiles * 16 * 32) return;
    const int c = i % 32, cell = (i / 32) % 16, tile = i / 512;
    const int cy = cell >> 2, cx = cell & 3;
    float s = 0.f;
    for (int y = 4 * cy; y < 4 * cy + 4; y++) s += rows[(((size_t)tile * F2 + y) * 4 + cx) * 32 + c];
    pool[i] = s;
}

// Column sums over `n` rows of `m` floats, fixed order: out[j] = sum_i in[i][j].
__global__ void sum_rows_kernel(const float *__restrict__ in, int n, int m, float *__restrict__ out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    float s = 0.f;
    for (int i = 0; i < n; i++) s += in[(size_t)i * m + j];
    out[j] = s;
}

// Input tiles: (cur - 128) / 64 and (ref - 128) / 64, bf16, into the
// interior of [tile][66][66][2]; samples beyond the picture are the nearest
// edge sample.  tiles[i] = {picture slot, tile x, tile y}; the picture pair
// of slot s is cur + s * slot_stride, ref + s * slot_stride.
__global__ void cnn_input_kernel(const uint8_t *__restrict__ cur, int cpitch, const uint8_t *__restrict__ ref,
                                 int rpitch, size_t slot_stride, int w, int h, const int3 *__restrict__ tiles,
                                 uint16_t *__restrict__ in, int ntiles)
{
    const int i = blockIdx.x, t = threadIdx.x;  // 256 threads: 16 pixels each
    if (i >= ntiles) return;
    const int3 tl = tiles[i];
    const uint8_t *c = cur + (size_t)tl.x * slot_stride, *r = ref + (size_t)tl.x * slot_stride;
    for (int p = t; p < T * T; p += 256) {
        const int yy = p / T, xx = p % T;
        const int gy = min(tl.z * T + yy, h - 1), gx = min(tl.y * T + xx, w - 1);
        const float a = ((int)c[(size_t)gy * cpitch + gx] - 128) * (1.f / 64.f);
        const float b = ((int)r[(size_t)gy * rpitch + gx] - 128) * (1.f / 64.f);
        *(uint32_t *)(in + (((size_t)i * TP + yy + 1) * TP + xx + 1) * 2) = (uint32_t)f2bf(a) | ((uint32_t)f2bf(b) << 16);
    }
}

// Input tiles of a batch of pictures: blockIdx = batch-wide tile.
__global__ void cnn_input_batch_kernel(const bc250_me_batch_t b, uint16_t *__restrict__ in)
{
    const int i = blockIdx.x, t = threadIdx.x;
    if (i >= b.ntile_total) return;
    const int f = batch_frame_of_tile(b, i);
    const bc250_batch_frame_t &fr = b.f[f];
    const int tw = (fr.wc / 16 + 3) / 4, lt = i - fr.tile_base, tx = lt % tw, ty = lt / tw;
    for (int p = t; p < T * T; p += 256) {
        const int yy = p / T, xx = p % T;
        const int gy = min(ty * T + yy, fr.hc - 1), gx = min(tx * T + xx, fr.wc - 1);
        const float a = ((int)fr.cur[(size_t)gy * fr.cpitch + gx] - 128) * (1.f / 64.f);
        const float c = ((int)fr.ref[(size_t)gy * fr.rpitch + gx] - 128) * (1.f / 64.f);
        *(uint32_t *)(in + (((size_t)i * TP + yy + 1) * TP + xx + 1) * 2) = (uint32_t)f2bf(a) | ((uint32_t)f2bf(c) << 16);
    }
}

// Head of one 16x16 block: z = Wh [pool/16 ; mv] + bh (12 outputs): pool is the
// sum over the block's 4x4 cell of conv4, pool/16 its mean.
__device__ __forceinline__ void head_block(const float *__restrict__ wh, const float *__restrict__ pool,
                                           const float *__restrict__ mv, float *z)
{
#pragma unroll
    for (int o = 0; o < BC250_CNN_NOUT; o++) {
        const float *row = wh + o * (BC250_CNN_NFEAT + 1);
        float s = row[BC250_CNN_NFEAT];
        for (int i = 0; i < 32; i++) s += row[i] * (pool[i] * POOL_MEAN);
        for (int i = 0; i < 32; i++) s += row[32 + i] * mv[i];
        z[o] = s;
    }
}

// Inference head: every block's 12 outputs into z [tile][16][12], one
// thread per (block, output) with the block's 64 inputs read as float4s
// (the same sums in the same order as head_block, each one fused
// multiply-add, so that the fused proposer of the learned search, which
// evaluates the head per macroblock, gives the same z bit for bit).
__global__ void cnn_head_kernel(const float *__restrict__ wh, const float *__restrict__ pool,
                                const float *__restrict__ mvin, float *__restrict__ z, int nblocks)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nblocks * BC250_CNN_NOUT) return;
    const int b = i / BC250_CNN_NOUT, o = i % BC250_CNN_NOUT;
    const float *row = wh + o * (BC250_CNN_NFEAT + 1);
    const float4 *pb = (const float4 *)(pool + (size_t)b * 32), *mb = (const float4 *)(mvin + (size_t)b * 32);
    float s = row[BC250_CNN_NFEAT];
    for (int q = 0; q < 8; q++) {
        const float4 v = pb[q];
        s = fmaf(row[4 * q], v.x * POOL_MEAN, s);
        s = fmaf(row[4 * q + 1], v.y * POOL_MEAN, s);
        s = fmaf(row[4 * q + 2], v.z * POOL_MEAN, s);
        s = fmaf(row[4 * q + 3], v.w * POOL_MEAN, s);
    }
    for (int q = 0; q < 8; q++) {
        const float4 v = mb[q];
        s = fmaf(row[32 + 4 * q], v.x, s);
        s = fmaf(row[32 + 4 * q + 1], v.y, s);
        s = fmaf(row[32 + 4 * q + 2], v.z, s);
        s = fmaf(row[32 + 4 * q + 3], v.w, s);
    }
    z[i] = s;
}

__device__ __forceinline__ float huber(float d, float delta, float *g)
{
    const float a = fabsf(d);
    if (a <= delta) {
        *g = d / delta;
        return 0.5f * d * d / delta;
    }
    *g = d > 0 ? 1.f : -1.f;
    return a - 0.5f * delta;
}

// Training head, one thread per tile: forward, winner-takes-most loss and
// backward of its 16 blocks.  target [tile][16][2] (pel); loss_out[tile];
// dpool [tile][16][32] (d loss / d block sum of conv4); part [tile] = the
// gradient of Wh and bh, laid out like wh.  Loss per block: the Huber loss
// (delta 1/16) of the hypothesis nearest the target, EPS times that of the
// others, and the cross entropy of the confidences against the nearest
// hypothesis' index.
__global__ void cnn_head_train_kernel(const float *__restrict__ wh, const float *__restrict__ pool,
                                      const float *__restrict__ mvin, const float *__restrict__ target,
                                      float *__restrict__ loss_out, float *__restrict__ dpool,
                                      float *__restrict__ part, int ntiles, float scale)
{
    const int tile = blockIdx.x * blockDim.x + threadIdx.x;
    if (tile >= ntiles) return;
    constexpr int NIN = BC250_CNN_NFEAT + 1;
    float *pt = part + (size_t)tile * BC250_CNN_NOUT * NIN;
    for (int i = 0; i < BC250_CNN_NOUT * NIN; i++) pt[i] = 0.f;
    float total = 0.f;
    for (int j = 0; j < 16; j++) {
        const size_t b = (size_t)tile * 16 + j;
        const float *pb = pool + b * 32, *mb = mvin + b * 32;
        float z[BC250_CNN_NOUT];
        head_block(wh, pb, mb, z);
        const float tx = target[b * 2] * (1.f / 16.f), ty = target[b * 2 + 1] * (1.f / 16.f);
        float l[4], gx[4], gy[4];
        int best = 0;
        for (int k = 0; k < 4; k++) {
            l[k] = huber(z[2 * k] - tx, 1.f / 16.f, &gx[k]) + huber(z[2 * k + 1] - ty, 1.f / 16.f, &gy[k]);
            if (l[k] < l[best]) best = k;
        }
        float mx = z[8];
        for (int k = 1; k < 4; k++) mx = fmaxf(mx, z[8 + k]);
        float e[4], se = 0.f;
        for (int k = 0; k < 4; k++) { e[k] = __expf(z[8 + k] - mx); se += e[k]; }
        float dz[BC250_CNN_NOUT];
        float lb = 0.f;
        for (int k = 0; k < 4; k++) {
            const float wk = k == best ? 1.f : BC250_CNN_EPS;
            lb += wk * l[k];
            dz[2 * k] = wk * gx[k];
            dz[2 * k + 1] = wk * gy[k];
            dz[8 + k] = e[k] / se - (k == best ? 1.f : 0.f);
        }
        lb += -(z[8 + best] - mx - __logf(se));
        total += lb;
        for (int o = 0; o < BC250_CNN_NOUT; o++) dz[o] *= scale;
        float feat[BC250_CNN_NFEAT];
        for (int i = 0; i < 32; i++) { feat[i] = pb[i] * POOL_MEAN; feat[32 + i] = mb[i]; }
        for (int o = 0; o < BC250_CNN_NOUT; o++) {
            float *row = pt + o * NIN;
            for (int i = 0; i < BC250_CNN_NFEAT; i++) row[i] += dz[o] * feat[i];
            row[BC250_CNN_NFEAT] += dz[o];
        }
        for (int i = 0; i < 32; i++) {
            float s = 0.f;
            for (int o = 0; o < BC250_CNN_NOUT; o++) s += wh[o * NIN + i] * dz[o];
            dpool[b * 32 + i] = s * POOL_MEAN;
        }
    }
    loss_out[tile] = total;
}

// d conv4 output -> its pre-activation gradient: dz4[p][c] = dpool[cell(p)][c]
// where conv4's output a4[p][c] > 0, else 0 (bf16, padded layout).
__global__ void cnn_pool_backward_kernel(const float *__restrict__ dpool, const uint16_t *__restrict__ a4,
                                         uint16_t *__restrict__ dz4, int ntiles)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)ntiles * F2 * F2 * 32;
    if (i >= n) return;
    const int c = (int)(i % 32);
    const size_t p = i / 32;
    const int tile = (int)(p / (F2 * F2)), yy = (int)(p % (F2 * F2)) / F2, xx = (int)(p % F2);
    const size_t o = (((size_t)tile * F2P + yy + 1) * F2P + xx + 1) * 32 + c;
    const float g = dpool[((size_t)tile * 16 + (yy >> 2) * 4 + (xx >> 2)) * 32 + c];
    dz4[o] = bf2f(a4[o]) > 0.f ? f2bf(g) : (uint16_t)0;
}

// Adam on the flat fp32 parameters (bias-corrected; t >= 1).
__global__ void adam_kernel(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, int n, float lr, float b1, float b2, float eps, int t)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi, vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.f - __powf(b1, (float)t)), vh = vi / (1.f - __powf(b2, (float)t));
    p[i] -= lr * mh / (sqrtf(vh) + eps);
}

// bf16 copies of the conv weights (fp32 [COUT][KPAD + 1], bias last) in both
// layouts: forward [COUT][KPAD] and, for input gradients, flipped and
// transposed [CIN][KPAD'] with K' = (8 - pos) * COUT + co.
__global__ void pack_weights_kernel(const float *__restrict__ w, int cin, int cout, uint16_t *__restrict__ fwd,
                                    uint16_t *__restrict__ flip)
{
    const int kp = kpad(cin), kq = kpad(cout);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cout * kp) fwd[i] = f2bf(w[(size_t)(i / kp) * (kp + 1) + i % kp]);
    if (flip && i < cin * kq) {
        const int ci = i / kq, k = i % kq, pos = k / cout, co = k % cout;
        flip[i] = pos < 9 ? f2bf(w[(size_t)co * (kp + 1) + (8 - pos) * cin + ci]) : (uint16_t)0;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI (include/bc250_learned.h)

extern "C" size_t bc250_cnn_param_count(void) { return BC250_CNN_NPARAM; }

extern "C" size_t bc250_cnn_tile_bytes(int layer)
{
    switch (layer) {
    case 0: return (size_t)TP * TP * 2 * 2;
    case 1: return (size_t)F1P * F1P * 16 * 2;
    case 2: case 3: case 4: return (size_t)F2P * F2P * 32 * 2;
    default: return 0;
    }
}

extern "C" int bc250_cnn_input(const uint8_t *cur, int cpitch, const uint8_t *ref, int rpitch, size_t slot_stride,
                               int w, int h, const int *tiles, int ntiles, uint16_t *in, void *stream)
{
    if (!cur || !ref || !tiles || !in || ntiles <= 0 || w <= 0 || h <= 0 || cpitch < w || rpitch < w) return -1;
    cnn_input_kernel<<<ntiles, 256, 0, (hipStream_t)stream>>>(cur, cpitch, ref, rpitch, slot_stride, w, h,
                                                            (const int3 *)tiles, in, ntiles);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Workgroups per convolution launch: enough to fill the chip, never more
// than there are rows; the cap is a tunable (bc250_cnn_set_grid_cap).
static int g_conv_grid_cap = 4096;
extern "C" int bc250_cnn_set_grid_cap(int cap)
{
    if (cap < 64 || cap > 65536) return -1;
    g_conv_grid_cap = cap;
    return 0;
}
static int conv_grid(int ntiles, int out_side)
{
    const int rows = ntiles * out_side;
    const int g = (rows + 3) / 4;
    return g < g_conv_grid_cap ? g : g_conv_grid_cap;
}

extern "C" int bc250_cnn_conv(int layer, int mode, const uint16_t *in, const uint16_t *w, const float *bias,
                              uint16_t *out, const uint16_t *mask, float *pool, int ntiles, void *stream)
{
    if (ntiles <= 0 || !in || !w || layer < 1 || layer > 4) return -1;
    // the bias is the last column of the layer's [COUT][KPAD + 1] parameter matrix
    const int bstride = kpad(layer == 1 ? 2 : layer == 2 ? 16 : 32) + 1;
    hipStream_t st = (hipStream_t)stream;
    // forward: mode 0 (ReLU) or 2 (ReLU + cell sums, layer 4 only);
    // input gradients: mode 1, layer = the layer whose input is differentiated
    // (layer 2's stride-2 input gradient reads its dz zero-upsampled)
    if (mode == MODE_RELU) {
        if (!out) return -1;
        switch (layer) {
        case 1:
            conv3x3_kernel<2, 16, MODE_RELU, 2, T, F1, 1><<<conv_grid(ntiles, F1), 256, 0, st>>>(
                in, w, bias, out, nullptr, nullptr, ntiles, bstride);
            break;
        case 2:
            conv3x3_kernel<16, 32, MODE_RELU, 2, F1, F2, 1><<<conv_grid(ntiles, F2), 256, 0, st>>>(
                in, w, bias, out, nullptr, nullptr, ntiles, bstride);
            break;
        case 3:
        case 4:
            conv3x3_kernel<32, 32, MODE_RELU, 1, F2, F2, 1><<<conv_grid(ntiles, F2), 256, 0, st>>>(
                in, w, bias, out, nullptr, nullptr, ntiles, bstride);
            break;
        default: return -1;
        }
    } else if (mode == MODE_RELU_POOL) {
        if (layer != 4 || !pool) return -1;
        conv3x3_kernel<32, 32, MODE_RELU_POOL, 1, F2, F2, 1><<<conv_grid(ntiles, F2), 256, 0, st>>>(
            in, w, bias, out, nullptr, pool, ntiles, bstride);
    } else if (mode == MODE_MASK) {
        if (!out || !mask) return -1;
        switch (layer) {
        case 2:
            conv3x3_kernel<32, 16, MODE_MASK, 1, F1, F1, 2><<<conv_grid(ntiles, F1), 256, 0, st>>>(
                in, w, nullptr, out, mask, nullptr, ntiles, 0);
            break;
        case 3:
        case 4:
            conv3x3_kernel<32, 32, MODE_MASK, 1, F2, F2, 1><<<conv_grid(ntiles, F2), 256, 0, st>>>(
                in, w, nullptr, out, mask, nullptr, ntiles, 0);
            break;
        default: return -1;
        }
    } else {
        return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// the default: in the pipeline, beside the previous frame's readback, the
// plain layout measured faster than the swizzled one (5201 vs 5180 fps at
// 1080p); the serving autotuner picks its own
static int g_forward_variant = 0;
extern "C" int bc250_cnn_forward(const uint16_t *in, const uint16_t *packed, const float *params, float *pool,
                                 int ntiles, void *stream)
{
    if (!in || !packed || !params || !pool || ntiles <= 0 || ntiles > (1 << 28)) return -1;
    const int cin[4] = {2, 16, 32, 32};
    FusedWeights fw;
    for (int l = 0; l < 4; l++) {
        fw.w[l] = packed + bc250_cnn_packed_offset(l, 0);
        fw.b[l] = params + bc250_cnn_w_offset(l) + kpad(cin[l]);
    }
    hipStream_t st = (hipStream_t)stream;
    switch (g_forward_variant) {
    case 0: cnn_forward_fused_kernel<0><<<ntiles, FwdShape<0>::NW * 64, 0, st>>>(in, fw, pool, ntiles); break;
    case 1: cnn_forward_fused_kernel<1><<<ntiles, FwdShape<1>::NW * 64, 0, st>>>(in, fw, pool, ntiles); break;
    default: cnn_forward_fused_kernel<2><<<ntiles, FwdShape<2>::NW * 64, 0, st>>>(in, fw, pool, ntiles); break;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

static FusedWeights fused_weights(const uint16_t *packed, const float *params)
{
    const int cin[4] = {2, 16, 32, 32};
    FusedWeights fw;
    for (int l = 0; l < 4; l++) {
        fw.w[l] = packed + bc250_cnn_packed_offset(l, 0);
        fw.b[l] = params + bc250_cnn_w_offset(l) + kpad(cin[l]);
    }
    return fw;
}

template <bool PROPOSE>
static int forward_picture(const uint8_t *cur, int cpitch, const uint8_t *ref, int rpitch, int wc, int hc,
                           const uint16_t *packed, const float *params, float *pool, const ProposeArgs &pa,
                           void *stream)
{
    if (!cur || !ref || !packed || !params || !pool || wc <= 0 || hc <= 0 || (wc & 15) || (hc & 15) ||
        cpitch < wc || rpitch < wc)
        return -1;
    FusedInput src{};
    src.cur = cur;
    src.ref = ref;
    src.cpitch = cpitch;
    src.rpitch = rpitch;
    src.w = wc;
    src.h = hc;
    src.tw = (wc / 16 + 3) / 4;
    const int ntiles = src.tw * ((hc / 16 + 3) / 4);
    const FusedWeights fw = fused_weights(packed, params);
    hipStream_t st = (hipStream_t)stream;
    switch (g_forward_variant) {
    case 0:
        cnn_forward_picture_kernel<0, PROPOSE><<<ntiles, FwdShape<0>::NW * 64, 0, st>>>(src, fw, pool, ntiles, pa);
        break;
    case 1:
        cnn_forward_picture_kernel<1, PROPOSE><<<ntiles, FwdShape<1>::NW * 64, 0, st>>>(src, fw, pool, ntiles, pa);
        break;
    default:
        cnn_forward_picture_kernel<2, PROPOSE><<<ntiles, FwdShape<2>::NW * 64, 0, st>>>(src, fw, pool, ntiles, pa);
        break;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int bc250_cnn_forward_picture(const uint8_t *cur, int cpitch, const uint8_t *ref, int rpitch, int wc,
                                         int hc, const uint16_t *packed, const float *params, float *pool,
                                         void *stream)
{
    return forward_picture<false>(cur, cpitch, ref, rpitch, wc, hc, packed, params, pool, ProposeArgs{}, stream);
}

extern "C" int bc250_cnn_forward_propose_picture(const uint8_t *cur, int cpitch, const uint8_t *ref, int rpitch,
                                                 int wc, int hc, const uint16_t *packed, const float *params,
                                                 float *pool, const uint32_t *prev16, int range, int *cand,
                                                 void *stream)
{
    if (!cand || range < 4 || range > 64) return -1;
    const ProposeArgs pa{params ? params + bc250_cnn_w_offset(4) : nullptr, (const uint4 *)prev16, (int2 *)cand,
                         range};
    return forward_picture<true>(cur, cpitch, ref, rpitch, wc, hc, packed, params, pool, pa, stream);
}

template <bool PROPOSE>
static int forward_batch(const bc250_me_batch_t *b, const uint16_t *packed, const float *params, float *pool,
                         void *stream)
{
    if (!b || b->n < 1 || b->n > BC250_BATCH_MAX || b->ntile_total <= 0 || !packed || !params || !pool) return -1;
    const FusedWeights fw = fused_weights(packed, params);
    const ProposeArgs pa{params + bc250_cnn_w_offset(4), nullptr, nullptr, b->range};
    hipStream_t st = (hipStream_t)stream;
    const int g = b->ntile_total;
    switch (g_forward_variant) {
    case 0: cnn_forward_batch_kernel<0, PROPOSE><<<g, FwdShape<0>::NW * 64, 0, st>>>(*b, fw, pool, pa); break;
    case 1: cnn_forward_batch_kernel<1, PROPOSE><<<g, FwdShape<1>::NW * 64, 0, st>>>(*b, fw, pool, pa); break;
    default: cnn_forward_batch_kernel<2, PROPOSE><<<g, FwdShape<2>::NW * 64, 0, st>>>(*b, fw, pool, pa); break;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int bc250_cnn_forward_batch(const bc250_me_batch_t *b, const uint16_t *packed, const float *params,
                                       float *pool, void *stream)
{
    return forward_batch<false>(b, packed, params, pool, stream);
}

extern "C" int bc250_cnn_forward_propose_batch(const bc250_me_batch_t *b, const uint16_t *packed,
                                               const float *params, float *pool, void *stream)
{
    if (b && (b->range < 4 || b->range > 64)) return -1;
    return forward_batch<true>(b, packed, params, pool, stream);
}

// The fused forward's launch shape (0, 1, 2: see FwdShape); a tunable.
extern "C" int bc250_cnn_set_forward_variant(int v)
{
    if (v < 0 || v > 2) return -1;
    g_forward_variant = v;
    return 0;
}

extern "C" int bc250_cnn_pool(const float *rows, float *pool, int ntiles, void *stream)
{
    if (!rows || !pool || ntiles <= 0) return -1;
    const int n = ntiles * 16 * 32;
    pool_rows_kernel<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>(rows, pool, ntiles);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int bc250_cnn_wgrad(int layer, const uint16_t *a, const uint16_t *dz, float *part, float *grad,
                               int ntiles, void *stream)
{
    if (ntiles <= 0 || !a || !dz || !part || !grad) return -1;
    hipStream_t st = (hipStream_t)stream;
    int cin, cout;
    switch (layer) {
    case 1: conv_wgrad_kernel<2, 16, 2, T, F1><<<ntiles, 256, 0, st>>>(a, dz, part); cin = 2; cout = 16; break;
    case 2: conv_wgrad_kernel<16, 32, 2, F1, F2><<<ntiles, 256, 0, st>>>(a, dz, part); cin = 16; cout = 32; break;
    case 3:
    case 4: conv_wgrad_kernel<32, 32, 1, F2, F2><<<ntiles, 256, 0, st>>>(a, dz, part); cin = 32; cout = 32; break;
    default: return -1;
    }
    const int m = cout * (kpad(cin) + 1);
    sum_rows_kernel<<<(m + 255) / 256, 256, 0, st>>>(part, ntiles, m, grad);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int bc250_cnn_head(const float *wh, const float *pool, const float *mvin, float *z, int ntiles,
                              void *stream)
{
    if (ntiles <= 0 || !wh || !pool || !mvin || !z) return -1;
    const int nb = ntiles * 16;
    cnn_head_kernel<<<(nb * BC250_CNN_NOUT + 255) / 256, 256, 0, (hipStream_t)stream>>>(wh, pool, mvin, z, nb);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int bc250_cnn_head_train(const float *wh, const float *pool, const float *mvin, const float *target,
                                    float *loss, float *dpool, float *part, float *grad, int ntiles, float scale,
                                    void *stream)
{
    if (ntiles <= 0 || !wh || !pool || !mvin || !target || !loss || !dpool || !part || !grad) return -1;
    hipStream_t st = (hipStream_t)stream;
    cnn_head_train_kernel<<<(ntiles + 63) / 64, 64, 0, st>>>(wh, pool, mvin, target, loss, dpool, part, ntiles, scale);
    const int m = BC250_CNN_NOUT * (BC250_CNN_NFEAT + 1);
    sum_rows_kernel<<<(m + 255) / 256, 256, 0, st>>>(part, ntiles, m, grad);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int bc250_cnn_pool_backward(const float *dpool, const uint16_t *a4, uint16_t *dz4, int ntiles,
                                       void *stream)
{
    if (ntiles <= 0 || !dpool || !a4 || !dz4) return -1;
    const size_t n = (size_t)ntiles * F2 * F2 * 32;
    cnn_pool_backward_kernel<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(dpool, a4, dz4, ntiles);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int bc250_cnn_adam(float *p, const float *g, float *m, float *v, int n, float lr, int t, void *stream)
{
    if (!p || !g || !m || !v || n <= 0 || t < 1) return -1;
    adam_kernel<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>(p, g, m, v, n, lr, 0.9f, 0.999f, 1e-8f, t);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int bc250_cnn_pack(const float *params, uint16_t *packed, void *stream)
{
    if (!params || !packed) return -1;
    hipStream_t st = (hipStream_t)stream;
    const int cin[4] = {2, 16, 32, 32}, cout[4] = {16, 32, 32, 32};
    for (int l = 0; l < 4; l++) {
        const int n = cout[l] * kpad(cin[l]), nf = cin[l] * kpad(cout[l]);
        const int m = n > nf ? n : nf;
        pack_weights_kernel<<<(m + 255) / 256, 256, 0, st>>>(params + bc250_cnn_w_offset(l), cin[l], cout[l],
                                                              packed + bc250_cnn_packed_offset(l, 0),
                                                              l > 0 ? packed + bc250_cnn_packed_offset(l, 1) : nullptr);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int bc250_cnn_input_batch(const bc250_me_batch_t *b, uint16_t *in, void *stream)
{
    if (!b || !in || b->n < 1 || b->n > BC250_BATCH_MAX || b->ntile_total <= 0) return -1;
    cnn_input_batch_kernel<<<b->ntile_total, 256, 0, (hipStream_t)stream>>>(*b, in);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
