// me_cnn.hip - the learned motion proposer's network on gfx950: a small CNN
// over 64x64 luma tiles (current picture + co-located reference) that
// proposes candidate vectors per 16x16 block, with its training kernels.
//
// Network (per 64x64 tile, "same" 3x3 convolutions with a zero border):
//   input  2 ch  : (cur - 128) / 64, (ref - 128) / 64
//   conv1  2->16, conv2 16->32, conv3 32->32, conv4 32->32, ReLU after each
//   head   per 16x16 block: the block mean of conv4 (32) and the 16
//          neighbour vectors (32 numbers, pel / 16) -> 12 outputs: 4
//          hypotheses (dx, dy) in units of 16 pel, then 4 confidence logits
// Activations and weights are bf16, every product accumulates in fp32.
//
// Layout: activations NHWC per tile with a one-pixel zero border,
// [tile][66][66][C] bf16; only the 64x64 interior is ever written, so the
// border stays the zero padding.  A convolution is an implicit GEMM,
// out^T (Cout x 16 pixels) = W (Cout x K) * im2col^T (K x 16 pixels), K =
// 9*Cin ordered (ky, kx, ci) and zero-padded to a multiple of 32, one
// v_mfma_f32_16x16x32_bf16 per 16 output channels per 32 K:
//   A (weights)  lane l: row co = l%16, k = 8*(l/16) .. +7
//   B (im2col)   lane l: col px = l%16, k = 8*(l/16) .. +7
//   D            lane l: col px = l%16, rows co = 4*(l/16) .. +3
// so a lane ends with 4 consecutive channels of one pixel (one 8-byte
// store).  A wave keeps its whole weight tile in VGPRs and walks 16-pixel
// row segments.  The same kernel computes input gradients (Cin and Cout
// swapped, weights flipped) with the ReLU mask of the layer below fused into
// the epilogue.  Weight gradients are a second MFMA kernel over pixel
// chunks staged in LDS, reduced per tile into a partial buffer and then
// across tiles in a fixed order, so training is deterministic.
#include "bc250_dev.h"
#include "bc250_learned.h"

#include <hip/hip_bf16.h>

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int T = 64;        // tile side
constexpr int TP = T + 2;    // padded side

__device__ __forceinline__ uint16_t f2bf(float f)
{
    // round to nearest even (inputs are finite)
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

constexpr int kpad(int cin) { return (9 * cin + 31) / 32 * 32; }

enum { MODE_RELU = 0, MODE_MASK = 1, MODE_RELU_POOL = 2 };

// im2col operand of one lane: pixel (y, x) of tile `in` (padded
// coordinates of the output pixel are (y+1, x+1)), k = k0 .. k0+7.
template <int CIN>
__device__ __forceinline__ bf16x8 im2col8(const uint16_t *__restrict__ in, int y, int x, int k0)
{
    bf16x8 v;
    if constexpr (CIN % 8 == 0) {
        const int pos = k0 / CIN, ci = k0 % CIN;
        if (pos < 9) {
            const int ky = pos / 3, kx = pos % 3;
            const uint4 q = *(const uint4 *)(in + ((size_t)(y + ky) * TP + (x + kx)) * CIN + ci);
            v = __builtin_bit_cast(bf16x8, q);
        } else {
            v = __builtin_bit_cast(bf16x8, make_uint4(0, 0, 0, 0));
        }
    } else {
        uint16_t e[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int k = k0 + j, pos = k / CIN, ci = k % CIN;
            e[j] = pos < 9 ? in[((size_t)(y + pos / 3) * TP + (x + pos % 3)) * CIN + ci] : (uint16_t)0;
        }
        v = __builtin_bit_cast(bf16x8, make_uint4(e[0] | (uint32_t)e[1] << 16, e[2] | (uint32_t)e[3] << 16,
                                                  e[4] | (uint32_t)e[5] << 16, e[6] | (uint32_t)e[7] << 16));
    }
    return v;
}

// One 3x3 convolution layer over `ntiles` tiles.  w: [COUT][KPAD] bf16, bias
// fp32, channel c at bias[c * bstride] (null: none).  MODE_RELU: out = relu(conv + b).  MODE_MASK:
// out = conv * (mask > 0) (input gradients through the ReLU below).
// MODE_RELU_POOL: relu(conv + b), its bf16 values summed over each 16-pixel
// row segment into rows [tile][64 rows][4 segments][COUT] fp32 (`pool`), out
// optional; pool_rows_kernel then sums 16 rows per block in a fixed order.
template <int CIN, int COUT, int MODE>
__global__ __launch_bounds__(256) void conv3x3_kernel(const uint16_t *__restrict__ in, const uint16_t *__restrict__ w,
                                                      const float *__restrict__ bias, uint16_t *__restrict__ out,
                                                      const uint16_t *__restrict__ mask, float *__restrict__ pool,
                                                      int ntiles, int bstride)
{
    constexpr int KS = kpad(CIN) / 32, NT = COUT / 16;
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

    // a task is one 64-pixel row of one tile
    for (int task = wave; task < ntiles * T; task += nwaves) {
        const int tile = task / T, y = task % T;
        const uint16_t *tin = in + (size_t)tile * TP * TP * CIN;
        for (int x0 = 0; x0 < T; x0 += 16) {
            f32x4 acc[NT];
#pragma unroll
            for (int t = 0; t < NT; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < KS; s++) {
                const bf16x8 b = im2col8<CIN>(tin, y, x0 + lr, s * 32 + 8 * lg);
#pragma unroll
                for (int t = 0; t < NT; t++) acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa[t][s], b, acc[t], 0, 0, 0);
            }
            const size_t opix = ((size_t)tile * TP + (y + 1)) * TP + (x0 + lr + 1);
#pragma unroll
            for (int t = 0; t < NT; t++) {
                float v[4];
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    v[r] = acc[t][r] + bv[t][r];
                    if (MODE != MODE_MASK) v[r] = fmaxf(v[r], 0.f);
                }
                const int c0 = t * 16 + 4 * lg;
                if (MODE == MODE_MASK) {
                    const uint2 m = *(const uint2 *)(mask + opix * COUT + c0);
                    const uint32_t mm[2] = {m.x, m.y};
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        if (bf2f((uint16_t)(mm[r >> 1] >> (16 * (r & 1)))) <= 0.f) v[r] = 0.f;
                }
                if (out) {
                    uint2 o;
                    o.x = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
                    o.y = (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
                    *(uint2 *)(out + opix * COUT + c0) = o;
                }
                if (MODE == MODE_RELU_POOL) {
                    // the 16 pixels of this segment are in one 16x16 block:
                    // sum the bf16-rounded values over the 16 lanes of a row
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        float s_ = bf2f(f2bf(v[r]));
#pragma unroll
                        for (int o = 8; o > 0; o >>= 1) s_ += __shfl_xor(s_, o, 16);
                        v[r] = s_;
                    }
                    if (lr == 0)
                        *(float4 *)(pool + (((size_t)tile * T + y) * 4 + (x0 >> 4)) * COUT + c0) =
                            make_float4(v[0], v[1], v[2], v[3]);
                }
            }
        }
    }
}

// Weight (and bias) gradients of one layer: dW[co][k] = sum over pixels of
// dz[p][co] * im2col(a)[p][k], db[co] = sum dz[p][co].  One workgroup per
// tile, pixel chunks of 32 staged in LDS; part[tile] = [COUT][KPAD + 1] fp32
// (the last column is db).
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void conv_wgrad_kernel(const uint16_t *__restrict__ a, const uint16_t *__restrict__ dz,
                                                         float *__restrict__ part)
{
    constexpr int KP = kpad(CIN), NT = COUT / 16, NK = KP / 16, NOUT = NT * NK;
    constexpr int PER_WAVE = (NOUT + 3) / 4;
    __shared__ uint16_t sa[32][KP + 8];   // im2col rows of 32 pixels
    __shared__ uint16_t sd[32][COUT + 8]; // dz of the same pixels
    const int tile = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 15, lg = lane >> 4;
    const uint16_t *ta = a + (size_t)tile * TP * TP * CIN;
    const uint16_t *td = dz + (size_t)tile * TP * TP * COUT;
    f32x4 acc[PER_WAVE];
#pragma unroll
    for (int i = 0; i < PER_WAVE; i++) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
    float dbs = 0.f;
    for (int chunk = 0; chunk < T * T / 32; chunk++) {
        const int y = chunk / 2, x0 = (chunk & 1) * 32;
        __syncthreads();
        for (int i = tid; i < 32 * KP; i += 256) {
            const int p = i / KP, k = i % KP, pos = k / CIN, ci = k % CIN;
            sa[p][k] = pos < 9 ? ta[((size_t)(y + pos / 3) * TP + (x0 + p + pos % 3)) * CIN + ci] : (uint16_t)0;
        }
        for (int i = tid; i < 32 * COUT; i += 256) {
            const int p = i / COUT, c = i % COUT;
            sd[p][c] = td[((size_t)(y + 1) * TP + (x0 + p + 1)) * COUT + c];
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

// Block sums of conv4 from its row-segment sums: pool[tile][by*4+bx][c] =
// sum over y in 16 by .. 16 by + 15 of rows[tile][y][bx][c], in order.
__global__ void pool_rows_kernel(const float *__restrict__ rows, float *__restrict__ pool, int ntiles)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;  // (tile, block, c)
    if (i >= ntiles * 16 * 32) return;
    const int c = i % 32, blk = (i / 32) % 16, tile = i / 512;
    const int by = blk >> 2, bx = blk & 3;
    float s = 0.f;
    for (int y = 16 * by; y < 16 * by + 16; y++) s += rows[(((size_t)tile * T + y) * 4 + bx) * 32 + c];
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

// Head of one 16x16 block: z = Wh [pool/256 ; mv] + bh (12 outputs).
__device__ __forceinline__ void head_block(const float *__restrict__ wh, const float *__restrict__ pool,
                                           const float *__restrict__ mv, float *z)
{
#pragma unroll
    for (int o = 0; o < BC250_CNN_NOUT; o++) {
        const float *row = wh + o * (BC250_CNN_NFEAT + 1);
        float s = row[BC250_CNN_NFEAT];
        for (int i = 0; i < 32; i++) s += row[i] * (pool[i] * (1.f / 256.f));
        for (int i = 0; i < 32; i++) s += row[32 + i] * mv[i];
        z[o] = s;
    }
}

// Inference head: every block's 12 outputs into z [tile][16][12], one
// thread per (block, output) with the block's 64 inputs read as float4s
// (the same sums in the same order as head_block).
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
        s += row[4 * q] * (v.x * (1.f / 256.f));
        s += row[4 * q + 1] * (v.y * (1.f / 256.f));
        s += row[4 * q + 2] * (v.z * (1.f / 256.f));
        s += row[4 * q + 3] * (v.w * (1.f / 256.f));
    }
    for (int q = 0; q < 8; q++) {
        const float4 v = mb[q];
        s += row[32 + 4 * q] * v.x;
        s += row[32 + 4 * q + 1] * v.y;
        s += row[32 + 4 * q + 2] * v.z;
        s += row[32 + 4 * q + 3] * v.w;
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
        for (int i = 0; i < 32; i++) { feat[i] = pb[i] * (1.f / 256.f); feat[32 + i] = mb[i]; }
        for (int o = 0; o < BC250_CNN_NOUT; o++) {
            float *row = pt + o * NIN;
            for (int i = 0; i < BC250_CNN_NFEAT; i++) row[i] += dz[o] * feat[i];
            row[BC250_CNN_NFEAT] += dz[o];
        }
        for (int i = 0; i < 32; i++) {
            float s = 0.f;
            for (int o = 0; o < BC250_CNN_NOUT; o++) s += wh[o * NIN + i] * dz[o];
            dpool[b * 32 + i] = s * (1.f / 256.f);
        }
    }
    loss_out[tile] = total;
}

// d conv4 output -> its pre-activation gradient: dz4[p][c] = dpool[block(p)][c]
// where conv4's output a4[p][c] > 0, else 0 (bf16, padded layout).
__global__ void cnn_pool_backward_kernel(const float *__restrict__ dpool, const uint16_t *__restrict__ a4,
                                         uint16_t *__restrict__ dz4, int ntiles)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)ntiles * T * T * 32;
    if (i >= n) return;
    const int c = (int)(i % 32);
    const size_t p = i / 32;
    const int tile = (int)(p / (T * T)), yy = (int)(p % (T * T)) / T, xx = (int)(p % T);
    const size_t o = (((size_t)tile * TP + yy + 1) * TP + xx + 1) * 32 + c;
    const float g = dpool[((size_t)tile * 16 + (yy >> 4) * 4 + (xx >> 4)) * 32 + c];
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

extern "C" size_t bc250_cnn_tile_bytes(int channels) { return (size_t)TP * TP * channels * 2; }

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
static int conv_grid(int ntiles)
{
    const int rows = ntiles * T;
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
    const int g = conv_grid(ntiles);
    // forward: mode 0 (ReLU) or 2 (ReLU + block sums, layer 4 only);
    // input gradients: mode 1, layer = the layer whose input is differentiated
    if (mode == MODE_RELU) {
        if (!out) return -1;
        switch (layer) {
        case 1: conv3x3_kernel<2, 16, MODE_RELU><<<g, 256, 0, st>>>(in, w, bias, out, nullptr, nullptr, ntiles, bstride); break;
        case 2: conv3x3_kernel<16, 32, MODE_RELU><<<g, 256, 0, st>>>(in, w, bias, out, nullptr, nullptr, ntiles, bstride); break;
        case 3:
        case 4: conv3x3_kernel<32, 32, MODE_RELU><<<g, 256, 0, st>>>(in, w, bias, out, nullptr, nullptr, ntiles, bstride); break;
        default: return -1;
        }
    } else if (mode == MODE_RELU_POOL) {
        if (layer != 4 || !pool) return -1;
        conv3x3_kernel<32, 32, MODE_RELU_POOL><<<g, 256, 0, st>>>(in, w, bias, out, nullptr, pool, ntiles, bstride);
    } else if (mode == MODE_MASK) {
        if (!out || !mask) return -1;
        switch (layer) {
        case 2: conv3x3_kernel<32, 16, MODE_MASK><<<g, 256, 0, st>>>(in, w, nullptr, out, mask, nullptr, ntiles, 0); break;
        case 3:
        case 4: conv3x3_kernel<32, 32, MODE_MASK><<<g, 256, 0, st>>>(in, w, nullptr, out, mask, nullptr, ntiles, 0); break;
        default: return -1;
        }
    } else {
        return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
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
    case 1: conv_wgrad_kernel<2, 16><<<ntiles, 256, 0, st>>>(a, dz, part); cin = 2; cout = 16; break;
    case 2: conv_wgrad_kernel<16, 32><<<ntiles, 256, 0, st>>>(a, dz, part); cin = 16; cout = 32; break;
    case 3:
    case 4: conv_wgrad_kernel<32, 32><<<ntiles, 256, 0, st>>>(a, dz, part); cin = 32; cout = 32; break;
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
    const size_t n = (size_t)ntiles * T * T * 32;
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
