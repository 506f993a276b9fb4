/*
 * bc250_learned.h - the learned motion proposer: a small bf16 CNN on MFMA
 * that proposes candidate vectors per 16x16 block, its training kernels, and
 * the candidate search that refines the proposals in +-4 windows with the
 * plain full search as the fallback.
 *
 * No counterpart in the reference: its nearest structure is the start-
 * candidate search of the HEVC encoder (hevc_enc_template.c:958-1043, start
 * candidates at :1293-1298), whose AMVP / merge / zero starts this network
 * replaces with learned ones.  Device pointers throughout; every launcher
 * enqueues on `stream` (a hipStream_t, NULL = the null stream) and returns 0,
 * or -1 for invalid arguments or a failed launch.
 */
#ifndef BC250_LEARNED_H
#define BC250_LEARNED_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- the network --------------------------------------------------------
 * Tiles of 64x64 luma; activations bf16 [tile][66][66][C] (a zero border
 * around the 64x64 interior, which only the kernels write); see
 * csrc/me_cnn.hip for the layer list.  Parameters are one fp32 vector of
 * BC250_CNN_NPARAM: per conv layer l (1..4) a [COUT][KPAD + 1] matrix
 * (weights with K = (ky, kx, ci) zero-padded to KPAD, the bias last), then
 * the head [12][65] (32 block-mean features, 32 neighbour-vector inputs,
 * bias).  bf16 copies for the kernels ("packed"): per layer the forward
 * [COUT][KPAD] and, for layers 2..4, the flipped [CIN][KPAD(COUT)]. */
#define BC250_CNN_NOUT 12
#define BC250_CNN_NFEAT 64
#define BC250_CNN_EPS 0.02f
#define BC250_CNN_NPARAM 24956
#define BC250_CNN_NPACKED 47104

static inline size_t bc250_cnn_w_offset(int l) /* l = 0..3 conv layers, 4 = head */
{
    static const size_t off[5] = {0, 528, 5680, 14928, 24176};
    return off[l];
}
static inline size_t bc250_cnn_packed_offset(int l, int flipped)
{
    static const size_t f[4] = {0, 512, 10240, 28672}, g[4] = {0, 5632, 19456, 37888};
    return flipped ? g[l] : f[l];
}

size_t bc250_cnn_param_count(void);
size_t bc250_cnn_tile_bytes(int channels);

/* Input tiles from 8-bit luma pictures: tiles[3 i .. 3 i + 2] = {picture
 * slot, tile x, tile y}; slot s reads cur / ref + s * slot_stride (pitches
 * cpitch / rpitch, w x h samples, edge samples repeated past the picture). */
int bc250_cnn_input(const uint8_t *cur, int cpitch, const uint8_t *ref, int rpitch, size_t slot_stride, int w,
                    int h, const int *tiles, int ntiles, uint16_t *in, void *stream);

/* One layer.  mode 0: out = relu(conv(in, w) + b); mode 2 (layer 4 only):
 * the same, plus the bf16 outputs summed over each 16-pixel row segment into
 * `pool` = rows [tile][64][4][32] fp32 (out may be NULL), which
 * bc250_cnn_pool reduces to the block sums [tile][16][32]; mode 1: input
 * gradient of `layer` -- out = conv(in = dz of layer, w = its flipped
 * weights) where mask (that layer's input activation) > 0, else 0.
 * bias: the layer's parameter matrix base + KPAD (stride KPAD + 1). */
int bc250_cnn_conv(int layer, int mode, const uint16_t *in, const uint16_t *w, const float *bias, uint16_t *out,
                   const uint16_t *mask, float *pool, int ntiles, void *stream);

/* Block sums from the row-segment sums, in a fixed order. */
int bc250_cnn_pool(const float *rows, float *pool, int ntiles, void *stream);

/* Weight and bias gradient of `layer` from its input activation a and its
 * pre-activation gradient dz, summed over the tiles in a fixed order into
 * grad ([COUT][KPAD + 1], the layer's parameter layout); part is a work
 * buffer of ntiles * COUT * (KPAD + 1) floats. */
int bc250_cnn_wgrad(int layer, const uint16_t *a, const uint16_t *dz, float *part, float *grad, int ntiles,
                    void *stream);

/* Head: z [tile][16][12] from the block sums of conv4 and mvin [tile][16][32]. */
int bc250_cnn_head(const float *wh, const float *pool, const float *mvin, float *z, int ntiles, void *stream);

/* Training head: forward, loss (target [tile][16][2] in pel), backward.
 * loss [tile] (unscaled sums over its 16 blocks); dpool [tile][16][32]; the
 * head's parameter gradient times `scale` into grad ([12][65]); part is a
 * work buffer of ntiles * 12 * 65 floats. */
int bc250_cnn_head_train(const float *wh, const float *pool, const float *mvin, const float *target, float *loss,
                         float *dpool, float *part, float *grad, int ntiles, float scale, void *stream);

/* dz of conv4's pre-activation from dpool: dpool[block] where a4 > 0. */
int bc250_cnn_pool_backward(const float *dpool, const uint16_t *a4, uint16_t *dz4, int ntiles, void *stream);

/* Adam step t (>= 1), beta 0.9 / 0.999, eps 1e-8, on n parameters. */
int bc250_cnn_adam(float *p, const float *g, float *m, float *v, int n, float lr, int t, void *stream);

/* fp32 parameters -> the bf16 packed weights (BC250_CNN_NPACKED). */
int bc250_cnn_pack(const float *params, uint16_t *packed, void *stream);

/* ---- the candidate search -------------------------------------------------
 * Tiles cover the MB grid 4 x 4 MBs at a time, tile index ty * tw + tx with
 * tw = ceil(mbw / 4); block j of a tile is MB (4 tx + j % 4, 4 ty + j / 4).
 * Vectors are {dx, dy, cost, sad} records (uint32 x 4) as the full search
 * writes them (include/bc250_stages.h). */

/* The head's neighbour input mvin [tile][16][32]: for each block the
 * previous picture's integer 16x16 vectors (prev16 [nmb], NULL = all zero) at
 * MB offsets -1..2 in x and y (clamped to the picture), pel / 16. */
int bc250_cnn_mvin(const uint32_t *prev16, int mbw, int mbh, float *mvin, void *stream);

/* Candidates cand [nmb][4] {dx, dy} (whole pels) from the head outputs z:
 * the four hypotheses ordered by confidence (ties: lower index), rounded to
 * nearest, clamped to +-range. */
int bc250_cnn_candidates(const float *z, int mbw, int mbh, int range, int *cand, void *stream);

/* Integer search over the +-4 windows around the first ncand candidates of
 * every MB (positions outside +-range skipped) with the full search's
 * costs (SAD + lambda8 |mv|_1 per 8x8 block, SAD + 4 lambda8 |mv|_1 per
 * 16x16) and argmin keys (cost, then the raster index in the +-range
 * square): out16 [nmb], out8 [4 nmb] as bc250_launch_me_fullsearch writes. */
int bc250_launch_cand_search(const uint8_t *cur, int cpitch, const uint8_t *ref, int rpitch, int wc, int hc,
                             int range, int lambda8, const int *cand, int ncand, uint32_t *out16, uint32_t *out8,
                             void *stream);

/* Fallback list: list[0 .. *count) = the MBs whose 16x16 SAD (r16[m].sad)
 * exceeds alpha_q4 / 16 times the lower median of its left / right / top /
 * bottom neighbours' SADs plus beta, or exceeds t_high (in no particular
 * order).  *count is cleared on the stream first. */
int bc250_fallback_select(const uint32_t *r16, int mbw, int mbh, int alpha_q4, int beta, int t_high, int *list,
                          int *count, void *stream);

/* ---- in the pipeline ------------------------------------------------------
 * With a proposer set, the integer search of every P picture (H.264 and
 * HEVC; +-search_range as set by gpu_compute_set_encode_params) is the
 * learned one: the network over the picture's 64x64 tiles (current luma and
 * the reference reconstruction; at 10 bits their top 8 bits), the first
 * ncand candidates of every MB searched in +-4 windows, and the MBs whose
 * SAD exceeds the neighbour bound (alpha_q4, beta, t_high as in
 * bc250_fallback_select) searched in full.  Everything after the integer
 * search is unchanged. */
typedef struct {
    int32_t ncand;     /* 1..4 */
    int32_t alpha_q4;  /* fallback bound: alpha_q4 / 16 * median of the neighbours' SADs ... */
    int32_t beta;      /* ... + beta */
    int32_t t_high;    /* and every MB above this SAD */
} bc250_proposer_cfg_t;

/* params: BC250_CNN_NPARAM host floats (copied), cfg: the search; either
 * NULL turns the proposer off (the plain full search again).  Returns 0, or
 * -1 for out-of-range fields or a failed allocation. */
struct bc250_gpu_context;
int gpu_compute_set_me_proposer(struct bc250_gpu_context *ctx, const float *params, const bc250_proposer_cfg_t *cfg);

/* MBs the frame in `slot` searched in full after the learned search
 * (after gpu_compute_sync_slot), -1 when the frame did not use it. */
int gpu_compute_get_me_fallbacks(struct bc250_gpu_context *ctx, int slot);

#ifdef __cplusplus
}
#endif
#endif
