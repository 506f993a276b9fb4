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
 * Tiles of 64x64 luma.  conv1 (2->16) and conv2 (16->32) have stride 2,
 * conv3 and conv4 (32->32) stride 1, so a tile's activations are bf16
 * [tile][66][66][2] (input), [tile][34][34][16] (conv1) and
 * [tile][18][18][32] (conv2..conv4), each with a zero border around the
 * interior, which only the kernels write; a 16x16 block of the picture is a
 * 4x4 cell of conv4.  See csrc/me_cnn.hip for the layer list.  Parameters
 * are one fp32 vector of BC250_CNN_NPARAM: per conv layer l (1..4) a
 * [COUT][KPAD + 1] matrix (weights with K = (ky, kx, ci) zero-padded to
 * KPAD, the bias last), then the head [12][65] (32 cell-mean features, 32
 * neighbour-vector inputs, bias).  bf16 copies for the kernels ("packed"):
 * per layer the forward [COUT][KPAD] and, for layers 2..4, the flipped
 * [CIN][KPAD(COUT)]. */
#define BC250_CNN_NOUT 12
#define BC250_CNN_NFEAT 64
#define BC250_CNN_EPS 0.02f
#define BC250_CNN_NPARAM 24956
#define BC250_CNN_NPACKED 47104
#define BC250_CNN_POOL_SCALE 0.0625f /* cell sum -> cell mean (4 x 4 pixels of conv4) */

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
/* Bytes of one tile's activation of layer 0 (input) .. 4 (0: bad layer). */
size_t bc250_cnn_tile_bytes(int layer);

/* Input tiles from 8-bit luma pictures: tiles[3 i .. 3 i + 2] = {picture
 * slot, tile x, tile y}; slot s reads cur / ref + s * slot_stride (pitches
 * cpitch / rpitch, w x h samples, edge samples repeated past the picture). */
int bc250_cnn_input(const uint8_t *cur, int cpitch, const uint8_t *ref, int rpitch, size_t slot_stride, int w,
                    int h, const int *tiles, int ntiles, uint16_t *in, void *stream);

/* One layer.  mode 0: out = relu(conv(in, w) + b); mode 2 (layer 4 only):
 * the same, plus the bf16 outputs summed over each 4-pixel row group into
 * `pool` = rows [tile][16][4][32] fp32 (out may be NULL), which
 * bc250_cnn_pool reduces to the cell sums [tile][16][32]; mode 1 (layers
 * 2..4): input gradient of `layer` -- out = conv(in = dz of layer, w = its
 * flipped weights) where mask (that layer's input activation) > 0, else 0;
 * for layer 2 (stride 2) the convolution runs over dz zero-upsampled to the
 * input's size (dz at the odd padded positions).  bias: the layer's
 * parameter matrix base + KPAD (stride KPAD + 1). */
int bc250_cnn_conv(int layer, int mode, const uint16_t *in, const uint16_t *w, const float *bias, uint16_t *out,
                   const uint16_t *mask, float *pool, int ntiles, void *stream);

/* Cell sums from the row-group sums, in a fixed order. */
int bc250_cnn_pool(const float *rows, float *pool, int ntiles, void *stream);

/* Inference in one launch: conv1..conv4 and the cell sums [tile][16][32]
 * of the input tiles `in` ([tile][66][66][2], zero borders), one tile per
 * workgroup, every intermediate activation kept on chip; packed: the forward weights
 * (bc250_cnn_pack), params: the fp32 parameters (biases).  The same bits as
 * bc250_cnn_conv layers 1..4 (mode 2 for layer 4) followed by
 * bc250_cnn_pool. */
int bc250_cnn_forward(const uint16_t *in, const uint16_t *packed, const float *params, float *pool, int ntiles,
                      void *stream);

/* The fused launch's shape: 0 four waves, 4-row tasks, plain NHWC in LDS;
 * 1 the same with channel chunks XOR-swizzled by pixel for 32-bank reads;
 * 2 eight waves, 2-row tasks, swizzled for 64 banks.  Process-wide (0 until
 * set); every variant gives the same bits. */
int bc250_cnn_set_forward_variant(int variant);

/* Weight and bias gradient of `layer` from its input activation a and its
 * pre-activation gradient dz, summed over the tiles in a fixed order into
 * grad ([COUT][KPAD + 1], the layer's parameter layout); part is a work
 * buffer of ntiles * COUT * (KPAD + 1) floats. */
int bc250_cnn_wgrad(int layer, const uint16_t *a, const uint16_t *dz, float *part, float *grad, int ntiles,
                    void *stream);

/* Head: z [tile][16][12] from the cell sums of conv4 and mvin [tile][16][32]. */
int bc250_cnn_head(const float *wh, const float *pool, const float *mvin, float *z, int ntiles, void *stream);

/* Training head: forward, loss (target [tile][16][2] in pel), backward.
 * loss [tile] (unscaled sums over its 16 blocks); dpool [tile][16][32]; the
 * head's parameter gradient times `scale` into grad ([12][65]); part is a
 * work buffer of ntiles * 12 * 65 floats. */
int bc250_cnn_head_train(const float *wh, const float *pool, const float *mvin, const float *target, float *loss,
                         float *dpool, float *part, float *grad, int ntiles, float scale, void *stream);

/* dz of conv4's pre-activation from dpool: dpool[cell] where a4 > 0. */
int bc250_cnn_pool_backward(const float *dpool, const uint16_t *a4, uint16_t *dz4, int ntiles, void *stream);

/* Adam step t (>= 1), beta 0.9 / 0.999, eps 1e-8, on n parameters. */
int bc250_cnn_adam(float *p, const float *g, float *m, float *v, int n, float lr, int t, void *stream);

/* Workgroups per convolution launch at most (a launch shape the serving
 * autotuner chooses; default 4096, 64..65536). */
int bc250_cnn_set_grid_cap(int cap);

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

/* bc250_cnn_mvin -> bc250_cnn_head -> bc250_cnn_candidates in one launch,
 * one thread per MB, from the block sums pool [tile][16][32]: the same
 * candidates, bit for bit, without mvin and z in memory. */
int bc250_cnn_propose(const float *pool, const float *params, const uint32_t *prev16, int mbw, int mbh, int range,
                      int *cand, void *stream);

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

/* The learned search of one picture in four steps: the network and the
 * proposer from the pictures (bc250_cnn_forward_propose_picture), the
 * candidate search (which also clears *count), the fallback list (which
 * also sets the split search's keys) and the split full search of the
 * listed MBs (bc250_launch_me_fullsearch_list_split).  Work space: pool
 * [ntiles][16][32] float, cand [nmb][4] int2, list [nmb], count, keys
 * [5 nmb]; the records out16 / out8 are bc250_launch_me_fullsearch's. */
int bc250_learned_search_picture(const uint8_t *cur, int cpitch, const uint8_t *ref, int rpitch, int wc, int hc,
                                 int range, int lambda8, const float *params, const uint16_t *packed,
                                 const uint32_t *prev16, int ncand, int alpha_q4, int beta, int t_high, float *pool,
                                 int *cand, int *list, int *count, uint32_t *keys, uint32_t *out16, uint32_t *out8,
                                 void *stream);
/* The same without the records of the MBs searched in full: every MB's keys
 * (5 per MB) start empty and those MBs' keys hold their search (out16 /
 * out8 keep the candidate search's records for them); the HEVC coding makes
 * the records from the keys (bc250_hevc_launch_refine_p_encode_pack_keys). */
int bc250_learned_search_picture_keys(const uint8_t *cur, int cpitch, const uint8_t *ref, int rpitch, int wc, int hc,
                                      int range, int lambda8, const float *params, const uint16_t *packed,
                                      const uint32_t *prev16, int ncand, int alpha_q4, int beta, int t_high,
                                      float *pool, int *cand, int *list, int *count, uint32_t *keys, uint32_t *out16,
                                      uint32_t *out8, void *stream);

/* ---- batches of pictures ---------------------------------------------------
 * One launch per stage for the integer searches of several pictures (of
 * different sessions and sizes): the batch travels by value as a kernel
 * argument.  Batch-wide MB index g belongs to the picture f with
 * f[k].mb_base <= g < f[k+1].mb_base (local MB g - mb_base); tiles likewise
 * with tile_base (local tiles ceil(mbw/4) per row, as above). */
#define BC250_BATCH_MAX 32
typedef struct {
    const uint8_t *cur, *ref;     /* 8-bit luma of the picture and of its reference (coded size) */
    int32_t cpitch, rpitch, wc, hc;
    uint32_t *out16, *out8;       /* its integer search records [nmb], [4 nmb] */
    const uint32_t *prev16;       /* the previous picture's 16x16 records (NULL: none) */
    int32_t *cand;                /* [nmb][4] {dx, dy} candidates */
    int32_t mb_base, tile_base;
} bc250_batch_frame_t;
typedef struct {
    int32_t n, range, lambda8, nmb_total, ntile_total, pad;
    bc250_batch_frame_t f[BC250_BATCH_MAX];
} bc250_me_batch_t;

#if defined(__HIPCC__) || defined(__HIP__)
/* The picture of batch-wide index g (MB or tile), by binary search. */
static __device__ __forceinline__ int batch_frame_of_mb(const bc250_me_batch_t &b, int g)
{
    int lo = 0, hi = b.n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (b.f[mid].mb_base <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
static __device__ __forceinline__ int batch_frame_of_tile(const bc250_me_batch_t &b, int t)
{
    int lo = 0, hi = b.n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (b.f[mid].tile_base <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
#endif

/* The same launch reading one picture pair directly (8-bit luma, coded
 * size wc x hc, multiples of 16; the tiles in raster order, (wc/16 + 3)/4
 * per row), or every picture of a batch (tiles in batch order): the block
 * sums bc250_cnn_input / bc250_cnn_input_batch + bc250_cnn_forward give,
 * without the input tiles in memory. */
int bc250_cnn_forward_picture(const uint8_t *cur, int cpitch, const uint8_t *ref, int rpitch, int wc, int hc,
                              const uint16_t *packed, const float *params, float *pool, void *stream);
int bc250_cnn_forward_batch(const bc250_me_batch_t *b, const uint16_t *packed, const float *params, float *pool,
                            void *stream);
/* The same launches ending with the proposer: each tile's workgroup also
 * runs bc250_cnn_propose for its 16 MBs from the cell sums it holds, so cand
 * [nmb][4] {dx, dy} gets the same candidates bit for bit (prev16: the
 * previous picture's 16x16 records or NULL; range 4..64).  The batch form
 * takes prev16 and cand of each picture, and the range, from the batch. */
int bc250_cnn_forward_propose_picture(const uint8_t *cur, int cpitch, const uint8_t *ref, int rpitch, int wc, int hc,
                                      const uint16_t *packed, const float *params, float *pool,
                                      const uint32_t *prev16, int range, int *cand, void *stream);
int bc250_cnn_forward_propose_batch(const bc250_me_batch_t *b, const uint16_t *packed, const float *params,
                                    float *pool, void *stream);

/* The plain full search of every MB of the batch (list / count: only the
 * listed batch-wide MBs, count read on the device). */
int bc250_launch_me_fullsearch_batch(const bc250_me_batch_t *b, const int *list, const int *count, void *stream);
/* The listed MBs' search split over several workgroups per MB (keys: 5 *
 * nmb_total uint32 of work space, set on the stream first unless keys_ready
 * says the fallback list already set them); the same records. */
int bc250_launch_me_fullsearch_batch_split(const bc250_me_batch_t *b, const int *list, const int *count,
                                           uint32_t *keys, int keys_ready, void *stream);

/* The learned search of every picture of the batch, one launch per stage:
 * the network and the proposer from the pictures
 * (bc250_cnn_forward_propose_batch), the candidate search, the fallback list
 * (batch-wide MB indices) and the split full search of the listed MBs.
 * work: bc250_learned_work_bytes(work_tiles, work_mbs) bytes, zeroed once before its first use and laid out
 * by those capacities (any batch up to them reuses it); params / packed as
 * for the single-picture path. */
size_t bc250_learned_work_bytes(int ntiles, int nmb);
int bc250_learned_search_batch(const bc250_me_batch_t *b, const float *params, const uint16_t *packed, int ncand,
                               int alpha_q4, int beta, int t_high, void *work, int work_tiles, int work_mbs,
                               void *stream);
/* The device int holding the fallback count of the last
 * bc250_learned_search_batch on `work` (of these capacities). */
const int *bc250_learned_fallback_count(const void *work, int work_tiles, int work_mbs);
/* Input tiles of every picture of the batch (batch-wide tile order). */
int bc250_cnn_input_batch(const bc250_me_batch_t *b, uint16_t *in, void *stream);

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

/* ---- for a batching caller (bc250_serve.h) ---------------------------------
 * imv16 / imv8 (device, [nmb] / [4 nmb] records) non-NULL: P pictures skip
 * their own integer search and refine these instead (the caller wrote them,
 * e.g. with a batched search, and orders that before the dispatch with
 * gpu_compute_stream_wait_event); NULL restores the context's own search. */
int gpu_compute_set_external_int_me(struct bc250_gpu_context *ctx, const uint32_t *imv16, const uint32_t *imv8);
/* The 8-bit luma the next P picture refers to (device) and its pitch. */
int gpu_compute_get_reference_luma(struct bc250_gpu_context *ctx, const uint8_t **y, int *pitch);
/* The context's compute stream waits for a hipEvent_t (device-side wait). */
int gpu_compute_stream_wait_event(struct bc250_gpu_context *ctx, void *event);
/* 1 when the frame submitted in `slot` has finished (stages and copies), 0
 * while it runs, -1 for an unused slot; never blocks. */
int gpu_compute_slot_done(struct bc250_gpu_context *ctx, int slot);

#ifdef __cplusplus
}
#endif
#endif
