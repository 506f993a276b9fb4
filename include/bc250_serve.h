/*
 * bc250_serve.h - many encode sessions on one GPU: admission by device
 * memory footprint, reference frames resident per session, and a request
 * loop with continuous batching.
 *
 * No counterpart in the reference: its driver gives each VA context its own
 * encoder but all of them share one GPU context, serialised by a lock
 * (va_backend.h:51-53; reallocation on every resolution change,
 * gpu_compute.c:2931-2935).  Here every session owns a gpu_compute context
 * (its buffers allocated once, its reconstructions kept in HBM), and the
 * server batches the integer motion search of all sessions whose next frame
 * is ready into one launch per stage (csrc/me_kernels.hip,
 * csrc/me_learned.hip), then runs each session's remaining stages on its
 * own streams, where they overlap.
 */
#ifndef BC250_SERVE_H
#define BC250_SERVE_H

#include <stddef.h>
#include <stdint.h>

#include "bc250_learned.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bc250_server bc250_server_t;

typedef struct {
    size_t hbm_budget;       /* device bytes the sessions may hold (0: 3/4 of the free memory at creation) */
    int32_t max_batch;       /* pictures per batched integer search, 1..BC250_BATCH_MAX */
    int32_t learned;         /* 1: the learned integer search (needs proposer params), 0: the plain full search */
    int32_t search_range;    /* 16, 32 or 64 */
    int32_t lambda8;         /* 0..15 */
    bc250_proposer_cfg_t proposer;
} bc250_server_cfg_t;

typedef struct {
    int32_t width, height;   /* 8-bit NV12 pictures */
    int32_t codec;           /* GPU_CODEC_H264 or GPU_CODEC_HEVC */
    int32_t qp;              /* CQP */
    int32_t num_slices;
    int32_t gop;             /* I-picture period (0: only the first) */
    int32_t ring;            /* source pictures the session keeps on the device (>= 1) */
} bc250_session_cfg_t;

/* One frame to encode: `source` names one of the session's uploaded source
 * pictures; t_arrival is in seconds on the server's clock (0 = the start of
 * bc250_server_run).  A session's requests are coded in the order given. */
typedef struct {
    int32_t session, frame, source, pad;
    double t_arrival;
} bc250_request_t;

typedef struct {
    int32_t session, frame;
    int32_t batch;           /* the batch it was launched in */
    int32_t fallbacks;       /* MBs its batch's learned search searched in full (-1: not in that search) */
    double t_arrival, t_start, t_done;  /* seconds on the server's clock */
    uint64_t checksum;       /* FNV-1a of the frame's staging (levels, vectors, modes) when asked for */
} bc250_frame_result_t;

typedef struct {
    int32_t batches, frames, max_batch_seen, sessions;
    double mean_batch;
    int32_t forward_variant; /* the autotuner's choice of the network's LDS layout (-1: not tuned) */
    double autotune_us[8];   /* its measurements, per candidate in bc250_server_autotune's order */
    size_t hbm_used;
} bc250_server_stats_t;

/* proposer_params: BC250_CNN_NPARAM floats (copied) when cfg->learned. */
bc250_server_t *bc250_server_create(const bc250_server_cfg_t *cfg, const float *proposer_params);
void bc250_server_destroy(bc250_server_t *srv);

/* Device bytes a session of this shape holds: its context's buffers (both
 * slots' outputs, two reconstructions, search and wavefront workspaces), its
 * source pictures, its integer-search records and, with the learned search,
 * its share of the network's activations. */
size_t bc250_server_session_bytes(const bc250_server_t *srv, const bc250_session_cfg_t *cfg);

/* Opens a session when its footprint fits the budget: returns its id, or -1
 * when it does not fit (yet: closing a session frees its share) or the
 * shape is invalid.  All its device memory is allocated here. */
int bc250_server_open(bc250_server_t *srv, const bc250_session_cfg_t *cfg);
int bc250_server_close(bc250_server_t *srv, int session);

/* Source picture `source` (0..ring-1) of a session, from host NV12 planes. */
int bc250_server_upload(bc250_server_t *srv, int session, int source, const uint8_t *y, int y_pitch,
                        const uint8_t *uv, int uv_pitch);

/* The request loop.  Requests are released when the server's clock passes
 * their arrival time; a frame starts when it has arrived and its session's
 * previous frame is done; each step gathers every startable frame (up to
 * max_batch, round robin over the sessions) into one batched integer search
 * and then dispatches each session's remaining stages on its own streams,
 * while earlier batches still run.  out[i] is request i's result; checksums
 * 1 hashes every frame's staging on the host (slow: for tests).  Returns 0,
 * or -1 when a frame failed. */
int bc250_server_run(bc250_server_t *srv, const bc250_request_t *reqs, int n, int checksums,
                     bc250_frame_result_t *out, double *wall_s);

/* Times the learned search on a batch of every open session with each
 * LDS layout of the fused network launch (bc250_cnn_set_forward_variant)
 * and keeps the fastest (a no-op for the plain search). */
int bc250_server_autotune(bc250_server_t *srv);

int bc250_server_stats(const bc250_server_t *srv, bc250_server_stats_t *st);

/* FNV-1a 64 of n bytes continuing from h (0xcbf29ce484222325 to start): the
 * frame checksum the request loop records. */
uint64_t bc250_fnv1a64(uint64_t h, const void *p, size_t n);

#ifdef __cplusplus
}
#endif
#endif
