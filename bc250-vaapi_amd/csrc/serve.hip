// serve.hip - many encode sessions on one GPU (include/bc250_serve.h):
// admission by device-memory footprint, per-session contexts whose
// reconstructions stay resident, and the request loop with continuous
// batching of the integer motion search.  Host code only; the kernels are
// the pipeline's own (gpu_compute.hip) and the batched searches
// (me_kernels.hip, me_learned.hip).
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <deque>
#include <vector>

#include "bc250_hevc.h"
#include "bc250_learned.h"
#include "bc250_serve.h"
#include "bc250_stages.h"
#include "gpu_compute.h"

#define HIP_OK(x) ((x) == hipSuccess)

namespace {

double now_s()
{
    timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + t.tv_nsec * 1e-9;
}

size_t al256(size_t v) { return (v + 255) & ~(size_t)255; }

struct Session {
    bool open = false;
    bc250_session_cfg_t cfg{};
    gpu_context_t ctx{};
    std::vector<gpu_image_t> img;
    std::vector<gpu_memory_t> mem;
    uint32_t *imv16[2] = {nullptr, nullptr};  // ping-pong: this picture's / the previous picture's records
    uint32_t *imv8 = nullptr;
    int32_t *cand = nullptr;
    int pb = 0;
    bool has_prev = false;      // the previous picture ran the batched search
    int wc = 0, hc = 0, nmb = 0, ntiles = 0;
    size_t bytes = 0;
    // request loop state
    std::deque<int> fifo;       // released requests, in order
    bool busy = false;
    int req = -1, slot = -1, batch = -1;
    int frames = 0;
};

}  // namespace

struct bc250_server {
    bc250_server_cfg_t cfg{};
    size_t budget = 0, used = 0;
    std::vector<Session> ses;
    hipStream_t stream = nullptr;   // batched searches
    hipEvent_t me_done = nullptr;
    float *params = nullptr;
    uint16_t *packed = nullptr;
    void *work = nullptr;           // learned-search work for every open session at once
    size_t work_bytes = 0;
    int work_tiles = 0, work_mbs = 0;
    static constexpr int FB_RING = 4096;
    int *fb_host = nullptr;         // pinned fallback counts of the last FB_RING batches, by batch number
    // statistics
    int batches = 0, frames = 0, max_batch_seen = 0;
    int forward_variant = -1;
    double tune_us[8] = {0};
};

static int session_dims(const bc250_session_cfg_t *c, int *wc, int *hc)
{
    if (!c || c->width <= 0 || c->height <= 0 || c->width > 8192 || c->height > 8192 || c->qp < 0 || c->qp > 51 ||
        c->num_slices < 1 || c->gop < 0 || c->ring < 1 || c->ring > 64 ||
        (c->codec != GPU_CODEC_H264 && c->codec != GPU_CODEC_HEVC))
        return -1;
    *wc = (c->width + 15) / 16 * 16;
    *hc = (c->height + 15) / 16 * 16;
    return 0;
}

extern "C" size_t bc250_server_session_bytes(const bc250_server_t *srv, const bc250_session_cfg_t *c)
{
    int wc, hc;
    if (!srv || session_dims(c, &wc, &hc) != 0) return 0;
    const size_t nmb = (size_t)(wc / 16) * (hc / 16);
    const size_t pitch = al256((size_t)wc);
    const size_t picture = pitch * hc + pitch * (hc / 2);
    size_t b = 0;
    b += 2 * picture;                                                   // reconstructions
    b += 2 * nmb * (768 + 96 + 96 + 4 + 16 + 64 + 16);                  // both slots' outputs
    b += nmb * (16 + 64 + 4 + 4);                                       // search records, CU flags / sizes
    b += bc250_intra_work_bytes(wc, hc) + bc250_deblock_work_bytes(wc, hc) + bc250_hevc_intra_work_bytes(wc, hc);
    b += (size_t)c->ring * picture;                                     // source pictures
    b += nmb * (2 * 16 + 64 + 32);                                      // the server's records and candidates
    if (srv->cfg.learned) {
        const int nt = ((wc / 16 + 3) / 4) * ((hc / 16 + 3) / 4);
        b += bc250_learned_work_bytes(nt, (int)nmb);
    }
    return b;
}

static void free_session(Session &s)
{
    if (s.ctx.hip) {
        for (size_t i = 0; i < s.img.size(); i++)
            if (s.mem[i].memory) gpu_compute_destroy_image(&s.ctx, s.img[i], s.mem[i]);
        bc250_gpu_destroy(&s.ctx);
    }
    for (void *p : {(void *)s.imv16[0], (void *)s.imv16[1], (void *)s.imv8, (void *)s.cand})
        if (p) hipFree(p);
    s = Session{};
}

extern "C" bc250_server_t *bc250_server_create(const bc250_server_cfg_t *cfg, const float *proposer_params)
{
    if (!cfg || cfg->max_batch < 1 || cfg->max_batch > BC250_BATCH_MAX ||
        (cfg->search_range != 16 && cfg->search_range != 32 && cfg->search_range != 64) || cfg->lambda8 < 0 ||
        cfg->lambda8 > 15 || (cfg->learned && !proposer_params))
        return nullptr;
    if (cfg->learned && (cfg->proposer.ncand < 1 || cfg->proposer.ncand > 4)) return nullptr;
    bc250_server_t *srv = new bc250_server();
    srv->cfg = *cfg;
    size_t fr = 0, tot = 0;
    if (!HIP_OK(hipMemGetInfo(&fr, &tot))) {
        delete srv;
        return nullptr;
    }
    srv->budget = cfg->hbm_budget ? cfg->hbm_budget : fr / 4 * 3;
    bool ok = HIP_OK(hipStreamCreateWithFlags(&srv->stream, hipStreamNonBlocking)) &&
              HIP_OK(hipEventCreateWithFlags(&srv->me_done, hipEventDisableTiming)) &&
              HIP_OK(hipHostMalloc((void **)&srv->fb_host, bc250_server::FB_RING * sizeof(int), hipHostMallocDefault));
    if (ok && cfg->learned) {
        ok = HIP_OK(hipMalloc(&srv->params, BC250_CNN_NPARAM * sizeof(float))) &&
             HIP_OK(hipMalloc(&srv->packed, BC250_CNN_NPACKED * sizeof(uint16_t))) &&
             HIP_OK(hipMemcpy(srv->params, proposer_params, BC250_CNN_NPARAM * sizeof(float),
                              hipMemcpyHostToDevice)) &&
             bc250_cnn_pack(srv->params, srv->packed, srv->stream) == 0 && HIP_OK(hipStreamSynchronize(srv->stream));
    }
    if (!ok) {
        bc250_server_destroy(srv);
        return nullptr;
    }
    return srv;
}

extern "C" void bc250_server_destroy(bc250_server_t *srv)
{
    if (!srv) return;
    if (srv->stream) hipStreamSynchronize(srv->stream);
    for (auto &s : srv->ses)
        if (s.open) free_session(s);
    for (void *p : {(void *)srv->params, (void *)srv->packed, srv->work})
        if (p) hipFree(p);
    if (srv->fb_host) hipHostFree(srv->fb_host);
    if (srv->me_done) hipEventDestroy(srv->me_done);
    if (srv->stream) hipStreamDestroy(srv->stream);
    delete srv;
}

// The learned search's work buffer sized for every open session at once,
// zeroed once (the activations' borders stay zero).
static int ensure_work(bc250_server_t *srv)
{
    if (!srv->cfg.learned) return 0;
    int nt = 0, nmb = 0;
    for (auto &s : srv->ses)
        if (s.open) {
            nt += s.ntiles;
            nmb += s.nmb;
        }
    if (nt <= srv->work_tiles && nmb <= srv->work_mbs) return 0;
    if (srv->work) {
        hipStreamSynchronize(srv->stream);
        hipFree(srv->work);
        srv->work = nullptr;
    }
    const size_t b = bc250_learned_work_bytes(nt, nmb);
    if (!HIP_OK(hipMalloc(&srv->work, b)) || !HIP_OK(hipMemset(srv->work, 0, b))) {
        srv->work = nullptr;
        srv->work_tiles = srv->work_mbs = 0;
        return -1;
    }
    srv->work_bytes = b;
    srv->work_tiles = nt;
    srv->work_mbs = nmb;
    return 0;
}

extern "C" int bc250_server_open(bc250_server_t *srv, const bc250_session_cfg_t *c)
{
    int wc, hc;
    if (!srv || session_dims(c, &wc, &hc) != 0) return -1;
    const size_t bytes = bc250_server_session_bytes(srv, c);
    if (srv->used + bytes > srv->budget) return -1;  // does not fit (yet)
    int id = -1;
    for (size_t i = 0; i < srv->ses.size(); i++)
        if (!srv->ses[i].open) { id = (int)i; break; }
    if (id < 0) {
        id = (int)srv->ses.size();
        srv->ses.emplace_back();
    }
    Session &s = srv->ses[id];
    s = Session{};
    s.cfg = *c;
    s.wc = wc;
    s.hc = hc;
    s.nmb = (wc / 16) * (hc / 16);
    s.ntiles = ((wc / 16 + 3) / 4) * ((hc / 16 + 3) / 4);
    s.bytes = bytes;
    bool ok = bc250_gpu_init(&s.ctx) == 0;
    gpu_encode_params_t p{srv->cfg.search_range, srv->cfg.lambda8, 0, 1, 1, c->codec, 0, 0, 4, 0};
    ok = ok && gpu_compute_set_encode_params(&s.ctx, &p) == 0;
    s.img.resize(c->ring);
    s.mem.resize(c->ring);
    for (int i = 0; ok && i < c->ring; i++)
        ok = gpu_compute_create_image(&s.ctx, c->width, c->height, GPU_IMAGE_NV12, &s.img[i], &s.mem[i]) == 0;
    ok = ok && HIP_OK(hipMalloc(&s.imv16[0], (size_t)s.nmb * 16)) && HIP_OK(hipMalloc(&s.imv16[1], (size_t)s.nmb * 16)) &&
         HIP_OK(hipMalloc(&s.imv8, (size_t)s.nmb * 64)) && HIP_OK(hipMalloc(&s.cand, (size_t)s.nmb * 32));
    // the context's buffers now, not at the first frame: encode a flat
    // picture once (it is an I picture and leaves nothing behind but the
    // allocation: the first request is an I picture again)
    if (ok) {
        std::vector<uint8_t> y((size_t)c->width * c->height, 128), uv((size_t)((c->width + 1) / 2) * 2 * ((c->height + 1) / 2), 128);
        ok = gpu_compute_upload_nv12(&s.ctx, &s.img[0], s.mem[0], y.data(), c->width, uv.data(),
                                     (c->width + 1) / 2 * 2, c->width, c->height) == 0 &&
             gpu_compute_begin_picture(&s.ctx, s.img[0]) == 0 &&
             gpu_compute_dispatch_encode_ext(&s.ctx, s.img[0], c->width, c->height, c->qp, 1, c->num_slices, 0,
                                             nullptr) == 0 &&
             gpu_compute_end_picture(&s.ctx) == 0 &&
             gpu_compute_sync_slot(&s.ctx, gpu_compute_submitted_slot(&s.ctx)) == 0;
    }
    s.open = true;
    ok = ok && ensure_work(srv) == 0;
    if (!ok) {
        fprintf(stderr, "[bc250-serve] session %dx%d could not be set up\n", c->width, c->height);
        free_session(s);
        return -1;
    }
    srv->used += bytes;
    return id;
}

extern "C" int bc250_server_close(bc250_server_t *srv, int id)
{
    if (!srv || id < 0 || id >= (int)srv->ses.size() || !srv->ses[id].open) return -1;
    Session &s = srv->ses[id];
    if (s.busy) gpu_compute_sync_slot(&s.ctx, s.slot);
    srv->used -= s.bytes;
    free_session(s);
    return 0;
}

extern "C" int bc250_server_upload(bc250_server_t *srv, int id, int source, const uint8_t *y, int y_pitch,
                                   const uint8_t *uv, int uv_pitch)
{
    if (!srv || id < 0 || id >= (int)srv->ses.size() || !srv->ses[id].open) return -1;
    Session &s = srv->ses[id];
    if (source < 0 || source >= s.cfg.ring || s.busy) return -1;
    return gpu_compute_upload_nv12(&s.ctx, &s.img[source], s.mem[source], y, y_pitch, uv, uv_pitch, s.cfg.width,
                                   s.cfg.height);
}

static uint64_t fnv(uint64_t h, const void *p, size_t n)
{
    const uint8_t *b = (const uint8_t *)p;
    for (size_t i = 0; i < n; i++) h = (h ^ b[i]) * 0x100000001b3ull;
    return h;
}

extern "C" uint64_t bc250_fnv1a64(uint64_t h, const void *p, size_t n) { return fnv(h, p, n); }

static uint64_t staging_checksum(Session &s, int slot)
{
    uint64_t h = 0xcbf29ce484222325ull;
    const bool hevc = s.cfg.codec == GPU_CODEC_HEVC;
    void *d = nullptr;
    size_t n = 0;
    if (gpu_compute_get_quant_staging_data_slot(&s.ctx, slot, &d, &n) == 0) h = fnv(h, d, n);
    if (gpu_compute_get_mv8_staging_data_slot(&s.ctx, slot, &d, &n) == 0) h = fnv(h, d, n);
    if (hevc) {
        if (gpu_compute_get_hevc_mode_staging_data_slot(&s.ctx, slot, &d, &n) == 0) h = fnv(h, d, n);
    } else {
        if (gpu_compute_get_dc_staging_data_slot(&s.ctx, slot, &d, &n) == 0) h = fnv(h, d, n);
        if (gpu_compute_get_pred_mode_staging_data_slot(&s.ctx, slot, &d, &n) == 0) h = fnv(h, d, n);
        if (gpu_compute_get_mv_staging_data_slot(&s.ctx, slot, &d, &n) == 0) h = fnv(h, d, n);
    }
    return h;
}

// One batched integer search for the P pictures of `members`, then every
// member's dispatch.  Returns 0 or -1.
static int launch_batch(bc250_server_t *srv, const std::vector<int> &members, const bc250_request_t *reqs,
                        bc250_frame_result_t *out, double t0)
{
    bc250_me_batch_t b;
    memset(&b, 0, sizeof b);
    b.range = srv->cfg.search_range;
    b.lambda8 = srv->cfg.lambda8;
    std::vector<char> intra(members.size());
    for (size_t k = 0; k < members.size(); k++) {
        Session &s = srv->ses[members[k]];
        const bc250_request_t &r = reqs[s.req];
        intra[k] = s.frames == 0 || (s.cfg.gop > 0 && s.frames % s.cfg.gop == 0);
        if (intra[k]) continue;
        bc250_batch_frame_t &f = b.f[b.n++];
        const uint8_t *ry = nullptr;
        int rp = 0;
        if (gpu_compute_get_reference_luma(&s.ctx, &ry, &rp) != 0) return -1;
        f.cur = (const uint8_t *)s.img[r.source].y_plane;
        f.cpitch = (int)s.img[r.source].y_pitch;
        f.ref = ry;
        f.rpitch = rp;
        f.wc = s.wc;
        f.hc = s.hc;
        f.out16 = s.imv16[s.pb];
        f.out8 = s.imv8;
        f.prev16 = s.has_prev ? s.imv16[s.pb ^ 1] : nullptr;
        f.cand = s.cand;
        f.mb_base = b.nmb_total;
        f.tile_base = b.ntile_total;
        b.nmb_total += s.nmb;
        b.ntile_total += s.ntiles;
    }
    int fb = -1;
    if (b.n > 0) {
        int rc;
        if (srv->cfg.learned) {
            const bc250_proposer_cfg_t &q = srv->cfg.proposer;
            rc = bc250_learned_search_batch(&b, srv->params, srv->packed, q.ncand, q.alpha_q4, q.beta, q.t_high,
                                            srv->work, srv->work_tiles, srv->work_mbs, srv->stream);
            if (rc == 0)
                rc = HIP_OK(hipMemcpyAsync(srv->fb_host + srv->batches % bc250_server::FB_RING,
                                           bc250_learned_fallback_count(srv->work, srv->work_tiles, srv->work_mbs),
                                           sizeof(int), hipMemcpyDeviceToHost, srv->stream)) ? 0 : -1;
            fb = 0;
        } else {
            rc = bc250_launch_me_fullsearch_batch(&b, nullptr, nullptr, srv->stream);
        }
        if (rc != 0 || !HIP_OK(hipEventRecord(srv->me_done, srv->stream))) return -1;
    }
    const double ts = now_s() - t0;
    for (size_t k = 0; k < members.size(); k++) {
        Session &s = srv->ses[members[k]];
        const bc250_request_t &r = reqs[s.req];
        if (!intra[k]) {
            if (gpu_compute_set_external_int_me(&s.ctx, s.imv16[s.pb], s.imv8) != 0 ||
                gpu_compute_stream_wait_event(&s.ctx, srv->me_done) != 0)
                return -1;
        }
        if (gpu_compute_begin_picture(&s.ctx, s.img[r.source]) != 0 ||
            gpu_compute_dispatch_encode_ext(&s.ctx, s.img[r.source], s.cfg.width, s.cfg.height, s.cfg.qp, intra[k],
                                            s.cfg.num_slices, 0, nullptr) != 0 ||
            gpu_compute_end_picture(&s.ctx) != 0)
            return -1;
        s.slot = gpu_compute_submitted_slot(&s.ctx);
        s.busy = true;
        s.batch = srv->batches;
        out[s.req].t_start = ts;
        out[s.req].batch = srv->batches;
        out[s.req].fallbacks = intra[k] ? -1 : fb;
        if (!intra[k]) {
            s.pb ^= 1;
            s.has_prev = true;
        } else {
            s.has_prev = false;
        }
        s.frames++;
    }
    srv->batches++;
    srv->frames += (int)members.size();
    srv->max_batch_seen = std::max(srv->max_batch_seen, (int)members.size());
    return 0;
}

extern "C" int bc250_server_run(bc250_server_t *srv, const bc250_request_t *reqs, int n, int checksums,
                                bc250_frame_result_t *out, double *wall_s)
{
    if (!srv || !reqs || !out || n <= 0) return -1;
    for (int i = 0; i < n; i++) {
        const int id = reqs[i].session;
        if (id < 0 || id >= (int)srv->ses.size() || !srv->ses[id].open || reqs[i].source < 0 ||
            reqs[i].source >= srv->ses[id].cfg.ring || (i > 0 && reqs[i].t_arrival < reqs[i - 1].t_arrival))
            return -1;
        out[i] = bc250_frame_result_t{};
        out[i].session = id;
        out[i].frame = reqs[i].frame;
        out[i].t_arrival = reqs[i].t_arrival;
    }
    for (auto &s : srv->ses) {
        s.fifo.clear();
        s.busy = false;
        s.frames = 0;
        s.has_prev = false;
    }
    const double t0 = now_s();
    int next = 0, done = 0, rr = 0, rc = 0;
    const int ns = (int)srv->ses.size();
    std::vector<int> members;
    while (done < n && rc == 0) {
        const double t = now_s() - t0;
        while (next < n && reqs[next].t_arrival <= t) {
            srv->ses[reqs[next].session].fifo.push_back(next);
            next++;
        }
        // retire finished frames
        for (auto &s : srv->ses) {
            if (!s.open || !s.busy) continue;
            const int d = gpu_compute_slot_done(&s.ctx, s.slot);
            if (d < 0) { rc = -1; break; }
            if (d == 0) continue;
            if (gpu_compute_sync_slot(&s.ctx, s.slot) != 0) { rc = -1; break; }
            bc250_frame_result_t &o = out[s.req];
            o.t_done = now_s() - t0;
            // the batch's count (its copy came before the event the frame's stages waited for)
            if (o.fallbacks >= 0) o.fallbacks = srv->fb_host[o.batch % bc250_server::FB_RING];
            if (checksums) o.checksum = staging_checksum(s, s.slot);
            s.busy = false;
            s.req = -1;
            done++;
        }
        // continuous batching: every session whose next frame has arrived and
        // whose previous frame is done joins this step's batch
        members.clear();
        for (int k = 0; k < ns && (int)members.size() < srv->cfg.max_batch; k++) {
            Session &s = srv->ses[(rr + k) % ns];
            if (s.open && !s.busy && !s.fifo.empty()) {
                s.req = s.fifo.front();
                s.fifo.pop_front();
                members.push_back((rr + k) % ns);
            }
        }
        if (!members.empty()) {
            rr = (members.back() + 1) % ns;
            if (launch_batch(srv, members, reqs, out, t0) != 0) rc = -1;
        } else if (done < n) {
            __builtin_ia32_pause();
        }
    }
    if (wall_s) *wall_s = now_s() - t0;
    for (auto &s : srv->ses)
        if (s.open) gpu_compute_set_external_int_me(&s.ctx, nullptr, nullptr);
    return rc;
}

extern "C" int bc250_server_autotune(bc250_server_t *srv)
{
    if (!srv) return -1;
    if (!srv->cfg.learned) return 0;
    // a batch of every open session, each picture searched against itself
    bc250_me_batch_t b;
    memset(&b, 0, sizeof b);
    b.range = srv->cfg.search_range;
    b.lambda8 = srv->cfg.lambda8;
    for (auto &s : srv->ses) {
        if (!s.open || b.n == BC250_BATCH_MAX) continue;
        bc250_batch_frame_t &f = b.f[b.n++];
        f.cur = f.ref = (const uint8_t *)s.img[0].y_plane;
        f.cpitch = f.rpitch = (int)s.img[0].y_pitch;
        f.wc = s.wc;
        f.hc = s.hc;
        f.out16 = s.imv16[0];
        f.out8 = s.imv8;
        f.cand = s.cand;
        f.mb_base = b.nmb_total;
        f.tile_base = b.ntile_total;
        b.nmb_total += s.nmb;
        b.ntile_total += s.ntiles;
    }
    if (b.n == 0) return -1;
    hipEvent_t e0, e1;
    if (!HIP_OK(hipEventCreate(&e0)) || !HIP_OK(hipEventCreate(&e1))) return -1;
    const bc250_proposer_cfg_t &q = srv->cfg.proposer;
    double best = 1e30;
    int best_v = 1, rc = 0;
    for (int i = 0; i < 3 && rc == 0; i++) {
        bc250_cnn_set_forward_variant(i);
        float ms = 0.f, tot = 0.f;
        for (int rep = 0; rep < 4 && rc == 0; rep++) {
            hipEventRecord(e0, srv->stream);
            rc = bc250_learned_search_batch(&b, srv->params, srv->packed, q.ncand, q.alpha_q4, q.beta, q.t_high,
                                            srv->work, srv->work_tiles, srv->work_mbs, srv->stream);
            hipEventRecord(e1, srv->stream);
            hipEventSynchronize(e1);
            hipEventElapsedTime(&ms, e0, e1);
            if (rep > 0) tot += ms;  // the first is a warm-up
        }
        srv->tune_us[i] = tot / 3 * 1e3;
        if (srv->tune_us[i] < best) {
            best = srv->tune_us[i];
            best_v = i;
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    bc250_cnn_set_forward_variant(best_v);
    srv->forward_variant = best_v;
    return rc;
}

extern "C" int bc250_server_stats(const bc250_server_t *srv, bc250_server_stats_t *st)
{
    if (!srv || !st) return -1;
    memset(st, 0, sizeof *st);
    st->batches = srv->batches;
    st->frames = srv->frames;
    st->max_batch_seen = srv->max_batch_seen;
    st->mean_batch = srv->batches ? (double)srv->frames / srv->batches : 0.0;
    for (auto &s : srv->ses) st->sessions += s.open ? 1 : 0;
    st->forward_variant = srv->forward_variant;
    for (int i = 0; i < 8; i++) st->autotune_us[i] = srv->tune_us[i];
    st->hbm_used = srv->used;
    return 0;
}
