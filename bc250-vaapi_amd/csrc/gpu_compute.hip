// gpu_compute.hip - HIP implementation of the gpu_compute_* C ABI
// (include/gpu_compute.h), the drop-in for the reference's Vulkan
// orchestration (approach1-compute-encoder/src/gpu_compute.c).
//
// One context = a compute stream and a copy stream (both default priority)
// + per-resolution device buffers + two slots.  A slot owns the device
// buffers its frame's outputs are written to and the pinned host staging
// they are copied into.  dispatch_encode_ext enqueues the stages of a frame
// on the compute stream and the staging copies on the copy stream (after
// the frame's stages), so the copies of frame N overlap the stages of frame
// N+1; end_picture "submits" the slot (records its completion event on the
// copy stream) and toggles the slot, as the reference's vkQueueSubmit + slot
// flip (gpu_compute.c:3523-3619); sync_slot waits on that event
// (vkWaitForFences, gpu_compute.c:3640-3712).
#include <hip/hip_runtime.h>

#include <errno.h>
#include <fcntl.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include "bc250_bitstream.h"
#include "bc250_dev.h"
#include "bc250_hevc.h"
#include "bc250_learned.h"
#include "bc250_stages.h"
#include "gpu_compute.h"

#define HIP_OK(x) ((x) == hipSuccess)

namespace {

constexpr int NSLOT = 2;
constexpr int NEV = 6;  // start, after integer ME, after refine, after encode, after deblock, after copies
constexpr int NSTAGE = NEV - 1;
constexpr unsigned kSpinLimit = 1u << 22;

struct Plane2 {
    uint8_t *y = nullptr, *uv = nullptr;
    int py = 0, puv = 0;
    void *base = nullptr;
};

// Device buffers of one slot's frame outputs (what the host reads back).
struct DevOut {
    int16_t *levels = nullptr;
    int32_t *dc = nullptr;
    uint32_t *nz = nullptr, *modes = nullptr, *mvs = nullptr, *mv8 = nullptr;
    uint8_t *hmodes = nullptr;
    // the levels of a P frame packed for the readback: every 16-level group
    // holding a nonzero level (lvdata, 32 bytes each) with its group index
    // (lvidx), in no particular order, and their count
    uint32_t *lvidx = nullptr;
    uint4 *lvdata = nullptr;
    int *lvcount = nullptr;
    unsigned *rbdone = nullptr;  // the readback kernel's finished blocks (0 between launches)
};

struct Slot {
    int16_t *levels = nullptr;
    int32_t *dc = nullptr;
    uint32_t *modes = nullptr, *nz = nullptr;
    gpu_mv_t *mvs = nullptr, *mv8 = nullptr;
    uint8_t *hmodes = nullptr;  // HEVC PU modes
    int *err = nullptr;  // 4 ints: intra wait error, deblock wait error, learned-search fallback MBs (-1: none)
    uint8_t *rbsp = nullptr;     // CAVLC slices of the frame (entropy = 1), pinned
    uint32_t *rbsp_info = nullptr;
    int slices = 0;              // their number; 0 = no slices for this frame
    hipEvent_t ev[NEV] = {};
    hipEvent_t cdone = nullptr;  // the frame's stages done (compute stream): its copies may start
    hipEvent_t cearly = nullptr; // the frame's outputs final before its last stages (see finish_frame)
    DevOut dev;
    bool pending = false, recorded = false, have_data = false, was_intra = false;
    bool mv8_half = false;     // mv8 arrived as {x | y << 16, sad} per CU8 in its first half
    bool hmodes_zero = false;  // the frame's PU modes are all zero and did not travel
    bool timed = false;          // the stage events ev[0 .. NEV-2] were recorded for this frame
    // packed levels in host memory (pinned): the copies of DevOut's, every
    // packed group; `levels` is rebuilt from them when read
    uint32_t *lvidx = nullptr;
    int16_t *lvdata = nullptr;
    int *lvcount = nullptr;
    bool lv_packed = false;        // this frame's levels arrived packed
    bool lv_settled = false;       // ... their count was checked
    bool lv_built = false;         // ... and `levels` holds them again
    uint32_t *built_idx = nullptr; // groups `levels` holds nonzero (host heap)
    int built_n = 0;
    bool built_all = true;         // `levels` may hold nonzero levels anywhere
};

// The learned proposer's weights and per-resolution buffers
// (gpu_compute_set_me_proposer).
struct Proposer {
    bool on = false;
    bc250_proposer_cfg_t cfg{};
    float *params = nullptr;     // fp32 [BC250_CNN_NPARAM]
    uint16_t *packed = nullptr;  // bf16 [BC250_CNN_NPACKED]
    float *pool = nullptr;       // block sums [tile][16][32] of the picture's 64x64 tiles
    int *cand = nullptr, *list = nullptr, *count = nullptr;
    uint32_t *keys = nullptr;    // the split fallback search's keys
    bool has_prev = false;       // the integer records hold the picture before the current one
    bool ran = false;            // the current picture ran the learned search
};

}  // namespace

struct bc250_hip_state {
    hipStream_t stream = nullptr;   // stages
    hipStream_t cstream = nullptr;  // staging copies
    void *xfer = nullptr;           // pinned bounce buffer of uploads and downloads
    size_t xfer_size = 0;
    // resolution-dependent state
    int w = 0, h = 0, wc = 0, hc = 0, mbw = 0, mbh = 0, nmb = 0;
    int bps = 1;           // bytes per sample of the reconstructions: 1 (NV12) or 2 (P010)
    Plane2 recon[2];
    uint8_t *me8 = nullptr;  // P010 pictures: 8-bit luma of the current and reference frames for the search
    int me8_pitch = 0;
    int ref_idx = 0;
    bool has_recon = false;
    // outputs of the frame being dispatched: the current slot's DevOut (bind_slot)
    int16_t *levels = nullptr;
    int32_t *dc = nullptr;
    uint32_t *nz = nullptr, *modes = nullptr;
    uint32_t *mvs = nullptr, *mv8 = nullptr;
    uint8_t *hmodes = nullptr;  // HEVC PU modes
    uint32_t *imv16 = nullptr, *imv8 = nullptr;
    int *sync_i = nullptr, *sync_d = nullptr, *sync_h = nullptr;
    uint8_t *hcbf = nullptr;     // per-CU coded-luma flags
    uint8_t *hculog2 = nullptr;                   // HEVC P: log2 CU size per CU8 (CTU32 CU choice)
    Slot slot[NSLOT];
    int cur = 0, submitted = 0;
    double last_latency_ms = 0.0;
    // stage events on the compute stream (gpu_compute_set_stage_timing; on
    // when BC250_PERF_STATS is set, as the reference's timestamp queries)
    bool timing = false;
    gpu_encode_params_t params{64, 1, 0, 1, 1, GPU_CODEC_H264, 0, 0, 4, 0};
    // CAVLC slices on the GPU (entropy = 1), allocated at the first such frame
    uint8_t *cavlc_out = nullptr;
    uint32_t *cavlc_info = nullptr;
    void *cavlc_work = nullptr;
    size_t cavlc_cap = 0, cavlc_work_size = 0;
    bc250_h264_slice_cfg_t hdr{};  // static header fields (gpu_compute_set_h264_slice_header)
    // the next frame's frame_num and POC (0 at an IDR, +1 / +2 per coded frame, written
    // modulo 2^log2_max_*; encoder_h264.c:2258-2260,3526-3527) and the last idr_pic_id
    unsigned frame_num = 0, poc = 0;
    int idr_pic_id = -1;
    int err_flags = 0;
    // per-stage GPU time accumulated at sync_slot: [stage] sums, frames seen
    double stage_sum[NSTAGE] = {0, 0, 0, 0, 0};
    double total_sum = 0;
    int stage_frames = 0, stage_p_frames = 0;
    double me_int_sum_p = 0;
    Proposer prop;
    const uint32_t *fs_keys = nullptr;  // this picture's learned search left its full searches as keys
    // integer search results written by someone else (a batched launch): P
    // pictures then skip their own integer search (gpu_compute_set_external_int_me)
    const uint32_t *ext16 = nullptr, *ext8 = nullptr;
    const int *fb_frame = nullptr;  // the frame being dispatched used the learned search: its fallback count
    // map_surface mirror
    void *map_host = nullptr;
    size_t map_size = 0;
    gpu_image_t map_image{};
};

static int align_up(int v, int a) { return (v + a - 1) / a * a; }

// The caller allocates the context; its layout is part of the ABI.
static_assert(sizeof(gpu_image_t) == 48 && sizeof(gpu_memory_t) == 24, "ABI struct size");
static_assert(offsetof(bc250_gpu_context_t, recon_image) == 24 && sizeof(bc250_gpu_context_t) == 104,
              "ABI context layout");

// ------------------------------------------------------------ tiny kernels

__global__ void zero_i32_kernel(int *p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

extern "C" int bc250_launch_zero_i32(int *p, int n, void *stream)
{
    if (n <= 0) return 0;
    zero_i32_kernel<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>(p, n);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// One thread per coded-area byte outside the visible region of a plane.
__global__ void pad_plane_kernel(uint8_t *p, int pitch, int w, int h, int wc, int hc, int bpp)
{
    const int xb = blockIdx.x * blockDim.x + threadIdx.x;  // byte column
    const int y = blockIdx.y;
    if (xb >= wc || y >= hc) return;
    if (xb < w && y < h) return;
    const int sx = xb < w ? xb : (w - bpp) + (xb % bpp);
    const int sy = y < h ? y : h - 1;
    p[(size_t)y * pitch + xb] = p[(size_t)sy * pitch + sx];
}

// bps: bytes per sample (1 NV12, 2 P010); the replicated unit is one luma
// sample and one Cb/Cr pair.
static int pad_edges(uint8_t *y, uint8_t *uv, int py, int puv, int w, int h, int wc, int hc, int bps,
                     hipStream_t st)
{
    if (w == wc && h == hc) return 0;
    // columns first within visible rows, then whole rows: a single pass reads
    // a source that is either visible or itself padded on an earlier row, so
    // run rows in two launches: visible rows, then the rows below.
    dim3 b(256);
    const int wb = w * bps, wcb = wc * bps;
    const int cwb = 2 * ((w + 1) / 2) * bps, ch = (h + 1) / 2;  // visible chroma bytes per row / rows
    if (w < wc) {
        pad_plane_kernel<<<dim3((wcb + 255) / 256, h), b, 0, st>>>(y, py, wb, h, wcb, h, bps);
        if (cwb < wcb) pad_plane_kernel<<<dim3((wcb + 255) / 256, ch), b, 0, st>>>(uv, puv, cwb, ch, wcb, ch, 2 * bps);
    }
    if (h < hc) {
        pad_plane_kernel<<<dim3((wcb + 255) / 256, hc), b, 0, st>>>(y, py, wcb, h, wcb, hc, bps);
        if (ch < hc / 2)
            pad_plane_kernel<<<dim3((wcb + 255) / 256, hc / 2), b, 0, st>>>(uv, puv, wcb, ch, wcb, hc / 2, 2 * bps);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int bc250_launch_pad_edges(uint8_t *y, uint8_t *uv, int py, int puv, int w, int h, int wc, int hc,
                                      void *stream)
{
    return pad_edges(y, uv, py, puv, w, h, wc, hc, 1, (hipStream_t)stream);
}

// ------------------------------------------------------------ buffers

static void free_res(bc250_hip_state *s)
{
    if (s->stream) hipStreamSynchronize(s->stream);
    if (s->cstream) hipStreamSynchronize(s->cstream);
    for (auto &r : s->recon) {
        if (r.base) hipFree(r.base);
        r = Plane2{};
    }
    if (s->me8) hipFree(s->me8);
    s->me8 = nullptr;
    void *dev[] = {s->imv16, s->imv8, s->sync_i, s->sync_d, s->sync_h, s->hcbf, s->hculog2};
    for (void *p : dev)
        if (p) hipFree(p);
    for (auto &sl : s->slot) {
        DevOut &o = sl.dev;
        void *odev[] = {o.levels, o.dc, o.nz, o.modes, o.mvs, o.mv8, o.hmodes, o.lvidx, o.lvdata, o.lvcount,
                        o.rbdone};
        for (void *p : odev)
            if (p) hipFree(p);
        o = DevOut{};
    }
    s->levels = nullptr; s->dc = nullptr; s->nz = nullptr; s->modes = nullptr; s->mvs = nullptr;
    s->mv8 = nullptr; s->imv16 = nullptr; s->imv8 = nullptr; s->sync_i = nullptr; s->sync_d = nullptr;
    s->sync_h = nullptr; s->hmodes = nullptr; s->hcbf = nullptr; s->hculog2 = nullptr;
    for (void *p : {(void *)s->cavlc_out, (void *)s->cavlc_info, s->cavlc_work})
        if (p) hipFree(p);
    {
        Proposer &q = s->prop;
        for (void *p : {(void *)q.pool, (void *)q.cand, (void *)q.list, (void *)q.count,
                        (void *)q.keys})
            if (p) hipFree(p);
        q.cand = q.list = q.count = nullptr;
        q.keys = nullptr;
        q.pool = nullptr;
        q.has_prev = false;
    }
    s->cavlc_out = nullptr; s->cavlc_info = nullptr; s->cavlc_work = nullptr;
    s->cavlc_cap = s->cavlc_work_size = 0;
    for (auto &sl : s->slot) {
        for (void *p : {(void *)sl.rbsp, (void *)sl.rbsp_info})
            if (p) hipHostFree(p);
        sl.rbsp = nullptr; sl.rbsp_info = nullptr; sl.slices = 0;
        void *host[] = {sl.levels, sl.dc, sl.modes, sl.nz, sl.mvs, sl.mv8, sl.hmodes, sl.err, sl.lvidx, sl.lvdata,
                        sl.lvcount};
        for (void *p : host)
            if (p) hipHostFree(p);
        free(sl.built_idx);
        sl.levels = nullptr; sl.dc = nullptr; sl.modes = nullptr; sl.nz = nullptr; sl.mvs = nullptr;
        sl.mv8 = nullptr; sl.hmodes = nullptr; sl.err = nullptr;
        sl.lvidx = nullptr; sl.lvdata = nullptr; sl.lvcount = nullptr; sl.built_idx = nullptr;
        sl.lv_packed = sl.lv_built = false;
        sl.built_n = 0;
        sl.built_all = true;
        sl.pending = sl.have_data = false;
    }
    s->w = s->h = s->wc = s->hc = s->mbw = s->mbh = s->nmb = 0;
    s->has_recon = false;
}

static int alloc_plane2(Plane2 *p, int wc, int hc, int bps)
{
    p->py = align_up(wc * bps, 256);
    p->puv = p->py;
    size_t ysz = (size_t)p->py * hc, uvsz = (size_t)p->puv * (hc / 2);
    if (!HIP_OK(hipMalloc(&p->base, ysz + uvsz))) return -1;
    p->y = (uint8_t *)p->base;
    p->uv = (uint8_t *)p->base + ysz;
    return 0;
}

// Lazily (re)allocates everything for a resolution, like
// allocate_encoding_buffers() at gpu_compute.c:233-532.
static int ensure_res(bc250_hip_state *s, int w, int h, int bps)
{
    if (s->w == w && s->h == h && s->bps == bps && s->levels) return 0;
    free_res(s);
    const int wc = align_up(w, 16), hc = align_up(h, 16);
    const int mbw = wc / 16, mbh = hc / 16, nmb = mbw * mbh;
    bool ok = alloc_plane2(&s->recon[0], wc, hc, bps) == 0 && alloc_plane2(&s->recon[1], wc, hc, bps) == 0;
    s->me8_pitch = align_up(wc, 256);
    if (bps == 2) ok = ok && HIP_OK(hipMalloc(&s->me8, (size_t)s->me8_pitch * hc * 2));
    for (auto &sl : s->slot) {
        DevOut &o = sl.dev;
        ok = ok && HIP_OK(hipMalloc(&o.levels, (size_t)nmb * 384 * 2)) && HIP_OK(hipMalloc(&o.dc, (size_t)nmb * 24 * 4)) &&
             HIP_OK(hipMalloc(&o.nz, (size_t)nmb * 24 * 4)) && HIP_OK(hipMalloc(&o.modes, (size_t)nmb * 4)) &&
             HIP_OK(hipMalloc(&o.mvs, (size_t)nmb * 16)) && HIP_OK(hipMalloc(&o.mv8, (size_t)nmb * 64)) &&
             HIP_OK(hipMalloc(&o.hmodes, (size_t)nmb * 16)) &&
             HIP_OK(hipMalloc(&o.lvidx, (size_t)nmb * 24 * 4)) && HIP_OK(hipMalloc(&o.lvdata, (size_t)nmb * 24 * 32)) &&
             HIP_OK(hipMalloc(&o.lvcount, 4)) && HIP_OK(hipMemset(o.lvcount, 0, 4)) &&
             HIP_OK(hipMalloc(&o.rbdone, 4)) && HIP_OK(hipMemset(o.rbdone, 0, 4));
    }
    ok = ok && HIP_OK(hipMalloc(&s->imv16, (size_t)nmb * 16)) && HIP_OK(hipMalloc(&s->imv8, (size_t)nmb * 64)) &&
         HIP_OK(hipMalloc(&s->sync_i, bc250_intra_work_bytes(wc, hc))) &&
         HIP_OK(hipMalloc(&s->sync_d, bc250_deblock_work_bytes(wc, hc))) &&
         HIP_OK(hipMalloc(&s->sync_h, bc250_hevc_intra_work_bytes(wc, hc))) &&
         HIP_OK(hipMalloc(&s->hcbf, (size_t)nmb * 4)) &&
         HIP_OK(hipMalloc(&s->hculog2, (size_t)nmb * 4));
    for (auto &sl : s->slot) {
        ok = ok && HIP_OK(hipHostMalloc((void **)&sl.levels, (size_t)nmb * 384 * 2, hipHostMallocDefault)) &&
             HIP_OK(hipHostMalloc((void **)&sl.dc, (size_t)nmb * 24 * 4, hipHostMallocDefault)) &&
             HIP_OK(hipHostMalloc((void **)&sl.modes, (size_t)nmb * 4, hipHostMallocDefault)) &&
             HIP_OK(hipHostMalloc((void **)&sl.nz, (size_t)nmb * 24 * 4, hipHostMallocDefault)) &&
             HIP_OK(hipHostMalloc((void **)&sl.mvs, (size_t)nmb * 16, hipHostMallocDefault)) &&
             HIP_OK(hipHostMalloc((void **)&sl.mv8, (size_t)nmb * 64, hipHostMallocDefault)) &&
             HIP_OK(hipHostMalloc((void **)&sl.hmodes, (size_t)nmb * 16, hipHostMallocDefault)) &&
             HIP_OK(hipHostMalloc((void **)&sl.err, 4 * sizeof(int), hipHostMallocDefault)) &&
             HIP_OK(hipHostMalloc((void **)&sl.lvidx, (size_t)nmb * 24 * 4, hipHostMallocDefault)) &&
             HIP_OK(hipHostMalloc((void **)&sl.lvdata, (size_t)nmb * 24 * 32, hipHostMallocDefault)) &&
             HIP_OK(hipHostMalloc((void **)&sl.lvcount, sizeof(int), hipHostMallocDefault));
        sl.built_idx = ok ? (uint32_t *)malloc((size_t)nmb * 24 * 4) : nullptr;
        ok = ok && sl.built_idx;
        sl.built_n = 0;
        sl.built_all = true;
        sl.lv_packed = sl.lv_built = false;
    }
    // wavefront workspaces start zeroed: no hand-off record or epoch holds a
    // value from whatever used the memory before
    ok = ok && HIP_OK(hipMemset(s->sync_i, 0, bc250_intra_work_bytes(wc, hc))) &&
         HIP_OK(hipMemset(s->sync_d, 0, bc250_deblock_work_bytes(wc, hc))) &&
         HIP_OK(hipMemset(s->sync_h, 0, bc250_hevc_intra_work_bytes(wc, hc)));
    if (!ok) {
        fprintf(stderr, "[bc250-hip] buffer allocation failed for %dx%d\n", w, h);
        free_res(s);
        return -1;
    }
    s->w = w; s->h = h; s->wc = wc; s->hc = hc; s->mbw = mbw; s->mbh = mbh; s->nmb = nmb; s->bps = bps;
    s->has_recon = false;
    s->ref_idx = 0;
    return 0;
}

// Points the context's output pointers at slot `k`'s device buffers.
static void bind_slot(bc250_hip_state *s, int k)
{
    const DevOut &o = s->slot[k].dev;
    s->levels = o.levels;
    s->dc = o.dc;
    s->nz = o.nz;
    s->modes = o.modes;
    s->mvs = o.mvs;
    s->mv8 = o.mv8;
    s->hmodes = o.hmodes;
}

// Grows the pinned bounce buffer of uploads / downloads to `bytes`.
static void *ensure_xfer(bc250_hip_state *s, size_t bytes)
{
    if (s->xfer_size >= bytes) return s->xfer;
    if (s->xfer) hipHostFree(s->xfer);
    s->xfer = nullptr;
    s->xfer_size = 0;
    if (!HIP_OK(hipHostMalloc(&s->xfer, bytes, hipHostMallocDefault))) return nullptr;
    s->xfer_size = bytes;
    return s->xfer;
}

// Packs a P frame's levels for the readback: each thread one 16-level group
// (32 bytes); the groups holding a nonzero level go to idx / data at slots
// taken one atomic add per wave.  *count starts at zero: it is cleared at
// allocation and by each readback of it.
__global__ __launch_bounds__(256) void pack_levels_kernel(const int16_t *__restrict__ levels, int ngroups,
                                                          uint32_t *__restrict__ idx, uint4 *__restrict__ data,
                                                          int *__restrict__ count)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x, lane = threadIdx.x & 63;
    uint4 a = make_uint4(0, 0, 0, 0), b = a;
    if (g < ngroups) {
        const uint4 *p = (const uint4 *)(levels + (size_t)g * 16);
        a = p[0];
        b = p[1];
    }
    const bool nz = (a.x | a.y | a.z | a.w | b.x | b.y | b.z | b.w) != 0;
    const uint64_t m = __ballot(nz);
    if (!m) return;
    const int first = __ffsll((unsigned long long)m) - 1;
    int base = 0;
    if (lane == first) base = atomicAdd(count, __popcll(m));
    base = __shfl(base, first);
    if (nz) {
        const int k = base + __popcll(m & ((1ull << lane) - 1));
        idx[k] = (uint32_t)g;
        data[2 * k] = a;
        data[2 * k + 1] = b;
    }
}

// The readback of a finished frame, one launch on the copy stream: the
// kernel writes the frame's outputs straight into the slot's pinned staging
// (host memory that the device addresses): the listed arrays whole, the
// packed levels as many groups as were packed (the count read here, on the
// device), and the three flag words.  The last block to finish clears the
// packed count for the slot's next frame.
constexpr int RB_SPANS = 8, RB_BLOCKS = 32;
struct RbSpan {
    const uint32_t *src;
    uint32_t *dst;
    uint32_t words;
};
struct RbArgs {
    RbSpan sp[RB_SPANS];
    int nsp;
    const uint4 *mv8;        // non-null: n8 records {x, y, sad, 0} go to mv8_dst as {x | y << 16, sad}
    uint2 *mv8_dst;
    int n8;
    const int *flag_src[3];  // null: the flag takes flag_def
    int flag_def[3];
    int *flag_dst;
    const int *lvcount;      // null: no packed levels
    int maxrec;
    const uint32_t *lvidx;
    const uint4 *lvdata;
    int *lvcount_dst;
    uint32_t *lvidx_dst;
    uint4 *lvdata_dst;
    unsigned *done;
};

__global__ __launch_bounds__(256) void readback_kernel(const RbArgs a)
{
    const int tid = blockIdx.x * blockDim.x + threadIdx.x, nth = gridDim.x * blockDim.x;
    const int cnt = a.lvcount ? min(max(*a.lvcount, 0), a.maxrec) : 0;
    for (int k = 0; k < a.nsp; k++) {
        const RbSpan sp = a.sp[k];
        if ((sp.words & 3) == 0) {
            const uint4 *s4 = (const uint4 *)sp.src;
            uint4 *d4 = (uint4 *)sp.dst;
            for (uint32_t i = tid; i < sp.words / 4; i += nth) d4[i] = s4[i];
        } else {
            for (uint32_t i = tid; i < sp.words; i += nth) sp.dst[i] = sp.src[i];
        }
    }
    for (int i = tid; i < a.n8; i += nth) {
        const uint4 v = a.mv8[i];
        a.mv8_dst[i] = make_uint2((v.x & 0xffffu) | (v.y << 16), v.z);
    }
    if (a.lvcount) {
        for (int i = tid; i < cnt; i += nth) a.lvidx_dst[i] = a.lvidx[i];
        for (int i = tid; i < 2 * cnt; i += nth) a.lvdata_dst[i] = a.lvdata[i];
    }
    if (tid == 0) {
        for (int i = 0; i < 3; i++) a.flag_dst[i] = a.flag_src[i] ? *a.flag_src[i] : a.flag_def[i];
        if (a.lvcount) *a.lvcount_dst = cnt;
    }
    if (a.lvcount) {
        // every thread of a block has read the count before the block is
        // counted, so the last block counted clears it after all reads
        __shared__ bool last;
        __syncthreads();
        if (threadIdx.x == 0) {
            __threadfence();
            last = atomicAdd(a.done, 1u) == gridDim.x - 1;
        }
        __syncthreads();
        if (last && threadIdx.x == 0) {
            atomicExch(const_cast<int *>(a.lvcount), 0);
            atomicExch(a.done, 0u);
        }
    }
}

// Device and pinned host buffers of the CAVLC writer: room for 2560 bytes
// per MB (the reference's CAVLC allowance) and 4 KB per slice.
static int ensure_cavlc(bc250_hip_state *s)
{
    if (s->cavlc_out) return 0;
    const size_t cap = align_up((int)((size_t)s->nmb * 2560 + 4096 * 64), 256);
    const size_t work = bc250_h264_cavlc_work_size(s->mbw, s->mbh);
    bool ok = work && HIP_OK(hipMalloc((void **)&s->cavlc_out, cap)) &&
              HIP_OK(hipMalloc((void **)&s->cavlc_info, (4 * 64 + 2) * 4)) && HIP_OK(hipMalloc(&s->cavlc_work, work));
    for (auto &sl : s->slot)
        ok = ok && HIP_OK(hipHostMalloc((void **)&sl.rbsp, cap, hipHostMallocDefault)) &&
             HIP_OK(hipHostMalloc((void **)&sl.rbsp_info, (4 * 64 + 2) * 4, hipHostMallocDefault));
    if (!ok) {
        fprintf(stderr, "[bc250-hip] CAVLC buffer allocation failed\n");
        return -1;
    }
    s->cavlc_cap = cap;
    s->cavlc_work_size = work;
    return 0;
}

// The proposer's per-resolution buffers (block sums of every 64x64 tile of
// the picture, candidates, fallback list, split-search keys), allocated at
// its first frame at a resolution.
static int ensure_prop(bc250_hip_state *s)
{
    Proposer &q = s->prop;
    if (q.pool) return 0;
    const int nt = ((s->mbw + 3) / 4) * ((s->mbh + 3) / 4);
    bool ok = HIP_OK(hipMalloc(&q.pool, (size_t)nt * 16 * 32 * 4)) &&
              HIP_OK(hipMalloc(&q.cand, (size_t)s->nmb * 4 * 8)) && HIP_OK(hipMalloc(&q.list, (size_t)s->nmb * 4)) &&
              HIP_OK(hipMalloc(&q.count, 4)) &&
              HIP_OK(hipMalloc(&q.keys, (size_t)s->nmb * 5 * 4));
    if (!ok) {
        fprintf(stderr, "[bc250-hip] proposer buffer allocation failed\n");
        return -1;
    }
    q.has_prev = false;
    return 0;
}

// The learned integer search of one P picture: the network's candidates
// refined in +-4 windows, the fallback list searched in full; the result is
// out16 / out8 in the full search's format, which keeps the 16x16 vectors
// as the next picture's neighbour input.
static int enqueue_learned(bc250_hip_state *s, const uint8_t *c8, int cp, const uint8_t *r8, int rp, uint32_t *out16,
                           uint32_t *out8, hipStream_t st, bool keys = false)
{
    Proposer &q = s->prop;
    if (ensure_prop(s) != 0) return -1;
    const gpu_encode_params_t &p = s->params;
    // out16 still holds the previous picture's records when this picture's
    // proposer reads them: the candidate search rewrites it only after that
    // launch on the same stream
    // keys: the CTU16 HEVC coding makes the records of the MBs searched in
    // full from the keys itself (one launch fewer)
    int r = (keys ? bc250_learned_search_picture_keys : bc250_learned_search_picture)(
        c8, cp, r8, rp, s->wc, s->hc, p.search_range, p.lambda8, q.params, q.packed, q.has_prev ? out16 : nullptr,
        q.cfg.ncand, q.cfg.alpha_q4, q.cfg.beta, q.cfg.t_high, q.pool, q.cand, q.list, q.count, q.keys, out16, out8, st);
    q.ran = r == 0;
    s->fs_keys = r == 0 && keys ? q.keys : nullptr;
    return r;
}

static bc250_hip_state *st_of(gpu_context_t *ctx) { return ctx ? ctx->hip : nullptr; }

// Describes the newest reconstruction in ctx->recon_image / recon_memory, at
// its coded size so that a map/unmap round trip pads nothing.
static void publish_recon(gpu_context_t *ctx, bc250_hip_state *s)
{
    const Plane2 &r = s->recon[s->ref_idx];
    gpu_image_t &im = ctx->recon_image;
    im.y_plane = r.y;
    im.uv_plane = r.uv;
    im.width = im.coded_width = (uint32_t)s->wc;
    im.height = im.coded_height = (uint32_t)s->hc;
    im.y_pitch = (uint32_t)r.py;
    im.uv_pitch = (uint32_t)r.puv;
    im.current_layout = 0;
    im.format = s->bps == 2 ? GPU_IMAGE_P010 : GPU_IMAGE_NV12;
    ctx->recon_memory.memory = r.base;
    ctx->recon_memory.size = (size_t)r.py * s->hc + (size_t)r.puv * (s->hc / 2);
    ctx->recon_memory.mapped_ptr = nullptr;
    ctx->has_recon_frame = true;
}

// ------------------------------------------------------------ lifecycle

extern "C" int bc250_gpu_init(bc250_gpu_context_t *ctx)
{
    if (!ctx) return -1;
    if (ctx->hip) return 0;
    int n = 0;
    if (!HIP_OK(hipGetDeviceCount(&n)) || n < 1) {
        fprintf(stderr, "[bc250-hip] no HIP device\n");
        return -1;
    }
    if (!HIP_OK(hipSetDevice(ctx->device))) return -1;
    bc250_hip_state *s = new bc250_hip_state();
    s->timing = getenv("BC250_PERF_STATS") != nullptr;
    if (!HIP_OK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking))) {
        delete s;
        return -1;
    }
    if (!HIP_OK(hipStreamCreateWithFlags(&s->cstream, hipStreamNonBlocking))) {
        hipStreamDestroy(s->stream);
        delete s;
        return -1;
    }
    bool ev_ok = true;
    for (auto &sl : s->slot) {
        for (auto &e : sl.ev) ev_ok = ev_ok && HIP_OK(hipEventCreate(&e));
        ev_ok = ev_ok && HIP_OK(hipEventCreateWithFlags(&sl.cdone, hipEventDisableTiming)) &&
                HIP_OK(hipEventCreateWithFlags(&sl.cearly, hipEventDisableTiming));
    }
    if (!ev_ok) {
        for (auto &sl : s->slot) {
            for (auto &e : sl.ev)
                if (e) hipEventDestroy(e);
            if (sl.cdone) hipEventDestroy(sl.cdone);
            if (sl.cearly) hipEventDestroy(sl.cearly);
        }
        hipStreamDestroy(s->cstream);
        hipStreamDestroy(s->stream);
        delete s;
        return -1;
    }
    ctx->hip = s;
    ctx->frames_submitted = 0;
    ctx->has_recon_frame = false;
    ctx->any_surface_exported = false;
    return 0;
}

extern "C" void bc250_gpu_destroy(bc250_gpu_context_t *ctx)
{
    bc250_hip_state *s = st_of(ctx);
    if (!s) return;
    free_res(s);
    if (s->map_host) hipHostFree(s->map_host);
    if (s->xfer) hipHostFree(s->xfer);
    if (s->prop.params) hipFree(s->prop.params);
    if (s->prop.packed) hipFree(s->prop.packed);
    for (auto &sl : s->slot) {
        for (auto &e : sl.ev)
            if (e) hipEventDestroy(e);
        if (sl.cdone) hipEventDestroy(sl.cdone);
        if (sl.cearly) hipEventDestroy(sl.cearly);
    }
    hipStreamDestroy(s->cstream);
    hipStreamDestroy(s->stream);
    delete s;
    ctx->hip = nullptr;
    ctx->recon_image = gpu_image_t{};
    ctx->recon_memory = gpu_memory_t{};
    ctx->has_recon_frame = false;
}

extern "C" void gpu_compute_report_cu_topology(bc250_gpu_context_t *ctx)
{
    hipDeviceProp_t p;
    if (!ctx || !HIP_OK(hipGetDeviceProperties(&p, ctx->device))) return;
    fprintf(stderr, "[bc250-hip] %s (%s): %d CUs, %d KB LDS/block, %.1f GB HBM\n", p.name, p.gcnArchName,
            p.multiProcessorCount, (int)(p.sharedMemPerBlock / 1024), p.totalGlobalMem / 1073741824.0);
}

extern "C" int gpu_compute_init(gpu_context_t *ctx) { return bc250_gpu_init(ctx); }
extern "C" void gpu_compute_terminate(gpu_context_t *ctx) { bc250_gpu_destroy(ctx); }

// ------------------------------------------------------------ surfaces

extern "C" int gpu_compute_create_image(gpu_context_t *ctx, int width, int height, int format, gpu_image_t *image,
                                        gpu_memory_t *memory)
{
    if (!ctx || !ctx->hip || !image || !memory || width <= 0 || height <= 0) return -1;
    const int bpp = format == GPU_IMAGE_P010 ? 2 : 1;
    const int wc = align_up(width, 16), hc = align_up(height, 16);
    const int pitch = align_up(wc * bpp, 256);
    const size_t ysz = (size_t)pitch * hc, total = ysz + (size_t)pitch * (hc / 2);
    void *base = nullptr;
    if (!HIP_OK(hipMalloc(&base, total))) return -1;
    memset(image, 0, sizeof(*image));
    image->y_plane = base;
    image->uv_plane = (uint8_t *)base + ysz;
    image->width = (uint32_t)width;
    image->height = (uint32_t)height;
    image->coded_width = (uint32_t)wc;
    image->coded_height = (uint32_t)hc;
    image->y_pitch = image->uv_pitch = (uint32_t)pitch;
    image->format = format;
    memory->memory = base;
    memory->size = total;
    memory->mapped_ptr = nullptr;
    return 0;
}

extern "C" void gpu_compute_destroy_image(gpu_context_t *ctx, gpu_image_t image, gpu_memory_t memory)
{
    (void)image;
    if (ctx && ctx->hip) {
        hipStreamSynchronize(ctx->hip->stream);
        hipStreamSynchronize(ctx->hip->cstream);
    }
    if (memory.memory) hipFree(memory.memory);
}

extern "C" int gpu_compute_get_nv12_layout(gpu_context_t *ctx, gpu_image_t *image, gpu_memory_t memory,
                                           gpu_nv12_layout_t *layout)
{
    if (!ctx || !image || !layout || !memory.memory) return -1;
    layout->y_pitch = image->y_pitch;
    layout->y_offset = 0;
    layout->uv_pitch = image->uv_pitch;
    layout->uv_offset = (uint64_t)((uint8_t *)image->uv_plane - (uint8_t *)image->y_plane);
    layout->total_size = memory.size;
    return 0;
}

extern "C" uint8_t *gpu_compute_map_surface(gpu_context_t *ctx, gpu_image_t *image, gpu_memory_t memory,
                                            gpu_nv12_layout_t *layout, bool *unmap)
{
    bc250_hip_state *s = st_of(ctx);
    if (!s || !image || !memory.memory) return nullptr;
    if (s->map_size < memory.size) {
        if (s->map_host) hipHostFree(s->map_host);
        s->map_host = nullptr;
        s->map_size = 0;
        if (!HIP_OK(hipHostMalloc(&s->map_host, memory.size, hipHostMallocDefault))) return nullptr;
        s->map_size = memory.size;
    }
    if (!HIP_OK(hipMemcpyAsync(s->map_host, memory.memory, memory.size, hipMemcpyDeviceToHost, s->stream)) ||
        !HIP_OK(hipStreamSynchronize(s->stream)))
        return nullptr;
    s->map_image = *image;
    if (layout) gpu_compute_get_nv12_layout(ctx, image, memory, layout);
    if (unmap) *unmap = true;
    return (uint8_t *)s->map_host;
}

extern "C" void gpu_compute_unmap_surface(gpu_context_t *ctx, gpu_memory_t memory, bool unmap)
{
    bc250_hip_state *s = st_of(ctx);
    if (!s || !unmap || !s->map_host || !memory.memory) return;
    hipMemcpyAsync(memory.memory, s->map_host, memory.size, hipMemcpyHostToDevice, s->stream);
    const gpu_image_t &im = s->map_image;
    pad_edges((uint8_t *)im.y_plane, (uint8_t *)im.uv_plane, im.y_pitch, im.uv_pitch, im.width, im.height,
              im.coded_width, im.coded_height, im.format == GPU_IMAGE_P010 ? 2 : 1, s->stream);
    hipStreamSynchronize(s->stream);
}

extern "C" int gpu_compute_upload_nv12(gpu_context_t *ctx, gpu_image_t *image, gpu_memory_t memory,
                                       const uint8_t *y_plane, int y_pitch, const uint8_t *uv_plane, int uv_pitch,
                                       int width, int height)
{
    bc250_hip_state *s = st_of(ctx);
    if (!s || !image || !memory.memory || !y_plane || !uv_plane) return -1;
    const int bpp = image->format == GPU_IMAGE_P010 ? 2 : 1;
    if (width <= 0 || height <= 0 || (uint32_t)width > image->width || (uint32_t)height > image->height) return -1;
    // the caller's (pageable) rows are packed into the context's pinned
    // buffer on the host, then copied by DMA in stream order
    const size_t yb = (size_t)width * bpp, uvb = (size_t)((width + 1) / 2) * 2 * bpp;
    const int ch = (height + 1) / 2;
    uint8_t *x = (uint8_t *)ensure_xfer(s, yb * height + uvb * ch);
    if (!x) return -1;
    for (int r = 0; r < height; r++) memcpy(x + yb * r, y_plane + (size_t)y_pitch * r, yb);
    for (int r = 0; r < ch; r++) memcpy(x + yb * height + uvb * r, uv_plane + (size_t)uv_pitch * r, uvb);
    bool ok = HIP_OK(hipMemcpy2DAsync(image->y_plane, image->y_pitch, x, yb, yb, height, hipMemcpyHostToDevice,
                                      s->stream)) &&
              HIP_OK(hipMemcpy2DAsync(image->uv_plane, image->uv_pitch, x + yb * height, uvb, uvb, ch,
                                      hipMemcpyHostToDevice, s->stream));
    if (ok)
        ok = pad_edges((uint8_t *)image->y_plane, (uint8_t *)image->uv_plane, image->y_pitch, image->uv_pitch, width,
                       height, image->coded_width, image->coded_height, bpp, s->stream) == 0;
    ok = ok && HIP_OK(hipStreamSynchronize(s->stream));
    return ok ? 0 : -1;
}

// P010 keeps 10-bit samples in the high bits of 16-bit words (reference
// gpu_compute.h:409-420); the shift happens on the host copy.
extern "C" int gpu_compute_upload_p010(gpu_context_t *ctx, gpu_image_t *image, gpu_memory_t memory,
                                       const uint16_t *y_plane, int y_pitch, const uint16_t *uv_plane, int uv_pitch,
                                       int width, int height)
{
    if (!image || image->format != GPU_IMAGE_P010 || !y_plane || !uv_plane || width <= 0 || height <= 0) return -1;
    const int cw = (width + 1) / 2, ch = (height + 1) / 2;
    uint16_t *ty = (uint16_t *)malloc((size_t)width * height * 2);
    uint16_t *tuv = (uint16_t *)malloc((size_t)cw * 2 * ch * 2);
    if (!ty || !tuv) {
        free(ty);
        free(tuv);
        return -1;
    }
    for (int y = 0; y < height; y++)
        for (int x = 0; x < width; x++)
            ty[(size_t)y * width + x] = (uint16_t)(((const uint16_t *)((const uint8_t *)y_plane + (size_t)y * y_pitch))[x] << 6);
    for (int y = 0; y < ch; y++)
        for (int x = 0; x < 2 * cw; x++)
            tuv[(size_t)y * 2 * cw + x] = (uint16_t)(((const uint16_t *)((const uint8_t *)uv_plane + (size_t)y * uv_pitch))[x] << 6);
    int r = gpu_compute_upload_nv12(ctx, image, memory, (const uint8_t *)ty, width * 2, (const uint8_t *)tuv, cw * 4,
                                    width, height);
    free(ty);
    free(tuv);
    return r;
}

extern "C" int gpu_compute_video_proc(gpu_context_t *, gpu_image_t *, const int *, gpu_image_t *, const int *)
{
    return -1;  // VideoProc entrypoint: not part of the encoder compute path
}

extern "C" void gpu_compute_copy_from_wc(void *dst, const void *src, size_t n) { memcpy(dst, src, n); }

// Device planes -> the caller's rows, through the pinned bounce buffer.
static int download_planes(bc250_hip_state *s, const uint8_t *dy, int dpy, const uint8_t *duv, int dpuv,
                           uint8_t *y_plane, int y_pitch, uint8_t *uv_plane, int uv_pitch, size_t row_bytes,
                           size_t uvb, int height)
{
    const int ch = (height + 1) / 2;
    uint8_t *x = (uint8_t *)ensure_xfer(s, row_bytes * height + uvb * ch);
    if (!x) return -1;
    bool ok = HIP_OK(hipMemcpy2DAsync(x, row_bytes, dy, dpy, row_bytes, height, hipMemcpyDeviceToHost, s->stream)) &&
              HIP_OK(hipMemcpy2DAsync(x + row_bytes * height, uvb, duv, dpuv, uvb, ch, hipMemcpyDeviceToHost,
                                      s->stream)) &&
              HIP_OK(hipStreamSynchronize(s->stream));
    if (!ok) return -1;
    for (int r = 0; r < height; r++) memcpy(y_plane + (size_t)y_pitch * r, x + row_bytes * r, row_bytes);
    for (int r = 0; r < ch; r++) memcpy(uv_plane + (size_t)uv_pitch * r, x + row_bytes * height + uvb * r, uvb);
    return 0;
}

extern "C" int gpu_compute_download_nv12(gpu_context_t *ctx, gpu_image_t *image, gpu_memory_t memory,
                                         uint8_t *y_plane, int y_pitch, uint8_t *uv_plane, int uv_pitch, int width,
                                         int height)
{
    bc250_hip_state *s = st_of(ctx);
    if (!s || !image || !memory.memory || !y_plane || !uv_plane || width <= 0 || height <= 0) return -1;
    const int bpp = image->format == GPU_IMAGE_P010 ? 2 : 1;
    if ((uint32_t)width > image->coded_width || (uint32_t)height > image->coded_height) return -1;
    return download_planes(s, (const uint8_t *)image->y_plane, image->y_pitch, (const uint8_t *)image->uv_plane,
                           image->uv_pitch, y_plane, y_pitch, uv_plane, uv_pitch, (size_t)width * bpp,
                           (size_t)((width + 1) / 2) * 2 * bpp, height);
}

extern "C" int gpu_compute_export_nv12_dmabuf(gpu_context_t *, gpu_memory_t, int *) { return -1; }
extern "C" int gpu_compute_import_dmabuf_image(gpu_context_t *, int, int, int, int, uint32_t, uint32_t,
                                               gpu_image_t *, gpu_memory_t *)
{
    return -1;
}
// Surfaces here have no foreign writers (no dma-buf import) and uploads
// complete before they return.  The GPU-side wait is therefore the
// reference's "capability missing" no-op (-1: proceed); the host-side wait
// reports the surface quiescent (0), which is what lets the CPU ME route read
// it (own GPU work is gated separately by gpu_compute_cpu_read_safe).
extern "C" int gpu_compute_wait_for_image_ready(gpu_context_t *, gpu_memory_t) { return -1; }
extern "C" int gpu_compute_wait_for_image_ready_host(gpu_context_t *ctx, gpu_memory_t memory)
{
    return st_of(ctx) && memory.memory ? 0 : -1;
}

extern "C" int gpu_compute_cpu_read_safe(gpu_context_t *ctx)
{
    bc250_hip_state *s = st_of(ctx);
    if (!s) return 0;
    return hipStreamQuery(s->stream) == hipSuccess ? 1 : 0;
}

// ------------------------------------------------------------ encode

// Waits for a frame's last event by polling it: a blocking wait can sleep the
// host thread past the frame still queued behind it, which leaves the GPU
// idle between frames.
static hipError_t wait_event(hipEvent_t ev)
{
    hipError_t e;
    while ((e = hipEventQuery(ev)) == hipErrorNotReady) __builtin_ia32_pause();
    return e;
}

extern "C" int gpu_compute_begin_picture(gpu_context_t *ctx, gpu_image_t render_target)
{
    (void)render_target;
    bc250_hip_state *s = st_of(ctx);
    if (!s) return -1;
    Slot &sl = s->slot[s->cur];
    if (sl.pending) {
        // the slot is about to be reused: its previous frame must be done
        if (!HIP_OK(wait_event(sl.ev[NEV - 1]))) return -1;
        sl.pending = false;
    }
    return 0;
}

// A stage boundary: its event, recorded only while stage timing is on (an
// event with a timestamp costs the stream a few microseconds).
static void mark(const bc250_hip_state *s, hipEvent_t e, hipStream_t st)
{
    if (s->timing) hipEventRecord(e, st);
}

// Starts a frame in the current slot: the compute stream first waits for the
// slot's previous copies (its device outputs are about to be rewritten), then
// the slot's device buffers become the frame's outputs.
static void start_frame(bc250_hip_state *s)
{
    Slot &sl = s->slot[s->cur];
    if (sl.recorded) hipStreamWaitEvent(s->stream, sl.ev[NEV - 1], 0);
    sl.timed = s->timing;
    bind_slot(s, s->cur);
    // the learned search's neighbour input is the previous picture's vectors
    // only when that picture ran it (not after an I picture)
    s->prop.has_prev = s->prop.ran;
    s->prop.ran = false;
    s->fb_frame = nullptr;
    s->fs_keys = nullptr;
}

// Ends a frame's stages: the wavefront error flags (e0 intra, e1 deblock;
// null = the stage did not run) and the learned search's fallback count
// (fb, null = not used) go to the slot's pinned flag words, and, with
// readback, the listed outputs to the slot's pinned staging, by one
// readback_kernel on the copy stream once the stages are done.
// RB_PACK (P frames): the levels travel packed (pack_levels_kernel first),
// and the staging getter rebuilds the full levels on the host from them.
// RB_PACKED: the stages packed the levels themselves (no packing launch).
// RB_MV8_HALF: the 8x8 vectors travel as 8 bytes each (quarter-sample
// components within int16) and the staging getter widens them.
// RB_HMODES_ZERO: the PU modes are all zero; the getter clears the staging.
// early: the outputs and flags were final at sl.cearly, recorded by the
// caller before the frame's last stages (deblocking, which changes only the
// reconstruction): the readback overlaps those stages, and the copy stream
// still waits for the whole frame before the frame's last event.
enum {
    RB_LEVELS = 1, RB_DC = 2, RB_MODES = 4, RB_MVS = 8, RB_NZ = 16, RB_MV8 = 32, RB_HMODES = 64, RB_PACK = 128,
    RB_PACKED = 256, RB_MV8_HALF = 512, RB_HMODES_ZERO = 1024
};
static int finish_frame(bc250_hip_state *s, Slot &sl, const int *e0, const int *e1, const int *fb, int what,
                        bool early = false)
{
    const size_t n = (size_t)s->nmb;
    const bool pack = s->params.readback && (what & RB_LEVELS) && (what & RB_PACK);
    if (pack && !(what & RB_PACKED)) {
        const int ng = s->nmb * 24;
        pack_levels_kernel<<<(ng + 255) / 256, 256, 0, s->stream>>>(sl.dev.levels, ng, sl.dev.lvidx, sl.dev.lvdata,
                                                                     sl.dev.lvcount);
    }
    if (!HIP_OK(hipGetLastError()) || !HIP_OK(hipEventRecord(sl.cdone, s->stream)) ||
        !HIP_OK(hipStreamWaitEvent(s->cstream, early ? sl.cearly : sl.cdone, 0)))
        return -1;
    const DevOut &o = sl.dev;
    RbArgs a{};
    a.flag_src[0] = e0;
    a.flag_src[1] = e1;
    a.flag_src[2] = fb;
    a.flag_def[2] = -1;
    a.flag_dst = sl.err;
    a.done = o.rbdone;
    if (s->params.readback) {
        sl.lv_packed = pack;
        sl.lv_settled = sl.lv_built = false;
        if (pack) {
            a.lvcount = o.lvcount;
            a.maxrec = s->nmb * 24;
            a.lvidx = o.lvidx;
            a.lvdata = o.lvdata;
            a.lvcount_dst = sl.lvcount;
            a.lvidx_dst = sl.lvidx;
            a.lvdata_dst = (uint4 *)sl.lvdata;
            what &= ~RB_LEVELS;
        } else if (what & RB_LEVELS) {
            sl.built_all = true;  // the whole array arrives: any group may be nonzero
        }
        auto add = [&](const void *src, void *dst, size_t bytes) {
            a.sp[a.nsp++] = RbSpan{(const uint32_t *)src, (uint32_t *)dst, (uint32_t)(bytes / 4)};
        };
        if (what & RB_LEVELS) add(o.levels, sl.levels, n * 384 * 2);
        if (what & RB_DC) add(o.dc, sl.dc, n * 24 * 4);
        if (what & RB_MODES) add(o.modes, sl.modes, n * 4);
        if (what & RB_MVS) add(o.mvs, sl.mvs, n * 16);
        if (what & RB_NZ) add(o.nz, sl.nz, n * 24 * 4);
        sl.mv8_half = (what & RB_MV8) && (what & RB_MV8_HALF);
        sl.hmodes_zero = (what & RB_HMODES) && (what & RB_HMODES_ZERO);
        if (sl.mv8_half) {
            a.mv8 = (const uint4 *)o.mv8;
            a.mv8_dst = (uint2 *)sl.mv8;
            a.n8 = (int)n * 4;
        } else if (what & RB_MV8) {
            add(o.mv8, sl.mv8, n * 64);
        }
        if ((what & RB_HMODES) && !sl.hmodes_zero) add(o.hmodes, sl.hmodes, n * 16);
    }
    readback_kernel<<<RB_BLOCKS, 256, 0, s->cstream>>>(a);
    if (early && !HIP_OK(hipStreamWaitEvent(s->cstream, sl.cdone, 0))) return -1;
    return HIP_OK(hipGetLastError()) ? 0 : -1;
}

static int enqueue_me(bc250_hip_state *s, const gpu_image_t &rt, const Plane2 &ref, int me_mode, hipEvent_t mid)
{
    const gpu_encode_params_t &p = s->params;
    if (p.diamond) {
        // the reference's diamond search: final quarter-sample vectors, no 8x8 ones
        int r = bc250_launch_diamond_me((const uint8_t *)rt.y_plane, rt.y_pitch, ref.y, ref.py, s->wc, s->hc,
                                        me_mode == 1 ? 8 : 5, s->mvs, s->stream);
        mark(s, mid, s->stream);
        if (r == 0) r = bc250_launch_zero_i32((int *)s->mv8, s->nmb * 16, s->stream);
        return r;
    }
    if (p.square_radius) {
        // the reference's Tier-2 search: final quarter-sample vectors, no 8x8 ones
        int r = bc250_launch_square_me((const uint8_t *)rt.y_plane, rt.y_pitch, ref.y, ref.py, s->w, s->h,
                                       p.square_radius, s->mvs, s->stream);
        mark(s, mid, s->stream);
        if (r == 0) r = bc250_launch_zero_i32((int *)s->mv8, s->nmb * 16, s->stream);
        return r;
    }
    int r = 0;
    const uint32_t *i16 = s->imv16, *i8 = s->imv8;
    if (s->ext16) {
        i16 = s->ext16;
        i8 = s->ext8;
    } else if (s->prop.on) {
        r = enqueue_learned(s, (const uint8_t *)rt.y_plane, rt.y_pitch, ref.y, ref.py, s->imv16, s->imv8, s->stream);
        s->fb_frame = s->prop.count;
    } else {
        r = bc250_launch_me_fullsearch((const uint8_t *)rt.y_plane, rt.y_pitch, ref.y, ref.py, s->wc, s->hc,
                                       p.search_range, p.lambda8, s->imv16, s->imv8, s->stream);
    }
    mark(s, mid, s->stream);
    if (r == 0)
        r = bc250_launch_me_refine((const uint8_t *)rt.y_plane, rt.y_pitch, ref.y, ref.py, s->wc, s->hc, i16, i8,
                                   s->mvs, s->mv8, p.me_8x8, s->stream);
    return r;
}

// HEVC stages of one frame (same events: start, integer search, refine,
// encode, "deblock" (none), copies).  I: CTU wavefront; P: full search,
// quarter-sample 8x8 refine with the HEVC interpolation, CU coding.
static int enqueue_hevc(bc250_hip_state *s, Slot &sl, const Plane2 &cur, const Plane2 &ref, Plane2 &out, int qp,
                        int is_intra, int num_slices)
{
    hipStream_t st = s->stream;
    const gpu_encode_params_t &p = s->params;
    const size_t n = (size_t)s->nmb;
    const int bd = s->bps == 2 ? 10 : 8;
    int r = 0;
    bool packed = false;  // the coding kernel packed the levels (CTU16 P with readback)
    mark(s, sl.ev[0], st);
    if (!is_intra) {
        // the integer search runs on 8-bit luma: at 10 bits the top 8 bits of each sample
        const uint8_t *c8 = cur.y, *r8 = ref.y;
        int cp8 = cur.py, rp8 = ref.py;
        if (bd == 10) {
            uint8_t *m0 = s->me8, *m1 = s->me8 + (size_t)s->me8_pitch * s->hc;
            r = bc250_hevc_p010_to_8bit(cur.y, cur.py, m0, s->me8_pitch, s->wc, s->hc, st);
            if (r == 0) r = bc250_hevc_p010_to_8bit(ref.y, ref.py, m1, s->me8_pitch, s->wc, s->hc, st);
            c8 = m0;
            r8 = m1;
            cp8 = rp8 = s->me8_pitch;
        }
        if (r == 0 && s->ext16) {
            // the integer search ran elsewhere
        } else if (r == 0 && s->prop.on) {
            r = enqueue_learned(s, c8, cp8, r8, rp8, s->imv16, s->imv8, st,
                                !is_intra && p.hevc_ctu_log2 == 4 && p.readback);
            s->fb_frame = s->prop.count;
        } else if (r == 0) {
            r = bc250_launch_me_fullsearch(c8, cp8, r8, rp8, s->wc, s->hc, p.search_range, p.lambda8, s->imv16,
                                           s->imv8, st);
        }
    }
    mark(s, sl.ev[1], st);
    // P pictures: the quarter-sample refinement runs fused with the CU coding
    // below (bc250_hevc_launch_refine_p_encode), so it is timed with the coding
    mark(s, sl.ev[2], st);
    if (r == 0) {
        if (is_intra) {
            r = bc250_hevc_launch_i_encode_ctu(cur.y, cur.uv, cur.py, cur.puv, out.y, out.uv, out.py, out.puv,
                                                   s->wc, s->hc, qp, num_slices,
                                                   4 /* rounding of intra CUs in an I picture */, s->hmodes,
                                                   (int16_t *)s->levels, s->sync_h, kSpinLimit, p.hevc_ctu_log2, bd,
                                                   st);
            if (r == 0) r = bc250_launch_zero_i32((int *)s->mv8, (int)n * 16, st);
        } else {
            // P modes: zero, or at CTU32 the log2 CU size of each CU8 at [16 t + z];
            // at CTU16 the coding kernel also writes the coded-luma flags
            const bool c32 = p.hevc_ctu_log2 == 5;
            if (!c32 && p.readback) {
                // CTU16: the coding kernel packs the levels for the readback itself
                const DevOut &o = sl.dev;
                if (s->fs_keys && !s->ext8)
                    r = bc250_hevc_launch_refine_p_encode_pack_keys(
                        cur.y, cur.uv, cur.py, cur.puv, ref.y, ref.uv, ref.py, ref.puv, out.y, out.uv, out.py, out.puv,
                        s->wc, s->hc, qp, 2 /* inter rounding */, s->imv8, s->mv8, (int16_t *)s->levels, s->hmodes,
                        s->hcbf, bd, o.lvidx, (int16_t *)o.lvdata, o.lvcount, s->fs_keys, s->imv16, p.search_range,
                        p.lambda8, st);
                else
                    r = bc250_hevc_launch_refine_p_encode_pack(
                        cur.y, cur.uv, cur.py, cur.puv, ref.y, ref.uv, ref.py, ref.puv, out.y, out.uv, out.py, out.puv,
                        s->wc, s->hc, qp, 2 /* inter rounding */, s->ext8 ? s->ext8 : s->imv8, s->mv8,
                        (int16_t *)s->levels, s->hmodes, s->hcbf, bd, o.lvidx, (int16_t *)o.lvdata, o.lvcount, st);
                packed = r == 0 && HIP_OK(hipEventRecord(sl.cearly, st));
            } else {
                r = bc250_hevc_launch_refine_p_encode(cur.y, cur.uv, cur.py, cur.puv, ref.y, ref.uv, ref.py, ref.puv,
                                                      out.y, out.uv, out.py, out.puv, s->wc, s->hc, qp,
                                                      2 /* inter rounding */, s->ext8 ? s->ext8 : s->imv8, s->mv8,
                                                      (int16_t *)s->levels,
                                                      p.hevc_ctu_log2, c32 ? s->hculog2 : nullptr, s->hmodes,
                                                      c32 ? nullptr : s->hcbf, bd, st);
            }
        }
    }
    mark(s, sl.ev[3], st);
    // 8.7.2 deblocking (the reference encoder's PPS turns it off: deblock_idc 1);
    // slices never filter across their borders, so idc 0 and 2 agree
    if (r == 0 && p.deblock_idc != 1) {
        const uint8_t *cus = !is_intra && p.hevc_ctu_log2 == 5 ? s->hculog2 : nullptr;
        if (cus) r = bc250_hevc_cu_cbf_ctu((const int16_t *)s->levels, s->wc, s->hc, cus, s->hcbf, st);
        if (r == 0)
            r = bc250_hevc_launch_deblock_ctu(out.y, out.uv, out.py, out.puv, s->wc, s->hc, qp, is_intra, s->hcbf,
                                              s->mv8, cus, num_slices, p.hevc_ctu_log2, bd, st);
    }
    mark(s, sl.ev[4], st);
    if (r == 0)
        r = finish_frame(s, sl, is_intra ? s->sync_h + 1 : nullptr, nullptr, s->fb_frame,
                         RB_LEVELS | RB_HMODES | RB_MV8 | (is_intra ? 0 : RB_PACK) |
                             (packed ? RB_PACKED | RB_HMODES_ZERO : 0) | RB_MV8_HALF);
    return r;
}

extern "C" int gpu_compute_dispatch_encode_ext(gpu_context_t *ctx, gpu_image_t rt, int width, int height, int qp,
                                               int is_intra, int num_slices, int me_mode, const gpu_mv_t *cpu_mvs)
{
    bc250_hip_state *s = st_of(ctx);
    if (!s || !rt.y_plane || qp < 0 || qp > 51) return -1;
    if (num_slices < 1) num_slices = 1;
    const gpu_encode_params_t &p0 = s->params;
    // 10-bit (P010) pictures: the HEVC Main10 pipeline only
    if (rt.format != GPU_IMAGE_NV12 && !(rt.format == GPU_IMAGE_P010 && p0.codec == GPU_CODEC_HEVC)) return -1;
    if (ensure_res(s, width, height, rt.format == GPU_IMAGE_P010 ? 2 : 1) != 0) return -1;
    if (!s->has_recon) {  // new resolution: no reconstruction yet
        ctx->recon_image = gpu_image_t{};
        ctx->recon_memory = gpu_memory_t{};
        ctx->has_recon_frame = false;
    }
    if (rt.coded_width != (uint32_t)s->wc || rt.coded_height != (uint32_t)s->hc) return -1;
    Slot &sl = s->slot[s->cur];
    hipStream_t st = s->stream;
    const gpu_encode_params_t &p = s->params;

    Plane2 cur_in;
    cur_in.y = (uint8_t *)rt.y_plane;
    cur_in.uv = (uint8_t *)rt.uv_plane;
    cur_in.py = rt.y_pitch;
    cur_in.puv = rt.uv_pitch;
    const Plane2 &ref = s->has_recon ? s->recon[s->ref_idx] : cur_in;
    const int out_idx = s->has_recon ? (s->ref_idx ^ 1) : 0;
    Plane2 &out = s->recon[out_idx];

    sl.slices = 0;
    if (p.codec == GPU_CODEC_HEVC) {
        if (p.entropy) {
            fprintf(stderr, "[bc250-hip] entropy = 1 writes H.264 CAVLC slices; HEVC has no slice writer\n");
            return -1;
        }
        if (cpu_mvs) {
            fprintf(stderr, "[bc250-hip] cpu_mvs is an H.264 route; HEVC searches on the GPU\n");
            return -1;
        }
        start_frame(s);
        if (enqueue_hevc(s, sl, cur_in, ref, out, qp, is_intra, num_slices) != 0) {
            fprintf(stderr, "[bc250-hip] HEVC dispatch failed (%dx%d qp %d intra %d)\n", width, height, qp, is_intra);
            return -1;
        }
        sl.have_data = p.readback != 0;
        sl.was_intra = is_intra != 0;
        sl.recorded = false;
        s->ref_idx = out_idx;
        s->has_recon = true;
        publish_recon(ctx, s);
        return 0;
    }

    if (p.entropy && (num_slices > 64 || num_slices > s->nmb)) return -1;
    start_frame(s);
    int r = 0;
    mark(s, sl.ev[0], st);
    if (!is_intra && !cpu_mvs) {
        r = enqueue_me(s, rt, ref, me_mode, sl.ev[1]);
    } else {
        mark(s, sl.ev[1], st);
        if (!is_intra)
            r = HIP_OK(hipMemcpyAsync(s->mvs, cpu_mvs, (size_t)s->nmb * 16, hipMemcpyHostToDevice, st)) ? 0 : -1;
    }
    mark(s, sl.ev[2], st);
    if (r == 0) {
        if (is_intra) {
            r = bc250_launch_i_encode(cur_in.y, cur_in.uv, cur_in.py, cur_in.puv, out.y, out.uv, out.py, out.puv,
                                          s->wc, s->hc, qp, num_slices, s->levels, s->dc, s->nz, s->modes, s->sync_i,
                                          kSpinLimit, st);
            if (r == 0) r = bc250_launch_zero_i32((int *)s->mvs, s->nmb * 4, st);
            if (r == 0) r = bc250_launch_zero_i32((int *)s->mv8, s->nmb * 16, st);
        } else {
            r = bc250_launch_p_encode(cur_in.y, cur_in.uv, cur_in.py, cur_in.puv, ref.y, ref.uv, ref.py, ref.puv,
                                      out.y, out.uv, out.py, out.puv, s->wc, s->hc, qp, s->mvs, s->levels, s->dc,
                                      s->nz, s->modes, st);
        }
    }
    mark(s, sl.ev[3], st);
    if (r == 0 && p.deblock_idc != 1) {
        r = bc250_launch_deblock(out.y, out.uv, out.py, out.puv, s->wc, s->hc, qp, is_intra, s->nz, s->mvs,
                                     num_slices, p.deblock_idc, s->sync_d, kSpinLimit, st);
    }
    mark(s, sl.ev[4], st);
    // frame_num / POC / idr_pic_id of this frame, committed once it is enqueued
    const unsigned fn = is_intra ? 0u : s->frame_num, poc = is_intra ? 0u : s->poc;
    const int idr_id = is_intra ? (s->idr_pic_id + 1) & 0xffff : (s->idr_pic_id < 0 ? 0 : s->idr_pic_id);
    if (r == 0 && p.entropy) {
        // CAVLC slices of the frame, numbered the IPPP way (IDR at intra frames)
        if (ensure_cavlc(s) != 0) {
            r = -1;
        } else {
            bc250_h264_slice_cfg_t c = s->hdr;
            c.width_in_mbs = s->mbw;
            c.height_in_mbs = s->mbh;
            c.qp = qp;
            c.is_idr = is_intra ? 1 : 0;
            c.num_slices = num_slices;
            // written modulo 2^log2_max (encoder_h264.c:3102,3108)
            c.frame_num = (int)(fn & ((1u << (c.log2_max_frame_num_minus4 + 4)) - 1));
            c.idr_pic_id = idr_id;
            c.poc = (int)(poc & ((1u << (c.log2_max_poc_lsb_minus4 + 4)) - 1));
            c.deblock_idc = p.deblock_idc;
            r = bc250_h264_cavlc_slices(s->levels, s->dc, s->modes, (const gpu_mv_t *)s->mvs, &c, s->cavlc_out,
                                        s->cavlc_cap, s->cavlc_info, s->cavlc_work, s->cavlc_work_size, st);
            if (r == 0)
                r = bc250_h264_cavlc_copy_out(s->cavlc_out, s->cavlc_info, num_slices, sl.rbsp, s->cavlc_cap,
                                              sl.rbsp_info, st);
            if (r == 0) sl.slices = num_slices;
        }
    }
    // wait-error flags of the wavefront kernels this frame ran (sync[1]);
    // with deblocking off the deblock workspace is never written
    if (r == 0)
        r = finish_frame(s, sl, is_intra ? s->sync_i + 1 : nullptr, p.deblock_idc != 1 ? s->sync_d + 1 : nullptr,
                         s->fb_frame,
                         RB_LEVELS | RB_DC | RB_MODES | RB_MVS | RB_NZ | RB_MV8 | (is_intra ? 0 : RB_PACK));
    sl.have_data = r == 0 && p.readback;
    sl.was_intra = is_intra != 0;
    sl.recorded = false;
    if (r != 0) {
        fprintf(stderr, "[bc250-hip] dispatch_encode failed (%dx%d qp %d intra %d)\n", width, height, qp, is_intra);
        return -1;
    }
    s->ref_idx = out_idx;
    s->has_recon = true;
    if (is_intra) s->idr_pic_id = idr_id;
    s->frame_num = fn + 1;
    s->poc = poc + 2;
    publish_recon(ctx, s);
    return 0;
}

extern "C" int gpu_compute_dispatch_encode(gpu_context_t *ctx, gpu_image_t rt, int width, int height, int qp,
                                           int is_intra, int num_slices)
{
    return gpu_compute_dispatch_encode_ext(ctx, rt, width, height, qp, is_intra, num_slices, 0, nullptr);
}

// ME alone (gpu_compute.c:3432-3521): MVs land in the current slot's staging.
extern "C" int gpu_compute_dispatch_me_only(gpu_context_t *ctx, gpu_image_t rt, int width, int height)
{
    bc250_hip_state *s = st_of(ctx);
    if (!s || !rt.y_plane || rt.format != GPU_IMAGE_NV12) return -1;
    if (ensure_res(s, width, height, 1) != 0) return -1;
    if (!s->has_recon) {  // new resolution: no reconstruction yet
        ctx->recon_image = gpu_image_t{};
        ctx->recon_memory = gpu_memory_t{};
        ctx->has_recon_frame = false;
    }
    Slot &sl = s->slot[s->cur];
    Plane2 cur_in;
    cur_in.y = (uint8_t *)rt.y_plane;
    cur_in.py = rt.y_pitch;
    const Plane2 &ref = s->has_recon ? s->recon[s->ref_idx] : cur_in;
    start_frame(s);
    mark(s, sl.ev[0], s->stream);
    int r = enqueue_me(s, rt, ref, 0, sl.ev[1]);
    for (int i = 2; i < NEV - 1; i++) mark(s, sl.ev[i], s->stream);
    if (r == 0) r = finish_frame(s, sl, nullptr, nullptr, s->fb_frame, RB_MVS | RB_MV8);
    sl.have_data = false;
    sl.recorded = false;
    sl.slices = 0;
    return r;
}

extern "C" double gpu_compute_get_last_latency_ms(const gpu_context_t *ctx)
{
    return ctx && ctx->hip ? ctx->hip->last_latency_ms : 0.0;
}

extern "C" int gpu_compute_end_picture(gpu_context_t *ctx)
{
    bc250_hip_state *s = st_of(ctx);
    if (!s) return -1;
    Slot &sl = s->slot[s->cur];
    // the copy stream has waited for the frame's stages: its last event ends the frame
    if (!HIP_OK(hipEventRecord(sl.ev[NEV - 1], s->cstream))) return -1;
    sl.pending = true;
    sl.recorded = true;
    s->submitted = s->cur;
    s->cur ^= 1;
    ctx->frames_submitted++;
    return 0;
}

extern "C" int gpu_compute_submitted_slot(gpu_context_t *ctx)
{
    bc250_hip_state *s = st_of(ctx);
    return s ? s->submitted : 0;
}

static int settle_levels(bc250_hip_state *s, Slot &sl);

extern "C" int gpu_compute_sync_slot(gpu_context_t *ctx, int slot)
{
    bc250_hip_state *s = st_of(ctx);
    if (!s || slot < 0 || slot >= NSLOT) return -1;
    Slot &sl = s->slot[slot];
    if (!sl.recorded) return -1;
    if (!HIP_OK(wait_event(sl.ev[NEV - 1]))) {
        sl.have_data = false;
        return -1;
    }
    sl.pending = false;
    if (settle_levels(s, sl) != 0) {
        sl.have_data = false;
        return -1;
    }
    float ms = 0.f;
    if (sl.timed && HIP_OK(hipEventElapsedTime(&ms, sl.ev[0], sl.ev[NEV - 1]))) {
        s->last_latency_ms = ms;
        s->total_sum += ms;
        for (int i = 0; i < NSTAGE; i++) {
            float t = 0.f;
            if (HIP_OK(hipEventElapsedTime(&t, sl.ev[i], sl.ev[i + 1]))) {
                s->stage_sum[i] += t;
                if (i == 0 && !sl.was_intra) s->me_int_sum_p += t;
            }
        }
        s->stage_frames++;
        if (!sl.was_intra) s->stage_p_frames++;
    }
    if (sl.err && (sl.err[0] | sl.err[1])) {
        s->err_flags |= (sl.err[0] ? 1 : 0) | (sl.err[1] ? 2 : 0);
        fprintf(stderr, "[bc250-hip] wavefront dependency wait timed out (slot %d)\n", slot);
        sl.have_data = false;
        return -1;
    }
    return 0;
}

extern "C" int gpu_compute_sync(gpu_context_t *ctx)
{
    return gpu_compute_sync_slot(ctx, gpu_compute_submitted_slot(ctx));
}

// A finished frame whose levels arrived packed: its count is checked.
static int settle_levels(bc250_hip_state *s, Slot &sl)
{
    if (!sl.lv_packed || sl.lv_settled) return 0;
    const int cnt = *sl.lvcount;
    if (cnt < 0 || cnt > s->nmb * 24) return -1;
    sl.lv_settled = true;
    return 0;
}

// The full levels of a slot whose frame arrived packed: the groups the
// array held nonzero before are cleared, the packed groups written.
static void build_levels(bc250_hip_state *s, Slot &sl)
{
    if (settle_levels(s, sl) != 0 || !sl.lv_packed || sl.lv_built) return;
    const size_t ng = (size_t)s->nmb * 24;
    if (sl.built_all) {
        memset(sl.levels, 0, ng * 32);
    } else {
        for (int i = 0; i < sl.built_n; i++) memset(sl.levels + (size_t)sl.built_idx[i] * 16, 0, 32);
    }
    const int cnt = *sl.lvcount;
    int kept = 0;
    for (int k = 0; k < cnt; k++) {
        const uint32_t g = sl.lvidx[k];
        if (g >= ng) continue;
        memcpy(sl.levels + (size_t)g * 16, sl.lvdata + (size_t)k * 16, 32);
        sl.built_idx[kept++] = g;
    }
    sl.built_n = kept;
    sl.built_all = false;
    sl.lv_built = true;
}

static int staging(gpu_context_t *ctx, int slot, int which, void **data, size_t *size)
{
    bc250_hip_state *s = st_of(ctx);
    if (!s || !data || slot < 0 || slot >= NSLOT) return -1;
    Slot &sl = s->slot[slot];
    const size_t n = (size_t)s->nmb;
    bool mv_ok = which == 3 || which == 5;  // MVs are valid after dispatch_me_only too
    if (!sl.levels || (!sl.have_data && !mv_ok)) {
        *data = nullptr;
        if (size) *size = 0;
        return -1;
    }
    switch (which) {
    case 0:
        build_levels(s, sl);
        *data = sl.levels;
        if (size) *size = n * 384 * 2;
        break;
    case 1: *data = sl.dc; if (size) *size = n * 24 * 4; break;
    case 2: *data = sl.modes; if (size) *size = n * 4; break;
    case 3: *data = sl.mvs; if (size) *size = n * 16; break;
    case 4: *data = sl.nz; if (size) *size = n * 24 * 4; break;
    case 5:
        if (sl.mv8_half) {
            // widen in place from the last record down: record i's 16 bytes
            // cover half records 2i and 2i + 1, read already (or now, at i = 0)
            const uint2 *h = (const uint2 *)sl.mv8;
            for (size_t i = n * 4; i-- > 0;) {
                const uint2 v = h[i];
                sl.mv8[i] = gpu_mv_t{(int16_t)(v.x & 0xffffu), (int16_t)(v.x >> 16), v.y, 0u};
            }
            sl.mv8_half = false;
        }
        *data = sl.mv8;
        if (size) *size = n * 64;
        break;
    default:
        if (sl.hmodes_zero) {
            memset(sl.hmodes, 0, n * 16);
            sl.hmodes_zero = false;
        }
        *data = sl.hmodes;
        if (size) *size = n * 16;
        break;
    }
    return 0;
}

extern "C" int gpu_compute_get_hevc_mode_staging_data_slot(gpu_context_t *c, int slot, void **d, size_t *z)
{
    return staging(c, slot, 6, d, z);
}
extern "C" int gpu_compute_get_quant_staging_data_slot(gpu_context_t *c, int slot, void **d, size_t *z) { return staging(c, slot, 0, d, z); }
extern "C" int gpu_compute_get_dc_staging_data_slot(gpu_context_t *c, int slot, void **d, size_t *z) { return staging(c, slot, 1, d, z); }
extern "C" int gpu_compute_get_pred_mode_staging_data_slot(gpu_context_t *c, int slot, void **d, size_t *z) { return staging(c, slot, 2, d, z); }
extern "C" int gpu_compute_get_mv_staging_data_slot(gpu_context_t *c, int slot, void **d, size_t *z) { return staging(c, slot, 3, d, z); }
extern "C" int gpu_compute_get_nz_staging_data_slot(gpu_context_t *c, int slot, void **d, size_t *z) { return staging(c, slot, 4, d, z); }
extern "C" int gpu_compute_get_mv8_staging_data_slot(gpu_context_t *c, int slot, void **d, size_t *z) { return staging(c, slot, 5, d, z); }

extern "C" int gpu_compute_get_quant_staging_data(gpu_context_t *c, void **d, size_t *z) { return staging(c, gpu_compute_submitted_slot(c), 0, d, z); }
extern "C" int gpu_compute_get_dc_staging_data(gpu_context_t *c, void **d, size_t *z) { return staging(c, gpu_compute_submitted_slot(c), 1, d, z); }
extern "C" int gpu_compute_get_pred_mode_staging_data(gpu_context_t *c, void **d, size_t *z) { return staging(c, gpu_compute_submitted_slot(c), 2, d, z); }
extern "C" int gpu_compute_get_mv_staging_data(gpu_context_t *c, void **d, size_t *z) { return staging(c, gpu_compute_submitted_slot(c), 3, d, z); }
extern "C" int gpu_compute_get_nz_staging_data(gpu_context_t *c, void **d, size_t *z) { return staging(c, gpu_compute_submitted_slot(c), 4, d, z); }
// The reference's generic staging getter returns the levels of the last frame.
extern "C" int gpu_compute_get_staging_data(gpu_context_t *c, void **d, size_t *z) { return staging(c, gpu_compute_submitted_slot(c), 0, d, z); }
extern "C" int gpu_compute_release_staging_data(gpu_context_t *ctx) { return ctx ? 0 : -1; }

extern "C" FILE *bc250_debug_dump_open(const char *name, const char *what)
{
    if (!name || !*name || strchr(name, '/')) return nullptr;
    char dir[4096];
    const char *d = getenv("BC250_DUMP_DIR");
    if (d && *d) {
        snprintf(dir, sizeof dir, "%s", d);
    } else {
        const char *rt = getenv("XDG_RUNTIME_DIR");
        if (!rt || !*rt) {
            fprintf(stderr, "[bc250-hip] %s: no BC250_DUMP_DIR or XDG_RUNTIME_DIR, not dumping\n",
                    what ? what : "dump");
            return nullptr;
        }
        snprintf(dir, sizeof dir, "%s/bc250_dump_frames", rt);
        mkdir(dir, 0700);
    }
    char path[4352];
    snprintf(path, sizeof path, "%s/%s", dir, name);
    const int fd = open(path, O_WRONLY | O_CREAT | O_TRUNC | O_NOFOLLOW | O_CLOEXEC, 0600);
    if (fd < 0) {
        fprintf(stderr, "[bc250-hip] %s: cannot create %s: %s\n", what ? what : "dump", path, strerror(errno));
        return nullptr;
    }
    FILE *f = fdopen(fd, "wb");
    if (!f) close(fd);
    return f;
}

extern "C" void bc250_debug_dump_nv12_frame(const uint8_t *, int, const uint8_t *, int, int, int) {}
extern "C" void gpu_compute_debug_dump_recon(gpu_context_t *, int, int) {}
extern "C" void gpu_compute_debug_dump_real_input(gpu_context_t *, gpu_image_t *, gpu_memory_t, int, int) {}

// ------------------------------------------------------------ extensions

extern "C" int gpu_compute_set_encode_params(gpu_context_t *ctx, const gpu_encode_params_t *p)
{
    bc250_hip_state *s = st_of(ctx);
    if (!s || !p) return -1;
    if (p->search_range != 16 && p->search_range != 32 && p->search_range != 64) return -1;
    if (p->lambda8 < 0 || p->lambda8 > 15 || p->deblock_idc < 0 || p->deblock_idc > 2) return -1;
    if (p->codec != GPU_CODEC_H264 && p->codec != GPU_CODEC_HEVC) return -1;
    if (p->square_radius != 0 && (p->square_radius < 2 || p->square_radius > 16)) return -1;
    if (p->diamond != 0 && p->diamond != 1) return -1;
    if (p->diamond && p->square_radius) return -1;  // one search at a time
    if (p->hevc_ctu_log2 != 0 && p->hevc_ctu_log2 != 4 && p->hevc_ctu_log2 != 5) return -1;
    if (p->entropy != 0 && p->entropy != 1) return -1;
    s->params = *p;
    if (!s->params.hevc_ctu_log2) s->params.hevc_ctu_log2 = 4;
    s->params.me_8x8 = p->me_8x8 ? 1 : 0;
    s->params.readback = p->readback ? 1 : 0;
    return 0;
}

extern "C" int gpu_compute_set_h264_slice_header(gpu_context_t *ctx, const bc250_h264_slice_cfg_t *cfg)
{
    bc250_hip_state *s = st_of(ctx);
    if (!s || !cfg) return -1;
    if (cfg->pps_id < 0 || cfg->pps_id > 255 || cfg->pic_init_qp_minus26 < -26 || cfg->pic_init_qp_minus26 > 25 ||
        cfg->log2_max_frame_num_minus4 < 0 || cfg->log2_max_frame_num_minus4 > 12 ||
        cfg->log2_max_poc_lsb_minus4 < 0 || cfg->log2_max_poc_lsb_minus4 > 12 ||
        (cfg->relax_chroma_skip != 0 && cfg->relax_chroma_skip != 1))
        return -1;
    s->hdr.pps_id = cfg->pps_id;
    s->hdr.pic_init_qp_minus26 = cfg->pic_init_qp_minus26;
    s->hdr.log2_max_frame_num_minus4 = cfg->log2_max_frame_num_minus4;
    s->hdr.log2_max_poc_lsb_minus4 = cfg->log2_max_poc_lsb_minus4;
    s->hdr.relax_chroma_skip = cfg->relax_chroma_skip;
    return 0;
}

extern "C" int gpu_compute_get_slices_slot(gpu_context_t *ctx, int slot, const uint8_t **rbsp, const uint32_t **info,
                                           int *num_slices)
{
    bc250_hip_state *s = st_of(ctx);
    if (!s || !rbsp || !info || !num_slices || slot < 0 || slot >= NSLOT) return -1;
    const Slot &sl = s->slot[slot];
    if (!sl.slices || sl.pending || !sl.rbsp) return -1;
    // a wavefront wait that gave up leaves levels built from stale neighbours
    if (sl.err && (sl.err[0] | sl.err[1])) return -1;
    if (sl.rbsp_info[4 * sl.slices + 1] || sl.rbsp_info[4 * sl.slices] > s->cavlc_cap) return -1;
    *rbsp = sl.rbsp;
    *info = sl.rbsp_info;
    *num_slices = sl.slices;
    return 0;
}

extern "C" int gpu_compute_set_me_proposer(gpu_context_t *ctx, const float *params, const bc250_proposer_cfg_t *cfg)
{
    bc250_hip_state *s = st_of(ctx);
    if (!s) return -1;
    Proposer &q = s->prop;
    if (!params || !cfg) {
        q.on = false;
        return 0;
    }
    if (cfg->ncand < 1 || cfg->ncand > 4 || cfg->alpha_q4 < 0 || cfg->beta < 0 || cfg->t_high < 0) return -1;
    if (!q.params && (!HIP_OK(hipMalloc(&q.params, BC250_CNN_NPARAM * sizeof(float))) ||
                      !HIP_OK(hipMalloc(&q.packed, BC250_CNN_NPACKED * sizeof(uint16_t)))))
        return -1;
    if (!HIP_OK(hipStreamSynchronize(s->stream)) ||
        !HIP_OK(hipMemcpy(q.params, params, BC250_CNN_NPARAM * sizeof(float), hipMemcpyHostToDevice)) ||
        bc250_cnn_pack(q.params, q.packed, s->stream) != 0 || !HIP_OK(hipStreamSynchronize(s->stream)))
        return -1;
    q.cfg = *cfg;
    q.on = true;
    return 0;
}

extern "C" int gpu_compute_set_external_int_me(gpu_context_t *ctx, const uint32_t *imv16, const uint32_t *imv8)
{
    bc250_hip_state *s = st_of(ctx);
    if (!s || (imv16 == nullptr) != (imv8 == nullptr)) return -1;
    s->ext16 = imv16;
    s->ext8 = imv8;
    return 0;
}

extern "C" int gpu_compute_get_reference_luma(gpu_context_t *ctx, const uint8_t **y, int *pitch)
{
    bc250_hip_state *s = st_of(ctx);
    if (!s || !y || !pitch || !s->has_recon) return -1;
    *y = s->bps == 2 ? nullptr : s->recon[s->ref_idx].y;
    *pitch = s->recon[s->ref_idx].py;
    return *y ? 0 : -1;
}

extern "C" int gpu_compute_stream_wait_event(gpu_context_t *ctx, void *event)
{
    bc250_hip_state *s = st_of(ctx);
    if (!s || !event) return -1;
    return HIP_OK(hipStreamWaitEvent(s->stream, (hipEvent_t)event, 0)) ? 0 : -1;
}

extern "C" int gpu_compute_slot_done(gpu_context_t *ctx, int slot)
{
    bc250_hip_state *s = st_of(ctx);
    if (!s || slot < 0 || slot >= NSLOT || !s->slot[slot].recorded) return -1;
    const hipError_t e = hipEventQuery(s->slot[slot].ev[NEV - 1]);
    return e == hipSuccess ? 1 : (e == hipErrorNotReady ? 0 : -1);
}

extern "C" int gpu_compute_get_me_fallbacks(gpu_context_t *ctx, int slot)
{
    bc250_hip_state *s = st_of(ctx);
    if (!s || slot < 0 || slot >= NSLOT) return -1;
    const Slot &sl = s->slot[slot];
    if (!sl.err || sl.pending || !sl.recorded) return -1;
    return sl.err[2];
}

extern "C" int gpu_compute_get_encode_params(gpu_context_t *ctx, gpu_encode_params_t *p)
{
    bc250_hip_state *s = st_of(ctx);
    if (!s || !p) return -1;
    *p = s->params;
    return 0;
}

extern "C" int gpu_compute_download_recon(gpu_context_t *ctx, uint8_t *y_plane, int y_pitch, uint8_t *uv_plane,
                                          int uv_pitch)
{
    bc250_hip_state *s = st_of(ctx);
    if (!s || !s->has_recon || !y_plane || !uv_plane) return -1;
    const Plane2 &r = s->recon[s->ref_idx];
    return download_planes(s, r.y, r.py, r.uv, r.puv, y_plane, y_pitch, uv_plane, uv_pitch, (size_t)s->w * s->bps,
                           (size_t)((s->w + 1) / 2) * 2 * s->bps, s->h);
}

extern "C" int gpu_compute_get_stage_times(gpu_context_t *ctx, int slot, float ms[6])
{
    bc250_hip_state *s = st_of(ctx);
    if (!s || !ms || slot < 0 || slot >= NSLOT || !s->slot[slot].recorded || !s->slot[slot].timed) return -1;
    Slot &sl = s->slot[slot];
    if (!HIP_OK(hipEventSynchronize(sl.ev[NEV - 1]))) return -1;
    for (int i = 0; i < NSTAGE; i++)
        if (!HIP_OK(hipEventElapsedTime(&ms[i], sl.ev[i], sl.ev[i + 1]))) return -1;
    return HIP_OK(hipEventElapsedTime(&ms[NSTAGE], sl.ev[0], sl.ev[NEV - 1])) ? 0 : -1;
}

extern "C" int gpu_compute_set_stage_timing(gpu_context_t *ctx, int on)
{
    bc250_hip_state *s = st_of(ctx);
    if (!s) return -1;
    s->timing = on != 0;
    return 0;
}

extern "C" int gpu_compute_get_stage_stats(gpu_context_t *ctx, gpu_stage_stats_t *st, int reset)
{
    bc250_hip_state *s = st_of(ctx);
    if (!s || !st) return -1;
    for (int i = 0; i < NSTAGE; i++) st->sum_ms[i] = s->stage_sum[i];
    st->total_ms = s->total_sum;
    st->frames = s->stage_frames;
    st->p_frames = s->stage_p_frames;
    st->me_int_ms_p = s->me_int_sum_p;
    if (reset) {
        for (int i = 0; i < NSTAGE; i++) s->stage_sum[i] = 0;
        s->total_sum = s->me_int_sum_p = 0;
        s->stage_frames = s->stage_p_frames = 0;
    }
    return 0;
}

extern "C" int gpu_compute_take_error_flags(gpu_context_t *ctx)
{
    bc250_hip_state *s = st_of(ctx);
    if (!s) return -1;
    int f = s->err_flags;
    s->err_flags = 0;
    return f;
}
