/* ref_hevc_decode.c - test infrastructure: drives the reference's own H.265
 * decoding functions (src/hevc_ps.c parameter sets and slice headers,
 * src/hevc_cu.c hevcd_read_ctu, src/hevc_filter.c hevcd_loop_filters, with
 * their residual, motion, prediction and transform files) over an Annex B
 * stream, so that a stream's decode can be compared with the reconstruction
 * the encoder kept.  The reference's top level (src/decoder_h265.c) cannot
 * be built here (it includes gpu_compute.h, hence Vulkan); this harness does
 * the part of it a single-reference IPPP stream needs - picture buffers, the
 * decoder's per-picture maps, the L0 list of one picture, the serial slice
 * walk - in the order that file gives (open_picture, prepare_picture,
 * build_lists, walk_slice, hevc_decoder_end_frame), and nothing else:
 * no tiles, no wavefront entry points, no long-term pictures. */
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

#include "hevc_dec_internal.h"
#include "hevc_ps.h"

#define SLOTS 3

typedef struct {
    hevcd_t d;
    hevcd_img_t buf[SLOTS];
    hevcd_progress_t progress;
    hevc_sps_t sps[16];
    hevc_pps_t pps[64];
    hevc_slice_t sl;
    int open;
    int prev_poc;
} dec_t;

/* EBSP -> RBSP (7.4.2): drop each 0x03 that follows two zero bytes. */
static size_t unescape(const uint8_t *in, size_t n, uint8_t *out)
{
    size_t o = 0;
    int zeros = 0;
    for (size_t i = 0; i < n; i++) {
        if (zeros >= 2 && in[i] == 3) {
            zeros = 0;
            continue;
        }
        out[o++] = in[i];
        zeros = in[i] == 0 ? zeros + 1 : 0;
    }
    return o;
}

static hevcd_img_t *find_poc(dec_t *x, int poc)
{
    for (int i = 0; i < SLOTS; i++)
        if (x->buf[i].is_valid && x->buf[i].poc == poc) return &x->buf[i];
    return NULL;
}

static int open_picture(dec_t *x, const hevc_sps_t *sps, int poc)
{
    hevcd_t *d = &x->d;
    const int w = sps->width, h = sps->height;
    const size_t bytes = sps->bit_depth_luma > 8 ? 2 : 1;
    hevcd_img_t *g = NULL;
    for (int i = 0; i < SLOTS && !g; i++)
        if (!x->buf[i].is_valid) g = &x->buf[i];
    if (!g) return -1;
    if (g->n_planes != (size_t)w * h * bytes) {
        for (int i = 0; i < 3; i++) free(g->plane[i]);
        g->plane[0] = malloc((size_t)w * h * bytes);
        g->plane[1] = malloc((size_t)(w / 2) * (h / 2) * bytes);
        g->plane[2] = malloc((size_t)(w / 2) * (h / 2) * bytes);
        g->n_planes = (size_t)w * h * bytes;
    }
    const size_t n_mvf = (size_t)(w >> 2) * (h >> 2);
    if (g->n_mvf != n_mvf) {
        free(g->mvf);
        g->mvf = calloc(n_mvf, sizeof *g->mvf);
        g->n_mvf = n_mvf;
    }
    const size_t ctbs = (size_t)sps->ctb_count;
    if (g->n_lists < ctbs) {
        free(g->lists);
        g->lists = calloc(ctbs, sizeof *g->lists);
        g->n_lists = ctbs;
    }
    if (g->n_slice_map < ctbs) {
        free(g->slice_of_ctb);
        g->slice_of_ctb = malloc(ctbs * sizeof *g->slice_of_ctb);
        g->n_slice_map = ctbs;
    }
    if (!g->plane[0] || !g->plane[1] || !g->plane[2] || !g->mvf || !g->lists || !g->slice_of_ctb) return -1;
    g->stride[0] = w;
    g->stride[1] = g->stride[2] = w / 2;
    g->poc = poc;
    g->is_valid = true;
    g->progress = &x->progress;
    atomic_store(&g->rows_ready, 0);
    d->current = g;
    d->slice_of_ctb = g->slice_of_ctb;
    d->n_slice_map = g->n_slice_map;
    d->mvf = g->mvf;
    for (int i = 0; i < 3; i++) {
        d->plane[i] = g->plane[i];
        d->stride[i] = g->stride[i];
    }
    d->n_planes = g->n_planes;
    return 0;
}

#define GROW(ptr, n, want, size)                                                                                    \
    do {                                                                                                            \
        if (!(ptr) || (n) < (want)) {                                                                               \
            free(ptr);                                                                                              \
            (ptr) = calloc((want), (size));                                                                         \
            (n) = (want);                                                                                           \
        }                                                                                                           \
        if (!(ptr)) return -1;                                                                                      \
    } while (0)

static int prepare_picture(dec_t *x, const hevc_sps_t *sps, const hevc_pps_t *pps)
{
    hevcd_t *d = &x->d;
    const size_t cb = (size_t)sps->min_cb_width * sps->min_cb_height;
    const size_t pu = (size_t)(sps->width >> 2) * (sps->height >> 2);
    const size_t edges = (size_t)((sps->width + 7) >> 3) * (size_t)((sps->height + 7) >> 3);
    GROW(d->ct_depth, d->n_ct_depth, cb, 1);
    GROW(d->intra_mode, d->n_intra_mode, pu, 1);
    GROW(d->edges, d->n_edges, edges, 1);
    GROW(d->no_filter, d->n_no_filter, cb, 1);
    GROW(d->sao, d->n_sao, (size_t)sps->ctb_count, sizeof *d->sao);
    GROW(d->cbf_map, d->n_cbf, pu, 1);
    GROW(d->skip, d->n_skip, cb, 1);
    GROW(d->qp_y_map, d->n_qp, cb, 1);
    memset(d->cbf_map, 0, pu);
    memset(d->ct_depth, 0, cb);
    memset(d->edges, 0, edges);
    memset(d->no_filter, 0, cb);
    memset(d->skip, 0, cb);
    memset(d->intra_mode, HEVCD_INTRA_DC, pu);
    memset(d->sao, 0, (size_t)sps->ctb_count * sizeof *d->sao);
    d->edges_stride = (sps->width + 7) >> 3;
    d->sps = sps;
    d->pps = pps;
    d->scaling_on = sps->scaling_list_enabled;
    if (d->scaling_on) return -2; /* the streams checked here send no matrices */
    if (hevcd_prepare_tiles(d) || hevcd_prepare_zscan(d)) return -1;
    for (int i = 0; i < sps->ctb_count; i++) d->slice_of_ctb[i] = -1;
    d->slice_now = -1;
    d->filters_done = false;
    d->have_segment_end = false;
    d->have_wpp_snapshot = false;
    return 0;
}

/* The end of a picture: units no slice reached lose their motion, the loop
 * filters run, every row is final. */
static void end_picture(dec_t *x)
{
    hevcd_t *d = &x->d;
    if (!x->open) return;
    const hevc_sps_t *sps = d->sps;
    for (int rs = 0; rs < sps->ctb_count; rs++)
        if (d->slice_of_ctb[rs] < 0) hevcd_clear_ctb_motion(d, rs % sps->ctb_width, rs / sps->ctb_width);
    if (d->slice) hevcd_loop_filters(d);
    hevcd_rows_ready(d->current, sps->ctb_height);
    x->open = 0;
}

/* One slice segment (independent; one reference picture, the one before). */
static int walk_slice(dec_t *x, const uint8_t *rbsp, size_t n)
{
    hevcd_t *d = &x->d;
    hevc_slice_t *sl = &x->sl;
    const hevc_sps_t *sps = d->sps;
    sl->poc = d->current->poc;
    d->slice = sl;
    d->slice_now++;
    /* L0: the picture POC - 1 (the only short-term set these streams use) */
    d->n_refs[0] = d->n_refs[1] = 0;
    d->col = NULL;
    if (sl->type != 2) {
        const hevcd_img_t *r = find_poc(x, sl->poc - 1);
        if (!r) return -8;
        for (int i = 0; i < sl->num_ref_idx[0] && i < 16; i++) {
            d->ref_pic[0][i] = r;
            d->ref_is_lt[0][i] = false;
        }
        d->n_refs[0] = sl->num_ref_idx[0];
    }
    hevcd_img_t *g = d->current;
    if ((size_t)d->slice_now < g->n_lists) {
        memset(&g->lists[d->slice_now], 0, sizeof g->lists[0]);
        g->lists[d->slice_now].n_list[0] = d->n_refs[0];
        for (int i = 0; i < d->n_refs[0]; i++) g->lists[d->slice_now].poc_list[0][i] = d->ref_pic[0][i]->poc;
    }
    if ((size_t)d->slice_now >= d->n_slice_filter) {
        const size_t want = (size_t)d->slice_now + 64;
        hevcd_slice_filter_t *p = realloc(d->slice_filter, want * sizeof *p);
        if (!p) return -1;
        d->slice_filter = p;
        d->n_slice_filter = want;
    }
    hevcd_slice_filter_t *f = &d->slice_filter[d->slice_now];
    f->beta_offset = (int16_t)sl->beta_offset;
    f->tc_offset = (int16_t)sl->tc_offset;
    f->disabled = sl->deblocking_filter_disabled ? 1 : 0;
    f->across_slices = sl->loop_filter_across_slices ? 1 : 0;
    d->min_pu_width = sps->width >> 2;
    d->min_pu_height = sps->height >> 2;
    d->qp_y = d->qp_y_pred = d->qp_y_prev = sl->qp;
    d->qg_restarts = true;
    d->slice_end = false;
    const size_t first = sl->data_bit_offset >> 3;
    if (first >= n) return -3;
    hevcd_cabac_init(&d->cabac, rbsp + first, n - first, sl->type, sl->cabac_init_flag, sl->qp);
    for (int rs = sl->segment_address; rs < sps->ctb_count; rs++) {
        const int xx = (rs % sps->ctb_width) << sps->log2_ctb, yy = (rs / sps->ctb_width) << sps->log2_ctb;
        if (hevcd_read_ctu(d, xx, yy)) return -4;
        if (hevcd_overrun(&d->cabac)) return -3;
        if (hevcd_terminate(&d->cabac)) return 0;
    }
    return -2; /* ran out of units before end_of_slice_segment_flag */
}

/* Decodes an Annex B stream of VPS / SPS / PPS and slice NAL units (IDR,
 * then pictures each predicting from the one before).  Each decoded picture
 * is appended to out as its Y, Cb and Cr planes at the coded size (uint8 at
 * 8 bits, uint16 above).  Returns the number of pictures, or a negative
 * code: -1 memory, -2 unsupported stream, -3 data overrun, -4 a unit
 * refused, -5 out too small, -6 header refused, -8 missing reference. */
int ref_hevc_decode(const uint8_t *s, size_t n, uint8_t *out, size_t cap, int *width, int *height, int *bit_depth)
{
    dec_t *x = calloc(1, sizeof *x);
    uint8_t *rbsp = malloc(n + 16);
    if (!x || !rbsp) {
        free(x);
        free(rbsp);
        return -1;
    }
    pthread_mutex_init(&x->progress.m, NULL);
    pthread_cond_init(&x->progress.cv, NULL);
    x->d.buf = x->buf;
    x->d.n_buf = SLOTS;
    x->d.progress = &x->progress;
    int npics = 0, rc = 0;
    size_t used = 0, i = 0;
    while (rc >= 0) {
        /* next start code */
        while (i + 3 <= n && !(s[i] == 0 && s[i + 1] == 0 && s[i + 2] == 1)) i++;
        if (i + 3 > n) break;
        const size_t a = i + 3;
        size_t b = a;
        while (b + 3 <= n && !(s[b] == 0 && s[b + 1] == 0 && (s[b + 2] == 1 || s[b + 2] == 0))) b++;
        if (b + 3 > n) b = n;
        i = b;
        const size_t m = unescape(s + a, b - a, rbsp);
        if (m < 2) continue;
        const int type = (rbsp[0] >> 1) & 63;
        if (type == HEVC_NAL_SPS) {
            hevc_sps_t t;
            if (hevc_ps_read_sps(&t, rbsp, m) || t.sps_id < 0 || t.sps_id > 15) rc = -6;
            else x->sps[t.sps_id] = t;
        } else if (type == HEVC_NAL_PPS) {
            hevc_pps_t t;
            if (hevc_ps_read_pps(&t, rbsp, m) || t.pps_id < 0 || t.pps_id > 63) rc = -6;
            else x->pps[t.pps_id] = t;
        } else if (type <= HEVC_NAL_CRA && hevc_nal_e_slice(type)) {
            hevc_slice_t sl;
            if (hevc_ps_read_slice(&sl, rbsp, m, type, x->sps, x->pps)) {
                rc = -6;
                break;
            }
            if (sl.dependent_slice_segment) {
                rc = -2;
                break;
            }
            const hevc_pps_t *pps = &x->pps[sl.pps_id];
            const hevc_sps_t *sps = &x->sps[pps->sps_id];
            if (sl.first_slice_in_pic) {
                if (x->open) {
                    end_picture(x);
                    /* hand the finished picture out */
                    const hevcd_img_t *g = x->d.current;
                    const size_t bytes = x->d.sps->bit_depth_luma > 8 ? 2 : 1;
                    const size_t ny = (size_t)x->d.sps->width * x->d.sps->height * bytes, nc = ny / 4;
                    if (used + ny + 2 * nc > cap) {
                        rc = -5;
                        break;
                    }
                    memcpy(out + used, g->plane[0], ny);
                    memcpy(out + used + ny, g->plane[1], nc);
                    memcpy(out + used + ny + nc, g->plane[2], nc);
                    used += ny + 2 * nc;
                    npics++;
                }
                /* POC (8.3.1) and the pictures the set keeps: the one before */
                int poc = 0;
                if (!hevc_nal_e_idr(type)) {
                    const int maxl = 1 << sps->log2_max_poc_lsb, prev_lsb = x->prev_poc & (maxl - 1);
                    int msb = x->prev_poc - prev_lsb;
                    if (sl.poc_lsb < prev_lsb && prev_lsb - sl.poc_lsb >= maxl / 2) msb += maxl;
                    else if (sl.poc_lsb > prev_lsb && sl.poc_lsb - prev_lsb > maxl / 2) msb -= maxl;
                    poc = msb + sl.poc_lsb;
                }
                for (int k = 0; k < SLOTS; k++)
                    if (hevc_nal_e_idr(type) || x->buf[k].poc != poc - 1) x->buf[k].is_valid = false;
                x->prev_poc = poc;
                if (open_picture(x, sps, poc) || prepare_picture(x, sps, pps)) {
                    rc = -1;
                    break;
                }
                x->open = 1;
            } else if (!x->open) {
                rc = -2;
                break;
            }
            x->sl = sl;
            rc = walk_slice(x, rbsp, m);
        }
    }
    if (rc >= 0 && x->open) {
        end_picture(x);
        const hevcd_img_t *g = x->d.current;
        const size_t bytes = x->d.sps->bit_depth_luma > 8 ? 2 : 1;
        const size_t ny = (size_t)x->d.sps->width * x->d.sps->height * bytes, nc = ny / 4;
        if (used + ny + 2 * nc > cap) {
            rc = -5;
        } else {
            memcpy(out + used, g->plane[0], ny);
            memcpy(out + used + ny, g->plane[1], nc);
            memcpy(out + used + ny + nc, g->plane[2], nc);
            npics++;
        }
    }
    if (x->d.sps) {
        if (width) *width = x->d.sps->width;
        if (height) *height = x->d.sps->height;
        if (bit_depth) *bit_depth = x->d.sps->bit_depth_luma;
    }
    if (x->d.pool) hevcd_pool_destroy(x->d.pool);
    hevcd_free_filters(&x->d);
    for (int k = 0; k < SLOTS; k++) {
        for (int p = 0; p < 3; p++) free(x->buf[k].plane[p]);
        free(x->buf[k].mvf);
        free(x->buf[k].lists);
        free(x->buf[k].slice_of_ctb);
    }
    free(x->d.ct_depth); free(x->d.intra_mode); free(x->d.edges); free(x->d.no_filter); free(x->d.sao);
    free(x->d.cbf_map); free(x->d.skip); free(x->d.qp_y_map); free(x->d.slice_filter);
    free(x->d.min_tb_addr_zs); free(x->d.zs_rs_to_ts); free(x->d.rs_to_ts); free(x->d.ts_to_rs);
    free(x->d.tile_of_ts);
    free(x);
    free(rbsp);
    return rc < 0 ? rc : npics;
}

/* The reference's parameter-set parser on one SPS / PPS NAL unit (without
 * start code, with emulation prevention): the fields the tests compare. */
int ref_hevc_read_sps(const uint8_t *nal, size_t n, int32_t *f)
{
    uint8_t *r = malloc(n);
    if (!r) return -1;
    hevc_sps_t s;
    const int rc = hevc_ps_read_sps(&s, r, unescape(nal, n, r));
    free(r);
    if (rc) return rc;
    const int32_t v[] = {s.width, s.height, s.crop_left, s.crop_right, s.crop_top, s.crop_bottom, s.bit_depth_luma,
                         s.bit_depth_chroma, s.log2_max_poc_lsb, s.log2_min_cb, s.log2_ctb, s.log2_min_tb, s.log2_max_tb,
                         s.max_transform_hierarchy_depth_inter, s.max_transform_hierarchy_depth_intra,
                         s.num_st_rps, s.temporal_mvp_enabled, s.amp_enabled, s.sao_enabled, s.chroma_format_idc};
    memcpy(f, v, sizeof v);
    return 0;
}

int ref_hevc_read_pps(const uint8_t *nal, size_t n, int32_t *f)
{
    uint8_t *r = malloc(n);
    if (!r) return -1;
    hevc_pps_t p;
    const int rc = hevc_ps_read_pps(&p, r, unescape(nal, n, r));
    free(r);
    if (rc) return rc;
    const int32_t v[] = {p.init_qp, p.deblocking_filter_disabled, p.loop_filter_across_slices, p.cu_qp_delta_enabled,
                         p.sign_data_hiding, p.entropy_coding_sync_enabled, p.tiles_enabled,
                         p.num_ref_idx_default[0], p.transform_skip_enabled, p.dependent_slice_segments_enabled};
    memcpy(f, v, sizeof v);
    return 0;
}
