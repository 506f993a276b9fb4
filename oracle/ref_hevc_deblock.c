/* ref_hevc_deblock.c - test harness (not the oracle, not the product): runs
 * the reference decoder's own HEVC deblocking filter (hevcd_filter_stage,
 * approach1-compute-encoder/src/hevc_filter.c:507, over
 * hevc_filter_template.c:299-412) on a picture of the pipeline's coding
 * shape, at 8 or 10 bits, so the oracle's orc_hevc_deblock_bd can be
 * checked against it.
 *
 * Built only into _ref/ from the reference tree (oracle/Makefile, target
 * ref): it includes the reference decoder's hevc_dec_internal.h and fills
 * the fields its loop filter reads, as the reference decoder would after
 * parsing such a picture:
 *   - 8x8 CUs (log2_min_cb 3) in 16x16 CTUs, every 8x8 cell with a CU and
 *     transform edge on its left and top (edge bits 1|2|4|8);
 *   - per 4x4: intra (pred_flag 0) or uni-L0 with the CU's vector and one
 *     reference picture; the CU's coded-luma flag;
 *   - QpY per CU = qp; slices as CTU-row ranges, deblocking on, not across
 *     slices, zero offsets; one tile.
 * Vertical edges over every CTU row first, then horizontal edges, as
 * hevcd_loop_filters runs its stages. */
#include <stdlib.h>
#include <string.h>

#include "hevc_dec_internal.h"

/* ctu_log2 4 or 5 (CTBs of 16 or 32, slices as CTB-row ranges); culog2 (may
 * be NULL): log2 CU size per CU8 (CU16 / CU32 of a CTU32 P picture): the
 * 8x8 cells inside such a CU carry no edge bits on their inner sides. */
int ref_hevc_deblock_ctu(void *y, void *uv, int wc, int hc, int qp, int is_intra, const uint8_t *cbf,
                         const int32_t *mv4, int num_slices, int bd, int ctu_log2, const uint8_t *culog2)
{
    const size_t bps = bd > 8 ? 2 : 1;  /* bytes per sample; planes hold the sample in the low bits */
    const int csz = 1 << ctu_log2;
    const int wctu = (wc + csz - 1) / csz, hctu = (hc + csz - 1) / csz, w4 = wc / 4, h4 = hc / 4, w8 = wc / 8,
              h8 = hc / 8, w16 = wc / 16;
    hevc_sps_t *sps = calloc(1, sizeof *sps);
    hevc_pps_t *pps = calloc(1, sizeof *pps);
    hevcd_t *d = calloc(1, sizeof *d);
    hevcd_mvf_t *mvf = calloc((size_t)w4 * h4, sizeof *mvf);
    uint8_t *cbf4 = calloc((size_t)w4 * h4, 1), *edges = malloc((size_t)w8 * h8);
    int8_t *qpm = malloc((size_t)w8 * h8);
    int32_t *slice_of = malloc(sizeof(int32_t) * wctu * hctu);
    hevcd_slice_filter_t *sf = calloc((size_t)(num_slices > 0 ? num_slices : 1), sizeof *sf);
    uint8_t *cb = malloc((size_t)(wc / 2) * (hc / 2) * bps), *cr = malloc((size_t)(wc / 2) * (hc / 2) * bps);
    int rc = -1;
    if (!sps || !pps || !d || !mvf || !cbf4 || !edges || !qpm || !slice_of || !sf || !cb || !cr) goto out;

    sps->width = wc;
    sps->height = hc;
    sps->bit_depth_luma = sps->bit_depth_chroma = bd;
    sps->log2_min_cb = 3;
    sps->log2_ctb = ctu_log2;
    sps->ctb_size = csz;
    sps->ctb_width = wctu;
    sps->ctb_height = hctu;
    sps->ctb_count = wctu * hctu;
    sps->min_cb_width = w8;
    sps->min_cb_height = h8;

    int ns = num_slices < 1 ? 1 : (num_slices > hctu ? hctu : num_slices);
    for (int r = 0; r < hctu; r++) {
        int s = 0;
        for (int k = 0; k < ns; k++)
            if (r >= (int)(((long long)hctu * k) / ns)) s = k;
        for (int c = 0; c < wctu; c++) slice_of[r * wctu + c] = s;
    }
    for (int k = 0; k < ns; k++) {
        sf[k].disabled = 0;
        sf[k].across_slices = 0;
    }
    memset(edges, 1 | 2 | 4 | 8, (size_t)w8 * h8);
    if (culog2)
        for (int cy = 0; cy < h8; cy++)
            for (int cx = 0; cx < w8; cx++) {
                const int cu = ((cy >> 1) * w16 + (cx >> 1)) * 4 + (cy & 1) * 2 + (cx & 1), L = culog2[cu];
                if (L < 4) continue;
                if ((cx * 8) % (1 << L)) edges[cy * w8 + cx] &= (uint8_t)~(1 | 4);
                if ((cy * 8) % (1 << L)) edges[cy * w8 + cx] &= (uint8_t)~(2 | 8);
            }
    memset(qpm, qp, (size_t)w8 * h8);
    for (int j = 0; j < h4; j++)
        for (int i = 0; i < w4; i++) {
            const int x = 4 * i, yy = 4 * j;
            const int cu = ((yy >> 4) * w16 + (x >> 4)) * 4 + ((yy >> 3) & 1) * 2 + ((x >> 3) & 1);
            hevcd_mvf_t *m = &mvf[j * w4 + i];
            if (!is_intra) {
                m->pred_flag = HEVCD_PF_L0;
                m->mv[0][0] = (int16_t)mv4[4 * cu];
                m->mv[0][1] = (int16_t)mv4[4 * cu + 1];
                m->ref_idx[0] = 0;
                m->ref_idx[1] = -1;
                m->ref_poc[0] = 0;
                cbf4[j * w4 + i] = cbf[cu] ? 1 : 0;
            }
        }
    for (int j = 0; j < hc / 2; j++)
        for (int i = 0; i < wc / 2; i++)
            for (int c = 0; c < 2; c++)
                memcpy((c ? cr : cb) + ((size_t)j * (wc / 2) + i) * bps,
                       (uint8_t *)uv + ((size_t)j * wc + 2 * i + c) * bps, bps);

    d->sps = sps;
    d->pps = pps;
    d->plane[0] = y;
    d->plane[1] = cb;
    d->plane[2] = cr;
    d->stride[0] = wc;
    d->stride[1] = d->stride[2] = wc / 2;
    d->n_planes = 3;
    d->mvf = mvf;
    d->min_pu_width = w4;
    d->min_pu_height = h4;
    d->cbf_map = cbf4;
    d->n_cbf = (size_t)w4 * h4;
    d->edges = edges;
    d->edges_stride = w8;
    d->n_edges = (size_t)w8 * h8;
    d->qp_y_map = qpm;
    d->n_qp = (size_t)w8 * h8;
    d->slice_of_ctb = slice_of;
    d->n_slice_map = (size_t)wctu * hctu;
    d->slice_filter = sf;
    d->n_slice_filter = (size_t)ns;
    d->n_tiles = 1;

    for (int r = 0; r < hctu; r++) hevcd_filter_stage(d, 0, r);
    for (int r = 0; r < hctu; r++) hevcd_filter_stage(d, 1, r);

    for (int j = 0; j < hc / 2; j++)
        for (int i = 0; i < wc / 2; i++)
            for (int c = 0; c < 2; c++)
                memcpy((uint8_t *)uv + ((size_t)j * wc + 2 * i + c) * bps,
                       (c ? cr : cb) + ((size_t)j * (wc / 2) + i) * bps, bps);
    rc = 0;
out:
    free(sps); free(pps); free(d); free(mvf); free(cbf4); free(edges); free(qpm); free(slice_of); free(sf);
    free(cb); free(cr);
    return rc;
}

int ref_hevc_deblock(void *y, void *uv, int wc, int hc, int qp, int is_intra, const uint8_t *cbf,
                     const int32_t *mv4, int num_slices, int bd)
{
    return ref_hevc_deblock_ctu(y, uv, wc, hc, qp, is_intra, cbf, mv4, num_slices, bd, 4, NULL);
}
