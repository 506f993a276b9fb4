/* orc_baseline.c - the CPU baseline encoder (test and measurement
 * infrastructure, like the rest of oracle/): the oracle's HEVC stages over
 * an IPPP sequence with the integer full search in AVX2 (vpsadbw, two rows
 * of the 16x16 block per instruction, the oracle's keys and tie order) and
 * the search and quarter-sample refine split over OpenMP threads; CU coding
 * and deblocking run per picture on one thread.  The reference's own CPU
 * path is a host-slice encoder with SIMD SAD (cpu_simd_me.c:136-141) and
 * the same per-frame stage order (encoder_h265.c:1879-1894); this one codes
 * what the GPU pipeline codes, so its output is the pipeline's. */
#include <immintrin.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

static int clampi(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }
static int iabs_(int v) { return v < 0 ? -v : v; }

/* The +-range integer search of one MB: orc_me_fullsearch_mb's records. */
__attribute__((target("avx2"))) static void fullsearch_mb_avx2(const uint8_t *cur, const uint8_t *ref, int py,
                                                                int wc, int hc, int range, int lambda8, int mbx,
                                                                int mby, uint8_t *win, orc_imv_t *o16,
                                                                orc_imv_t *o8)
{
    const int n = 2 * range + 1, ww = 16 + 2 * range;
    const int x0 = mbx * 16, y0 = mby * 16;
    for (int y = 0; y < ww; y++) {
        const uint8_t *row = ref + (size_t)clampi(y0 - range + y, 0, hc - 1) * py;
        for (int x = 0; x < ww; x++) win[(size_t)y * ww + x] = row[clampi(x0 - range + x, 0, wc - 1)];
    }
    __m256i c2[8];
    for (int k = 0; k < 8; k++)
        c2[k] = _mm256_inserti128_si256(
            _mm256_castsi128_si256(_mm_loadu_si128((const __m128i *)(cur + (size_t)(y0 + 2 * k) * py + x0))),
            _mm_loadu_si128((const __m128i *)(cur + (size_t)(y0 + 2 * k + 1) * py + x0)), 1);
    uint64_t best16 = UINT64_MAX, best8[4] = {UINT64_MAX, UINT64_MAX, UINT64_MAX, UINT64_MAX};
    uint32_t sad16b = 0, sad8b[4] = {0, 0, 0, 0};
    for (int dy = -range; dy <= range; dy++) {
        const uint8_t *r0 = win + (size_t)(dy + range) * ww;
        for (int dx = -range; dx <= range; dx++) {
            const uint8_t *r = r0 + dx + range;
            __m256i top = _mm256_setzero_si256(), bot = _mm256_setzero_si256();
            for (int k = 0; k < 8; k++) {
                const __m256i rr = _mm256_inserti128_si256(
                    _mm256_castsi128_si256(_mm_loadu_si128((const __m128i *)(r + (size_t)(2 * k) * ww))),
                    _mm_loadu_si128((const __m128i *)(r + (size_t)(2 * k + 1) * ww)), 1);
                const __m256i s = _mm256_sad_epu8(rr, c2[k]);
                if (k < 4) top = _mm256_add_epi64(top, s);
                else bot = _mm256_add_epi64(bot, s);
            }
            /* lanes: [row even left, row even right, row odd left, row odd right] */
            uint64_t t[4], b[4];
            _mm256_storeu_si256((__m256i *)t, top);
            _mm256_storeu_si256((__m256i *)b, bot);
            const uint32_t q[4] = {(uint32_t)(t[0] + t[2]), (uint32_t)(t[1] + t[3]), (uint32_t)(b[0] + b[2]),
                                   (uint32_t)(b[1] + b[3])};
            const uint32_t l1 = (uint32_t)(iabs_(dx) + iabs_(dy));
            const uint64_t idx = (uint64_t)((dy + range) * n + (dx + range));
            const uint32_t s16 = q[0] + q[1] + q[2] + q[3];
            const uint64_t k16 = ((uint64_t)(s16 + 4u * lambda8 * l1) << 15) | idx;
            if (k16 < best16) { best16 = k16; sad16b = s16; }
            for (int bb = 0; bb < 4; bb++) {
                const uint64_t k = ((uint64_t)(q[bb] + lambda8 * l1) << 15) | idx;
                if (k < best8[bb]) { best8[bb] = k; sad8b[bb] = q[bb]; }
            }
        }
    }
    const int i16 = (int)(best16 & 0x7fff);
    o16->dx = i16 % n - range;
    o16->dy = i16 / n - range;
    o16->cost = (uint32_t)(best16 >> 15);
    o16->sad = sad16b;
    for (int bb = 0; bb < 4; bb++) {
        const int i = (int)(best8[bb] & 0x7fff);
        o8[bb].dx = i % n - range;
        o8[bb].dy = i / n - range;
        o8[bb].cost = (uint32_t)(best8[bb] >> 15);
        o8[bb].sad = sad8b[bb];
    }
}

int orc_baseline_has_avx2(void) { return __builtin_cpu_supports("avx2"); }

/* The whole-picture search, MBs over nthreads threads (simd 0: the scalar
 * oracle's per-MB search). */
void orc_baseline_fullsearch(const uint8_t *cur, const uint8_t *ref, int py, int wc, int hc, int range, int lambda8,
                             int nthreads, int simd, orc_imv_t *int16, orc_imv_t *int8)
{
    const int mbw = wc / 16, nmb = mbw * (hc / 16), ww = 16 + 2 * range;
    const int avx2 = simd && orc_baseline_has_avx2();
#pragma omp parallel num_threads(nthreads)
    {
        uint8_t *win = (uint8_t *)malloc((size_t)ww * ww + 32);
#pragma omp for schedule(dynamic, 4)
        for (int m = 0; m < nmb; m++) {
            if (avx2) fullsearch_mb_avx2(cur, ref, py, wc, hc, range, lambda8, m % mbw, m / mbw, win, &int16[m],
                                         &int8[4 * m]);
            else orc_me_fullsearch_mb(cur, ref, py, wc, hc, range, lambda8, m % mbw, m / mbw, &int16[m], &int8[4 * m]);
        }
        free(win);
    }
}

/* An IPPP sequence of n NV12 pictures (wc x hc, pitch wc) in the HEVC
 * pipeline's stage order (the oracle's hevc_encode_frame): I -> CTU16 intra
 * coding; P -> +-range search of the 8x8 blocks, quarter-sample refine of
 * every 8x8 CU, inter CU coding; then deblocking.  frame_ms[i] receives each
 * picture's wall time; out_y / out_uv (optional) the last reconstruction. */
int orc_baseline_hevc_seq(const uint8_t *ys, const uint8_t *uvs, int n, int wc, int hc, int qp, int range,
                          int lambda8, int ns, int nthreads, int simd, double *frame_ms, uint8_t *out_y,
                          uint8_t *out_uv)
{
    const int mbw = wc / 16, t = mbw * (hc / 16);
    const size_t ysz = (size_t)wc * hc, csz = ysz / 4;
    uint8_t *ry = (uint8_t *)malloc(ysz), *ruv = (uint8_t *)malloc(ysz / 2);
    uint8_t *ny = (uint8_t *)malloc(ysz), *nuv = (uint8_t *)malloc(ysz / 2);
    uint8_t *scb = (uint8_t *)malloc(csz), *scr = (uint8_t *)malloc(csz);
    uint8_t *fcb = (uint8_t *)malloc(csz), *fcr = (uint8_t *)malloc(csz);
    uint8_t *rcb = (uint8_t *)malloc(csz), *rcr = (uint8_t *)malloc(csz);
    orc_imv_t *i16 = (orc_imv_t *)malloc(sizeof(orc_imv_t) * t), *i8 = (orc_imv_t *)malloc(sizeof(orc_imv_t) * t * 4);
    int32_t *mv = (int32_t *)malloc(sizeof(int32_t) * t * 8), *mv4 = (int32_t *)calloc((size_t)t * 16, 4);
    int16_t *coef = (int16_t *)malloc(sizeof(int16_t) * t * 384);
    uint8_t *modes = (uint8_t *)malloc((size_t)t * 16), *cbf = (uint8_t *)calloc((size_t)t * 4, 1);
    uint8_t *culog2 = (uint8_t *)malloc((size_t)t * 4);
    if (!ry || !ruv || !ny || !nuv || !scb || !scr || !fcb || !fcr || !rcb || !rcr || !i16 || !i8 || !mv || !mv4 ||
        !coef || !modes || !cbf || !culog2)
        return -1;
    for (int f = 0; f < n; f++) {
        const double t0 = omp_get_wtime();
        const uint8_t *cy = ys + ysz * f, *cuv = uvs + (ysz / 2) * f;
        for (size_t i = 0; i < csz; i++) {
            scb[i] = cuv[2 * i];
            scr[i] = cuv[2 * i + 1];
        }
        if (f == 0) {
            orc_hevc_encode_i(cy, scb, scr, ny, rcb, rcr, 8, wc, hc, qp, ns, 4, modes, coef);
            memset(cbf, 0, (size_t)t * 4);
            memset(mv4, 0, (size_t)t * 16 * 4);
        } else {
            orc_baseline_fullsearch(cy, ry, wc, wc, hc, range, lambda8, nthreads, simd, i16, i8);
            for (size_t i = 0; i < csz; i++) {
                fcb[i] = ruv[2 * i];
                fcr[i] = ruv[2 * i + 1];
            }
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 16)
            for (int cu = 0; cu < t * 4; cu++) {
                const int ctu = cu / 4, z = cu % 4;
                const int x0 = (ctu % mbw) * 16 + (z & 1) * 8, y0 = (ctu / mbw) * 16 + (z >> 1) * 8;
                int mx = i8[cu].dx * 4, my = i8[cu].dy * 4;
                orc_hevc_refine8(cy, ry, 8, wc, hc, x0, y0, &mx, &my);
                mv[2 * cu] = mx;
                mv[2 * cu + 1] = my;
                mv4[4 * cu] = mx;
                mv4[4 * cu + 1] = my;
            }
            orc_hevc_encode_p_ctu(cy, scb, scr, ry, fcb, fcr, ny, rcb, rcr, 8, wc, hc, qp, 2, mv, 4, coef, culog2);
            for (int cu = 0; cu < t * 4; cu++) {
                const int16_t *c = coef + 384 * (cu / 4) + 64 * (cu % 4);
                int any = 0;
                for (int k = 0; k < 64 && !any; k++) any = c[k] != 0;
                cbf[cu] = (uint8_t)any;
            }
        }
        for (size_t i = 0; i < csz; i++) {
            nuv[2 * i] = rcb[i];
            nuv[2 * i + 1] = rcr[i];
        }
        orc_hevc_deblock_ctu(ny, nuv, wc, wc, wc, hc, qp, f == 0, cbf, mv4, ns, 4, 8, NULL);
        uint8_t *sw = ry; ry = ny; ny = sw;
        sw = ruv; ruv = nuv; nuv = sw;
        if (frame_ms) frame_ms[f] = (omp_get_wtime() - t0) * 1e3;
    }
    if (out_y) memcpy(out_y, ry, ysz);
    if (out_uv) memcpy(out_uv, ruv, ysz / 2);
    free(ry); free(ruv); free(ny); free(nuv); free(scb); free(scr); free(fcb); free(fcr); free(rcb); free(rcr);
    free(i16); free(i8); free(mv); free(mv4); free(coef); free(modes); free(cbf); free(culog2);
    return 0;
}
