/*
 * framegen.c - seeded synthetic NV12 frames for tests and benchmarks.
 *
 * All randomness comes from splitmix64 applied to (seed, stream, index)
 * tuples, so a frame is a pure function of (seed, frame index) and rows
 * can be produced on any number of OpenMP threads.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "bc250_host.h"

#define NSIN 8
#define PI2 6.28318530717958647692

struct rect { int w, h, x0, y0; uint8_t *tex; };

struct bc250_gen {
    uint64_t seed;
    int w, h, nrect;
    double per[NSIN], ct[NSIN], st[NSIN], ph[NSIN], amp[NSIN];
    struct rect *rects;
    /* cumulative motion, lazily extended: offsets in qpel, rect positions */
    int nstate;
    int *gx, *gy;     /* [nstate] */
    int *rx, *ry;     /* [nstate * nrect] */
};

static uint64_t mix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
static uint64_t h3(uint64_t seed, uint64_t a, uint64_t b, uint64_t c)
{
    return mix64(mix64(mix64(seed ^ 0x5851f42d4c957f2dull) ^ a) ^ (b * 0x632be59bd9b4e019ull) ^ mix64(c));
}
static double u01(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }
static int randint(uint64_t r, int lo, int hi) { return lo + (int)(r % (uint64_t)(hi - lo + 1)); }

bc250_gen_t *bc250_gen_create(uint64_t seed, int width, int height, int nrect)
{
    if (width <= 0 || height <= 0 || nrect < 0) return NULL;
    bc250_gen_t *g = (bc250_gen_t *)calloc(1, sizeof(*g));
    if (!g) return NULL;
    g->seed = seed;
    g->w = width;
    g->h = height;
    g->nrect = nrect;
    double wsum = 0, wk[NSIN];
    for (int k = 0; k < NSIN; k++) {
        g->per[k] = 8.0 * pow(16.0, u01(h3(seed, 1, k, 0)));
        double th = PI2 * u01(h3(seed, 2, k, 0));
        g->ct[k] = cos(th);
        g->st[k] = sin(th);
        g->ph[k] = PI2 * u01(h3(seed, 3, k, 0));
        wk[k] = 0.5 + u01(h3(seed, 4, k, 0));
        wsum += wk[k];
    }
    for (int k = 0; k < NSIN; k++) g->amp[k] = 40.0 * wk[k] / wsum;
    g->rects = (struct rect *)calloc((size_t)(nrect ? nrect : 1), sizeof(struct rect));
    for (int i = 0; i < nrect; i++) {
        struct rect *r = &g->rects[i];
        r->w = randint(h3(seed, 5, i, 0), 32, 128);
        r->h = randint(h3(seed, 5, i, 1), 32, 128);
        if (r->w > width) r->w = width;
        if (r->h > height) r->h = height;
        r->x0 = randint(h3(seed, 6, i, 0), 0, width - r->w);
        r->y0 = randint(h3(seed, 6, i, 1), 0, height - r->h);
        r->tex = (uint8_t *)malloc((size_t)r->w * r->h);
        for (int p = 0; p < r->w * r->h; p++) r->tex[p] = (uint8_t)randint(h3(seed, 7, i, p), 16, 235);
    }
    return g;
}

void bc250_gen_destroy(bc250_gen_t *g)
{
    if (!g) return;
    for (int i = 0; i < g->nrect; i++) free(g->rects[i].tex);
    free(g->rects);
    free(g->gx); free(g->gy); free(g->rx); free(g->ry);
    free(g);
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* Motion state of frames 0..f. */
static int extend(bc250_gen_t *g, int f)
{
    if (f < g->nstate) return 0;
    int n = f + 1;
    int *gx = (int *)realloc(g->gx, sizeof(int) * n), *gy = (int *)realloc(g->gy, sizeof(int) * n);
    if (gx) g->gx = gx;
    if (gy) g->gy = gy;
    int *rx = (int *)realloc(g->rx, sizeof(int) * (size_t)n * (g->nrect ? g->nrect : 1));
    if (rx) g->rx = rx;
    int *ry = (int *)realloc(g->ry, sizeof(int) * (size_t)n * (g->nrect ? g->nrect : 1));
    if (ry) g->ry = ry;
    if (!gx || !gy || !rx || !ry) return -1;
    for (int i = g->nstate; i < n; i++) {
        if (i == 0) {
            g->gx[0] = g->gy[0] = 0;
            for (int k = 0; k < g->nrect; k++) { g->rx[k] = g->rects[k].x0; g->ry[k] = g->rects[k].y0; }
            continue;
        }
        g->gx[i] = g->gx[i - 1] + randint(h3(g->seed, 8, i, 0), -16, 16);
        g->gy[i] = g->gy[i - 1] + randint(h3(g->seed, 8, i, 1), -16, 16);
        for (int k = 0; k < g->nrect; k++) {
            const struct rect *r = &g->rects[k];
            int x = g->rx[(size_t)(i - 1) * g->nrect + k] + randint(h3(g->seed, 9, i, 2 * k), -8, 8);
            int y = g->ry[(size_t)(i - 1) * g->nrect + k] + randint(h3(g->seed, 9, i, 2 * k + 1), -8, 8);
            g->rx[(size_t)i * g->nrect + k] = clampi(x, 0, g->w - r->w);
            g->ry[(size_t)i * g->nrect + k] = clampi(y, 0, g->h - r->h);
        }
    }
    g->nstate = n;
    return 0;
}

void bc250_gen_offset_qpel(bc250_gen_t *g, int f, int *ox, int *oy)
{
    if (!g || f < 0 || extend(g, f)) { if (ox) *ox = 0; if (oy) *oy = 0; return; }
    if (ox) *ox = g->gx[f];
    if (oy) *oy = g->gy[f];
}

int bc250_gen_frame(bc250_gen_t *g, int f, uint8_t *y, int yp, uint8_t *uv, int uvp)
{
    if (!g || f < 0 || !y || !uv) return -1;
    if (extend(g, f)) return -1;
    const double ox = g->gx[f] * 0.25, oy = g->gy[f] * 0.25;
    const int W = g->w, H = g->h;
    const int *rx = g->rx + (size_t)f * g->nrect, *ry = g->ry + (size_t)f * g->nrect;
#pragma omp parallel for schedule(static)
    for (int yy = 0; yy < H; yy++) {
        uint8_t *row = y + (size_t)yy * yp;
        for (int xx = 0; xx < W; xx++) {
            double v = 128.0;
            const double px = xx - ox, pyy = yy - oy;
            for (int k = 0; k < NSIN; k++)
                v += g->amp[k] * sin(PI2 * (px * g->ct[k] + pyy * g->st[k]) / g->per[k] + g->ph[k]);
            for (int k = 0; k < g->nrect; k++) {
                const struct rect *r = &g->rects[k];
                int tx = xx - rx[k], ty = yy - ry[k];
                if (tx >= 0 && ty >= 0 && tx < r->w && ty < r->h) v = r->tex[ty * r->w + tx];
            }
            /* Box-Muller, sigma 2 */
            uint64_t a = h3(g->seed, 10 + (uint64_t)f, (uint64_t)yy, (uint64_t)xx);
            double u1 = (double)((a >> 11) + 1) * (1.0 / 9007199254740993.0);
            double u2 = (double)(mix64(a) >> 11) * (1.0 / 9007199254740992.0);
            v += 2.0 * sqrt(-2.0 * log(u1)) * cos(PI2 * u2);
            int iv = (int)floor(v + 0.5);
            row[xx] = (uint8_t)clampi(iv, 16, 235);
        }
    }
    const int cw = (W + 1) / 2, ch = (H + 1) / 2;
#pragma omp parallel for schedule(static)
    for (int cy = 0; cy < ch; cy++) {
        const uint8_t *r0 = y + (size_t)clampi(2 * cy, 0, H - 1) * yp;
        const uint8_t *r1 = y + (size_t)clampi(2 * cy + 1, 0, H - 1) * yp;
        uint8_t *o = uv + (size_t)cy * uvp;
        for (int cx = 0; cx < cw; cx++) {
            int x0 = clampi(2 * cx, 0, W - 1), x1 = clampi(2 * cx + 1, 0, W - 1);
            int m = (r0[x0] + r0[x1] + r1[x0] + r1[x1] + 2) >> 2;
            int u = 128 + ((m - 128) * 20) / 108;
            int v = 128 - ((m - 128) * 14) / 108;
            o[2 * cx] = (uint8_t)clampi(u, 108, 148);
            o[2 * cx + 1] = (uint8_t)clampi(v, 108, 148);
        }
    }
    return 0;
}
