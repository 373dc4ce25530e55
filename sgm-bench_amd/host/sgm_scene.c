/* sgm_scene.c -- synthetic stereo scenes, PGM I/O, FNV-1a, clock.
 *
 * Built with -ffp-contract=off: sgm_scene_reference must round every float
 * operation exactly as the reference generator's x86-64 build does, or the
 * regenerated inputs (and so every golden computed on them) drift by a byte.
 */
#include "sgm_scene.h"
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------------ */
/* Reference scene: common/gen_synthetic.c                                   */
/* ------------------------------------------------------------------------ */

/* xorshift32 (13, 17, 5), gen_synthetic.c:21-22 */
static inline uint32_t xs32(uint32_t *s)
{
    uint32_t v = *s;
    v ^= v << 13;
    v ^= v >> 17;
    v ^= v << 5;
    *s = v;
    return v;
}

/* Three octaves of bilinearly interpolated lattice noise at cell sizes 2, 8,
 * 32 with amplitudes 1, 1/2, 1/4, scaled by 255/1.75 and truncated
 * (gen_synthetic.c:25-44).  The float expression order is the reference's. */
static int value_noise(uint8_t *im, int W, int H, uint32_t seed)
{
    uint32_t s = seed ? seed : 1u;
    float *acc = (float *)calloc((size_t)W * H, sizeof(float));
    if (!acc)
        return -1;
    int cell = 2;
    for (int oct = 0; oct < 3; oct++, cell *= 4) {
        const int gw = W / cell + 2, gh = H / cell + 2;
        float *lat = (float *)malloc(sizeof(float) * (size_t)gw * gh);
        if (!lat) {
            free(acc);
            return -1;
        }
        for (int i = 0; i < gw * gh; i++)
            lat[i] = (float)(xs32(&s) & 0xffffu) / 65535.0f;
        const float amp = 1.0f / (float)(1 << oct);
        for (int y = 0; y < H; y++) {
            const float fy = (float)y / (float)cell;
            const int iy = (int)fy;
            const float ty = fy - (float)iy;
            for (int x = 0; x < W; x++) {
                const float fx = (float)x / (float)cell;
                const int ix = (int)fx;
                const float tx = fx - (float)ix;
                const float *g0 = lat + (size_t)iy * gw + ix, *g1 = g0 + gw;
                float v = g0[0] * (1.0f - tx) * (1.0f - ty);
                v = v + g0[1] * tx * (1.0f - ty);
                v = v + g1[0] * (1.0f - tx) * ty;
                v = v + g1[1] * tx * ty;
                acc[(size_t)y * W + x] += v * amp;
            }
        }
        free(lat);
    }
    for (size_t i = 0; i < (size_t)W * H; i++) {
        const float v = acc[i] / 1.75f * 255.0f;
        im[i] = v < 0.0f ? 0 : v > 255.0f ? 255 : (uint8_t)v;
    }
    free(acc);
    return 0;
}

int sgm_scene_reference(int W, int H, int D, uint32_t seed,
                        uint8_t *left, uint8_t *right, uint8_t *gt)
{
    if (W < 8 || H < 8 || D < 8 || D > 256 || !left || !right)
        return -1;
    const size_t N = (size_t)W * H;
    uint8_t *slab = (uint8_t *)malloc(N), *dr = (uint8_t *)malloc(N);
    uint8_t *g = gt ? gt : (uint8_t *)malloc(N);
    int rc = -1;
    if (!slab || !dr || !g)
        goto out;
    if (value_noise(right, W, H, seed) || value_noise(slab, W, H, seed * 7919u + 17u))
        goto out;

    /* disparity in right-image coordinates: background D/8, then 12 slabs,
     * slab s drawn from band s of [D/8+1, D-1], the last pinned to D-1; a
     * nearer slab (larger d) replaces both disparity and texture
     * (gen_synthetic.c:57-99) */
    const int bg = D / 8, nslab = 12;
    memset(dr, bg, N);
    uint32_t s = seed * 31u + 5u;
    const int dlo = bg + 1, dhi = D - 1;
    for (int k = 0; k < nslab; k++) {
        const int sw = (int)((unsigned)(W / 6) + xs32(&s) % (unsigned)(W / 4));
        const int sh = (int)((unsigned)(H / 6) + xs32(&s) % (unsigned)(H / 4));
        const int sx = (int)(xs32(&s) % (unsigned)(W - sw));
        const int sy = (int)(xs32(&s) % (unsigned)(H - sh));
        const int band_lo = dlo + (dhi - dlo) * k / nslab;
        int band_hi = dlo + (dhi - dlo) * (k + 1) / nslab;
        if (band_hi <= band_lo)
            band_hi = band_lo + 1;
        const int d = k == nslab - 1 ? dhi
                                     : band_lo + (int)(xs32(&s) % (unsigned)(band_hi - band_lo));
        for (int y = sy; y < sy + sh; y++)
            for (int x = sx; x < sx + sw; x++) {
                const size_t i = (size_t)y * W + x;
                if (d > dr[i]) {
                    dr[i] = (uint8_t)d;
                    right[i] = slab[i];
                }
            }
    }
    /* left(x + d) = right(x), the larger d wins a collision (the sweep goes
     * left to right and ties go to the later pixel); holes take the
     * background match (gen_synthetic.c:100-112) */
    memset(g, 0, N);
    for (int y = 0; y < H; y++) {
        uint8_t *lrow = left + (size_t)y * W, *grow = g + (size_t)y * W;
        const uint8_t *rrow = right + (size_t)y * W, *drow = dr + (size_t)y * W;
        memset(lrow, 0, (size_t)W);
        for (int x = 0; x < W; x++) {
            const int d = drow[x], xl = x + d;
            if (xl < W && d >= grow[xl]) {
                lrow[xl] = rrow[x];
                grow[xl] = (uint8_t)d;
            }
        }
        for (int x = 0; x < W; x++)
            if (grow[x] == 0) {
                const int xs = x - bg < 0 ? 0 : x - bg;
                lrow[x] = rrow[xs];
                grow[x] = (uint8_t)bg;
            }
    }
    rc = 0;
out:
    free(slab);
    free(dr);
    if (g != gt)
        free(g);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Piecewise-planar scenes (no reference counterpart)                        */
/* ------------------------------------------------------------------------ */

/* splitmix64 (Steele, Lea, Flood 2014) */
static inline uint64_t sm64(uint64_t *s)
{
    uint64_t z = (*s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

/* uniform integer in [a, b] */
static inline int urange(uint64_t *s, int a, int b)
{
    return b <= a ? a : a + (int)(sm64(s) % (uint64_t)(b - a + 1));
}

int sgm_scene_planes(const sgm_scene_spec *sp, uint8_t *left, uint8_t *right, uint8_t *gt)
{
    if (!sp || !left || !right)
        return -1;
    const int W = sp->W, H = sp->H, dmax = sp->dmax;
    if (W < 16 || H < 16 || dmax < 1 || dmax > 255 || sp->planes_lo < 0 ||
        sp->planes_hi < sp->planes_lo || sp->occl_permille < 0 || sp->occl_permille > 1000 ||
        sp->noise < 0 || sp->noise > 64)
        return -1;
    const size_t N = (size_t)W * H;
    uint8_t *raw = (uint8_t *)malloc(N);
    uint8_t *g = gt ? gt : (uint8_t *)malloc(N);
    if (!raw || !g) {
        free(raw);
        if (g != gt)
            free(g);
        return -1;
    }
    uint64_t s = 0x53474d5343454eULL ^ (uint64_t)sp->seed; /* "SGMSCEN" ^ seed */
    for (size_t i = 0; i < N; i++)
        raw[i] = (uint8_t)(sm64(&s) >> 56);
    /* left = 3x3 box filter of the noise, clamped borders, rounded */
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            int acc = 0;
            for (int oy = -1; oy <= 1; oy++) {
                const int yy = y + oy < 0 ? 0 : y + oy >= H ? H - 1 : y + oy;
                for (int ox = -1; ox <= 1; ox++) {
                    const int xx = x + ox < 0 ? 0 : x + ox >= W ? W - 1 : x + ox;
                    acc += raw[(size_t)yy * W + xx];
                }
            }
            left[(size_t)y * W + x] = (uint8_t)((acc + 4) / 9);
        }
    /* disparity field: a low background, then planes painted in order */
    memset(g, (uint8_t)urange(&s, 0, dmax / 4), N);
    const int n = urange(&s, sp->planes_lo, sp->planes_hi);
    for (int k = 0; k < n; k++) {
        const int rw = urange(&s, W / 16 > 1 ? W / 16 : 1, W / 3);
        const int rh = urange(&s, H / 16 > 1 ? H / 16 : 1, H / 3);
        const int rx = urange(&s, 0, W - rw), ry = urange(&s, 0, H - rh);
        const int d0 = k == n - 1 ? dmax : urange(&s, 0, dmax);
        int gx = 0, gy = 0;
        if (sp->slanted) {
            gx = urange(&s, -48, 48); /* slope in 1/1024 disparity per pixel */
            gy = urange(&s, -48, 48);
        }
        const int cx = rx + rw / 2, cy = ry + rh / 2;
        for (int y = ry; y < ry + rh; y++)
            for (int x = rx; x < rx + rw; x++) {
                const int64_t num = (int64_t)d0 * 1024 + (int64_t)gx * (x - cx) + (int64_t)gy * (y - cy);
                int64_t d = (num + 512) >> 10;
                d = d < 0 ? 0 : d > dmax ? dmax : d;
                g[(size_t)y * W + x] = (uint8_t)d;
            }
    }
    /* right(x) = left(x + d(x)) + noise; occluded pixels get fresh texture */
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const size_t i = (size_t)y * W + x;
            const uint64_t r = sm64(&s);
            const int occluded = (int)(r % 1000) < sp->occl_permille;
            int xs = x + g[i];
            if (xs > W - 1)
                xs = W - 1;
            int v = left[(size_t)y * W + xs];
            if (sp->noise)
                v += (int)((r >> 20) % (uint64_t)(2 * sp->noise + 1)) - sp->noise;
            if (occluded)
                v = (int)((r >> 40) & 0xff);
            right[i] = (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v);
        }
    free(raw);
    if (g != gt)
        free(g);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Tie-rich scenes (no reference counterpart)                                */
/* ------------------------------------------------------------------------ */

/* A texture that repeats every `period` columns (one random pattern per band
 * of 8 rows, K = 6 patterns), flat rectangles of one grey level painted over
 * it, and rectangular planes of integer disparity.  Where a plane's disparity
 * is g, right(x) = left(x + g) with no noise, so the matching cost is the
 * same at g, g + period, g + 2 period, ... and flat windows cost the same at
 * every disparity: the path sums have many tied minima, which is what pins
 * the argmin's lowest-d rule (common/sgm_ref.c:120-124). */
int sgm_scene_periodic(int W, int H, int dmax, int period, uint32_t seed,
                       uint8_t *left, uint8_t *right, uint8_t *gt)
{
    if (!left || !right || W < 16 || H < 16 || dmax < 1 || dmax > 255 || period < 2 || period > 64)
        return -1;
    const size_t N = (size_t)W * H;
    uint8_t *g = gt ? gt : (uint8_t *)malloc(N);
    if (!g)
        return -1;
    uint64_t s = 0x5449455343454eULL ^ (uint64_t)seed; /* "TIESCEN" ^ seed */
    enum { K = 6 };
    uint8_t pat[K][64];
    for (int k = 0; k < K; k++)
        for (int i = 0; i < period; i++)
            pat[k][i] = (uint8_t)(sm64(&s) >> 56);
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++)
            left[(size_t)y * W + x] = pat[(y / 8) % K][x % period];
    const int nflat = urange(&s, 2, 4);
    for (int k = 0; k < nflat; k++) {
        const int rw = urange(&s, W / 10 > 1 ? W / 10 : 1, W / 4), rh = urange(&s, H / 10 > 1 ? H / 10 : 1, H / 4);
        const int rx = urange(&s, 0, W - rw), ry = urange(&s, 0, H - rh);
        const uint8_t v = (uint8_t)urange(&s, 16, 240);
        for (int y = ry; y < ry + rh; y++)
            memset(left + (size_t)y * W + rx, v, (size_t)rw);
    }
    memset(g, (uint8_t)urange(&s, 0, dmax / 4), N);
    const int n = urange(&s, 3, 5);
    for (int k = 0; k < n; k++) {
        const int rw = urange(&s, W / 8 > 1 ? W / 8 : 1, W / 3), rh = urange(&s, H / 8 > 1 ? H / 8 : 1, H / 3);
        const int rx = urange(&s, 0, W - rw), ry = urange(&s, 0, H - rh);
        const uint8_t d = (uint8_t)(k == n - 1 ? dmax : urange(&s, 0, dmax));
        for (int y = ry; y < ry + rh; y++)
            memset(g + (size_t)y * W + rx, d, (size_t)rw);
    }
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const size_t i = (size_t)y * W + x;
            const int xs = x + g[i] > W - 1 ? W - 1 : x + g[i];
            right[i] = left[(size_t)y * W + xs];
        }
    if (g != gt)
        free(g);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* PGM, hash, clock                                                          */
/* ------------------------------------------------------------------------ */

static int pgm_token(FILE *f, int *out)
{
    int c;
    for (;;) {
        c = fgetc(f);
        if (c == EOF)
            return -1;
        if (c == '#') {
            while (c != '\n' && c != EOF)
                c = fgetc(f);
            continue;
        }
        if (!isspace(c))
            break;
    }
    ungetc(c, f);
    return fscanf(f, "%d", out) == 1 ? 0 : -1;
}

uint8_t *sgm_pgm_read(const char *path, int *W, int *H)
{
    FILE *f = fopen(path, "rb");
    if (!f) {
        fprintf(stderr, "pgm_read: cannot open %s\n", path);
        return NULL;
    }
    char magic[2];
    int w, h, maxv;
    if (fread(magic, 1, 2, f) != 2 || magic[0] != 'P' || magic[1] != '5') {
        fprintf(stderr, "pgm_read: %s is not binary PGM (P5)\n", path);
        fclose(f);
        return NULL;
    }
    if (pgm_token(f, &w) || pgm_token(f, &h) || pgm_token(f, &maxv) || w <= 0 || h <= 0) {
        fprintf(stderr, "pgm_read: bad header in %s\n", path);
        fclose(f);
        return NULL;
    }
    fgetc(f); /* the single whitespace byte after maxval */
    if (maxv != 255) {
        fprintf(stderr, "pgm_read: %s maxval %d, need 255\n", path, maxv);
        fclose(f);
        return NULL;
    }
    const size_t n = (size_t)w * h;
    uint8_t *img = (uint8_t *)malloc(n);
    if (!img || fread(img, 1, n, f) != n) {
        fprintf(stderr, "pgm_read: short read in %s\n", path);
        free(img);
        fclose(f);
        return NULL;
    }
    fclose(f);
    *W = w;
    *H = h;
    return img;
}

int sgm_pgm_write(const char *path, const uint8_t *img, int W, int H)
{
    FILE *f = fopen(path, "wb");
    if (!f) {
        fprintf(stderr, "pgm_write: cannot open %s\n", path);
        return -1;
    }
    fprintf(f, "P5\n%d %d\n255\n", W, H);
    const size_t n = (size_t)W * H;
    const int ok = fwrite(img, 1, n, f) == n;
    return (fclose(f) == 0 && ok) ? 0 : -1;
}

void sgm_free(void *p) { free(p); }

uint64_t sgm_fnv1a64(const void *data, size_t n)
{
    const uint8_t *p = (const uint8_t *)data;
    uint64_t h = 0xcbf29ce484222325ULL; /* FNV-1a 64 offset basis */
    for (size_t i = 0; i < n; i++)
        h = (h ^ p[i]) * 0x100000001b3ULL; /* FNV 64 prime */
    return h;
}

double sgm_now_ms(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec * 1e3 + (double)t.tv_nsec * 1e-6;
}
