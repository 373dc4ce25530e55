/* sgm_oracle.c -- CPU restatement of the reference SGM oracles (test
 * infrastructure only; see sgm_oracle.h for what may use it and what pins it).
 *
 * Plain C, integer only. The scalar entry points (orc_sgm_plain, orc_sgm_ms) use
 * one thread and no vector intrinsics, mirroring the reference oracle's build
 * (Makefile:17-18: -O2, no -march). The *_omp variants produce byte-identical
 * output with the reference's output-neutral OpenMP decomposition
 * (multiscale/sgm_ms_ref.c:88-92) and serve as the multithreaded CPU baseline.
 */
#include "sgm_oracle.h"
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* FNV-1a 64, common/util.c:54-59 (offset basis and prime are the published
 * FNV constants). */
uint64_t orc_fnv1a64(const void *data, size_t n)
{
    const uint8_t *p = (const uint8_t *)data;
    uint64_t h = 0xcbf29ce484222325ULL;
    for (size_t i = 0; i < n; i++) {
        h ^= p[i];
        h *= 0x100000001b3ULL;
    }
    return h;
}

static inline int clip(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }

/* common/sgm_ref.c:23-41; multiscale/sgm_ms_ref.c:55-65 (identical rule). */
void orc_census9x7(const uint8_t *im, int W, int H, uint64_t *out)
{
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const int centre = im[(size_t)y * W + x];
            uint64_t code = 0;
            int bit = 0;
            for (int oy = -3; oy <= 3; oy++) {
                const uint8_t *row = im + (size_t)clip(y + oy, 0, H - 1) * W;
                for (int ox = -4; ox <= 4; ox++) {
                    if (oy == 0 && ox == 0)
                        continue;
                    code |= (uint64_t)(row[clip(x + ox, 0, W - 1)] < centre) << bit;
                    bit++;
                }
            }
            out[(size_t)y * W + x] = code;
        }
}

/* common/sgm_ref.c:44-56 (invalid = SGM_COST_INVALID = 62) and
 * multiscale/sgm_ms_ref.c:66-72 (invalid = 63). */
void orc_cost(const uint64_t *cl, const uint64_t *cr, int W, int H, int D,
              int invalid, uint8_t *C)
{
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const uint64_t a = cl[(size_t)y * W + x];
            uint8_t *c = C + ((size_t)y * W + x) * D;
            for (int d = 0; d < D; d++)
                c[d] = x - d < 0 ? (uint8_t)invalid
                                 : (uint8_t)__builtin_popcountll(a ^ cr[(size_t)y * W + x - d]);
        }
}

/* common/sgm_ref.c:61-72 and multiscale/sgm_ms_ref.c:74-84.
 * m = min(Lp(d), Lp(d-1)+P1, Lp(d+1)+P1, min_k Lp(k)+P2), each add saturating at
 * 255, out-of-range neighbours 255; cur = C + m - min_k Lp(k) (saturating when
 * sat != 0). */
void orc_step(const uint8_t *prev, const uint8_t *c, uint8_t *cur, int D,
              int P1, int P2, int sat)
{
    int lo = 255;
    for (int d = 0; d < D; d++)
        if (prev[d] < lo)
            lo = prev[d];
    const int jump = lo + P2 > 255 ? 255 : lo + P2;
    for (int d = 0; d < D; d++) {
        int best = prev[d];
        const int left = d > 0 ? (prev[d - 1] + P1 > 255 ? 255 : prev[d - 1] + P1) : 255;
        const int right = d < D - 1 ? (prev[d + 1] + P1 > 255 ? 255 : prev[d + 1] + P1) : 255;
        if (left < best) best = left;
        if (right < best) best = right;
        if (jump < best) best = jump;
        int v = c[d] + best - lo;
        if (sat && v > 255)
            v = 255;
        cur[d] = (uint8_t)v;
    }
}

/* Direction tables: sgm_ref.c:96-99 (L->R, R->L, T->B, B->T, then diagonals)
 * and sgm_ms_ref.c:85-87 (dx, dy, P1). */
const int orc_dirs_plain[8][2] = {
    {1, 0}, {-1, 0}, {0, 1}, {0, -1}, {1, 1}, {-1, 1}, {1, -1}, {-1, -1}};
const int orc_dirs_ms[8][3] = {
    {1, 0, 40}, {-1, 0, 40}, {0, 1, 20}, {0, -1, 20},
    {1, 1, 10}, {-1, -1, 10}, {1, -1, 10}, {-1, 1, 10}};

/* One pixel of the march: predecessor (x-dx, y-dy) if inside the image, else
 * this pixel starts a path and L = C (sgm_ref.c:82-91). */
static inline void march_pixel(const uint8_t *C, uint8_t *L, int W, int H, int D,
                               int x, int y, int dx, int dy, int P1, int P2, int sat)
{
    const size_t o = ((size_t)y * W + x) * D;
    const int px = x - dx, py = y - dy;
    if (px >= 0 && px < W && py >= 0 && py < H)
        orc_step(L + ((size_t)py * W + px) * D, C + o, L + o, D, P1, P2, sat);
    else
        memcpy(L + o, C + o, (size_t)D);
}

/* The whole L volume of one direction is kept, so the predecessor is read from
 * it directly; row order follows dy and column order follows dx, which visits
 * every predecessor first (sgm_ref.c:77-93). */
void orc_path_L(const uint8_t *C, int W, int H, int D, int dx, int dy,
                int P1, int P2, int sat, uint8_t *Lout)
{
    for (int yi = 0; yi < H; yi++) {
        const int y = dy < 0 ? H - 1 - yi : yi;
        for (int xi = 0; xi < W; xi++) {
            const int x = dx < 0 ? W - 1 - xi : xi;
            march_pixel(C, Lout, W, H, D, x, y, dx, dy, P1, P2, sat);
        }
    }
}

static void path_L_omp(const uint8_t *C, int W, int H, int D, int dx, int dy,
                       int P1, int P2, int sat, uint8_t *Lout)
{
    if (dy == 0) {
        /* every row is an independent recurrence */
#pragma omp parallel for schedule(static)
        for (int y = 0; y < H; y++)
            for (int xi = 0; xi < W; xi++) {
                const int x = dx < 0 ? W - 1 - xi : xi;
                march_pixel(C, Lout, W, H, D, x, y, dx, dy, P1, P2, sat);
            }
    } else {
        /* within a row every column is independent: predecessors lie in the
         * previous row of the march */
        for (int yi = 0; yi < H; yi++) {
            const int y = dy < 0 ? H - 1 - yi : yi;
#pragma omp parallel for schedule(static)
            for (int x = 0; x < W; x++)
                march_pixel(C, Lout, W, H, D, x, y, dx, dy, P1, P2, sat);
        }
    }
}

static void add_L(const uint8_t *L, size_t n, uint16_t *S, int parallel)
{
    if (parallel) {
#pragma omp parallel for schedule(static)
        for (long i = 0; i < (long)n; i++)
            S[i] = (uint16_t)(S[i] + L[i]);
    } else {
        for (size_t i = 0; i < n; i++)
            S[i] = (uint16_t)(S[i] + L[i]);
    }
}

void orc_aggregate_dir(const uint8_t *C, int W, int H, int D, int dx, int dy,
                       int P1, int P2, int sat, uint16_t *S)
{
    const size_t n = (size_t)W * H * D;
    uint8_t *L = (uint8_t *)malloc(n);
    if (!L)
        return;
    orc_path_L(C, W, H, D, dx, dy, P1, P2, sat, L);
    add_L(L, n, S, 0);
    free(L);
}

/* sgm_ref.c:120-124: strict '<', lowest d wins ties. */
void orc_argmin(const uint16_t *S, size_t npix, int D, uint8_t *disp)
{
    for (size_t i = 0; i < npix; i++) {
        const uint16_t *s = S + i * (size_t)D;
        int best = 0;
        for (int d = 1; d < D; d++)
            if (s[d] < s[best])
                best = d;
        disp[i] = (uint8_t)best;
    }
}

static void argmin_omp(const uint16_t *S, size_t npix, int D, uint8_t *disp)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)npix; i++)
        orc_argmin(S + (size_t)i * D, 1, D, disp + i);
}

/* ---- Configuration A: common/sgm_ref.c:101-144 ---- */
#define PLAIN_P1 20      /* sgm_params.h:23 */
#define P_WIN1 PLAIN_P1
#define P_WIN2 PLAIN_P2
#define PLAIN_P2 192     /* sgm_params.h:24 */
#define PLAIN_INVALID 62 /* sgm_params.h:27-29: 9*7-1 */

static int sgm_plain_impl(const uint8_t *left, const uint8_t *right, int W, int H,
                          int D, int paths, uint8_t *disp, int parallel)
{
    if (W <= 0 || H <= 0 || D <= 0 || D > 256 || (paths != 4 && paths != 8))
        return -1;
    const size_t N = (size_t)W * H;
    uint64_t *cl = (uint64_t *)malloc(N * 8), *cr = (uint64_t *)malloc(N * 8);
    uint8_t *C = (uint8_t *)malloc(N * D), *L = (uint8_t *)malloc(N * D);
    uint16_t *S = (uint16_t *)calloc(N * D, 2);
    int rc = -1;
    if (!cl || !cr || !C || !L || !S)
        goto out;
    orc_census9x7(left, W, H, cl);
    orc_census9x7(right, W, H, cr);
    orc_cost(cl, cr, W, H, D, PLAIN_INVALID, C);
    for (int p = 0; p < paths; p++) {
        const int dx = orc_dirs_plain[p][0], dy = orc_dirs_plain[p][1];
        if (parallel)
            path_L_omp(C, W, H, D, dx, dy, PLAIN_P1, PLAIN_P2, 0, L);
        else
            orc_path_L(C, W, H, D, dx, dy, PLAIN_P1, PLAIN_P2, 0, L);
        add_L(L, N * D, S, parallel);
    }
    if (parallel)
        argmin_omp(S, N, D, disp);
    else
        orc_argmin(S, N, D, disp);
    rc = 0;
out:
    free(cl); free(cr); free(C); free(L); free(S);
    return rc;
}

int orc_sgm_plain(const uint8_t *left, const uint8_t *right, int W, int H,
                  int D, int paths, uint8_t *disp)
{
    return sgm_plain_impl(left, right, W, H, D, paths, disp, 0);
}

int orc_sgm_plain_omp(const uint8_t *left, const uint8_t *right, int W, int H,
                      int D, int paths, uint8_t *disp, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    return sgm_plain_impl(left, right, W, H, D, paths, disp, 1);
}

int orc_omp_max_threads(void)
{
#ifdef _OPENMP
    int n = 0;
#pragma omp parallel
    {
#pragma omp single
        n = omp_get_num_threads(); /* the team size observed inside the region */
    }
    return n;
#else
    return 1;
#endif
}

/* ---- Configurations B/C: multiscale/sgm_ms_ref.c ---- */
#define MS_P2 200      /* sgm_ms_ref.c:39 */
#define MS_INVALID 63  /* sgm_ms_ref.c:41 */

/* sgm_ms_ref.c:47-54: Sobel-X with clamped coordinates, out = clamp(128 +
 * (gx >> 3)), arithmetic shift of a possibly negative gx. */
void orc_sobelx(const uint8_t *im, int W, int H, uint8_t *out)
{
    for (int y = 0; y < H; y++) {
        const uint8_t *up = im + (size_t)clip(y - 1, 0, H - 1) * W;
        const uint8_t *mid = im + (size_t)y * W;
        const uint8_t *dn = im + (size_t)clip(y + 1, 0, H - 1) * W;
        for (int x = 0; x < W; x++) {
            const int l = clip(x - 1, 0, W - 1), r = clip(x + 1, 0, W - 1);
            const int gx = (up[r] - up[l]) + 2 * (mid[r] - mid[l]) + (dn[r] - dn[l]);
            out[(size_t)y * W + x] = (uint8_t)clip(128 + (gx >> 3), 0, 255);
        }
    }
}

/* sgm_ms_ref.c:145-150: 2x2 box mean with +2 rounding, from the raw image. */
void orc_half(const uint8_t *im, int W, int H, uint8_t *out)
{
    const int w = W / 2, h = H / 2;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const uint8_t *a = im + (size_t)(2 * y) * W + 2 * x;
            out[(size_t)y * w + x] = (uint8_t)((a[0] + a[1] + a[W] + a[W + 1] + 2) >> 2);
        }
}

/* sgm_ms_ref.c:135-144: Sobel -> census -> cost at one scale. */
static uint8_t *ms_stage_cost(const uint8_t *L, const uint8_t *R, int W, int H, int D)
{
    const size_t N = (size_t)W * H;
    uint8_t *sl = (uint8_t *)malloc(N), *sr = (uint8_t *)malloc(N);
    uint64_t *cl = (uint64_t *)malloc(N * 8), *cr = (uint64_t *)malloc(N * 8);
    uint8_t *C = (uint8_t *)malloc(N * D);
    if (sl && sr && cl && cr && C) {
        orc_sobelx(L, W, H, sl);
        orc_sobelx(R, W, H, sr);
        orc_census9x7(sl, W, H, cl);
        orc_census9x7(sr, W, H, cr);
        orc_cost(cl, cr, W, H, D, MS_INVALID, C);
    } else {
        free(C);
        C = NULL;
    }
    free(sl); free(sr); free(cl); free(cr);
    return C;
}

/* sgm_ms_ref.c:93-133: eight weighted, saturating directions. */
static int ms_aggregate(const uint8_t *C, int W, int H, int D, uint16_t *S, int parallel)
{
    const size_t n = (size_t)W * H * D;
    uint8_t *L = (uint8_t *)malloc(n);
    if (!L)
        return -1;
    for (int p = 0; p < 8; p++) {
        const int dx = orc_dirs_ms[p][0], dy = orc_dirs_ms[p][1], p1 = orc_dirs_ms[p][2];
        if (parallel)
            path_L_omp(C, W, H, D, dx, dy, p1, MS_P2, 1, L);
        else
            orc_path_L(C, W, H, D, dx, dy, p1, MS_P2, 1, L);
        add_L(L, n, S, parallel);
    }
    free(L);
    return 0;
}

static int sgm_ms_impl(const uint8_t *left, const uint8_t *right, int W, int H, int D,
                       uint8_t *disp, uint8_t *disp1, int parallel)
{
    if (W < 2 || H < 2 || (W & 1) || (H & 1) || D <= 0 || D > 256)
        return -1;
    const int hw = W / 2, hh = H / 2;
    const size_t N = (size_t)W * H, Nh = (size_t)hw * hh;
    uint8_t *Lh = (uint8_t *)malloc(Nh), *Rh = (uint8_t *)malloc(Nh);
    uint8_t *C1 = NULL, *C2 = NULL;
    uint16_t *S1 = NULL, *S2 = NULL;
    uint8_t *d1 = NULL;
    int rc = -1;
    if (!Lh || !Rh)
        goto out;
    /* stage 1: full SGM at half resolution (sgm_ms_ref.c:170-181) */
    orc_half(left, W, H, Lh);
    orc_half(right, W, H, Rh);
    C1 = ms_stage_cost(Lh, Rh, hw, hh, D);
    S1 = (uint16_t *)calloc(Nh * D, 2);
    d1 = disp1 ? disp1 : (uint8_t *)malloc(Nh);
    if (!C1 || !S1 || !d1 || ms_aggregate(C1, hw, hh, D, S1, parallel))
        goto out;
    if (parallel)
        argmin_omp(S1, Nh, D, d1);
    else
        orc_argmin(S1, Nh, D, d1);
    free(S1);
    S1 = NULL;
    /* stage 2: full resolution, cost averaged with the half-resolution cost at
     * (x/2, y/2, d/2) (sgm_ms_ref.c:183-188) */
    C2 = ms_stage_cost(left, right, W, H, D);
    if (!C2)
        goto out;
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            uint8_t *c2 = C2 + ((size_t)y * W + x) * D;
            const uint8_t *c1 = C1 + ((size_t)(y / 2) * hw + x / 2) * D;
            for (int d = 0; d < D; d++)
                c2[d] = (uint8_t)((c2[d] + c1[d / 2] + 1) >> 1);
        }
    S2 = (uint16_t *)calloc(N * D, 2);
    if (!S2 || ms_aggregate(C2, W, H, D, S2, parallel))
        goto out;
    /* argmin at (2y, 2x) -> (W/2)x(H/2) map (sgm_ms_ref.c:194-199) */
    for (int y = 0; y < hh; y++)
        for (int x = 0; x < hw; x++)
            orc_argmin(S2 + ((size_t)(2 * y) * W + 2 * x) * D, 1, D, disp + (size_t)y * hw + x);
    rc = 0;
out:
    free(Lh); free(Rh); free(C1); free(C2); free(S1); free(S2);
    if (d1 != disp1)
        free(d1);
    return rc;
}

int orc_sgm_ms(const uint8_t *left, const uint8_t *right, int W, int H, int D,
               uint8_t *disp, uint8_t *disp1)
{
    return sgm_ms_impl(left, right, W, H, D, disp, disp1, 0);
}

int orc_sgm_ms_omp(const uint8_t *left, const uint8_t *right, int W, int H, int D,
                   uint8_t *disp, uint8_t *disp1, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    return sgm_ms_impl(left, right, W, H, D, disp, disp1, 1);
}

/* ---- windowed SGM (build-defined, see sgm_oracle.h) ---- */
#define WIN_ABSENT 100000

static inline int tile_lo(const uint8_t *blk, int tiles_x, int x, int y)
{
    return 16 * (blk[(y >> 4) * tiles_x + (x >> 4)] & 15);
}
static inline int tile_hi(const uint8_t *blk, int tiles_x, int x, int y)
{
    return 16 * ((blk[(y >> 4) * tiles_x + (x >> 4)] >> 4) + 1);
}

int orc_sgm_window(const uint8_t *left, const uint8_t *right, int W, int H, int D,
                   int paths, const uint8_t *blk, uint8_t *disp)
{
    if (W <= 0 || H <= 0 || D <= 0 || D > 256 || D % 16 || (paths != 4 && paths != 8))
        return -1;
    const int tiles_x = (W + 15) / 16, tiles_y = (H + 15) / 16, nb = D / 16;
    for (int t = 0; t < tiles_x * tiles_y; t++)
        if ((blk[t] & 15) > (blk[t] >> 4) || (blk[t] >> 4) >= nb)
            return -1;
    const size_t N = (size_t)W * H;
    uint64_t *cl = (uint64_t *)malloc(N * 8), *cr = (uint64_t *)malloc(N * 8);
    uint8_t *C = (uint8_t *)malloc(N * D), *L = (uint8_t *)malloc(N * D);
    uint16_t *S = (uint16_t *)calloc(N * D, 2);
    int rc = -1;
    if (!cl || !cr || !C || !L || !S)
        goto out;
    orc_census9x7(left, W, H, cl);
    orc_census9x7(right, W, H, cr);
    orc_cost(cl, cr, W, H, D, PLAIN_INVALID, C);
    for (int p = 0; p < paths; p++) {
        const int dx = orc_dirs_plain[p][0], dy = orc_dirs_plain[p][1];
        for (int yi = 0; yi < H; yi++) {
            const int y = dy < 0 ? H - 1 - yi : yi;
            for (int xi = 0; xi < W; xi++) {
                const int x = dx < 0 ? W - 1 - xi : xi;
                const size_t o = ((size_t)y * W + x) * D;
                const int lo = tile_lo(blk, tiles_x, x, y), hi = tile_hi(blk, tiles_x, x, y);
                const int px = x - dx, py = y - dy;
                if (!(px >= 0 && px < W && py >= 0 && py < H)) {
                    for (int d = lo; d < hi; d++)
                        L[o + d] = C[o + d];
                    continue;
                }
                const size_t po = ((size_t)py * W + px) * D;
                const int plo = tile_lo(blk, tiles_x, px, py), phi = tile_hi(blk, tiles_x, px, py);
                int mn = WIN_ABSENT;
                for (int d = plo; d < phi; d++)
                    if (L[po + d] < mn)
                        mn = L[po + d];
                for (int d = lo; d < hi; d++) {
                    const int self = d >= plo && d < phi ? L[po + d] : WIN_ABSENT;
                    const int below = d - 1 >= plo && d - 1 < phi ? L[po + d - 1] + P_WIN1 : WIN_ABSENT;
                    const int above = d + 1 >= plo && d + 1 < phi ? L[po + d + 1] + P_WIN1 : WIN_ABSENT;
                    int m = self;
                    if (below < m) m = below;
                    if (above < m) m = above;
                    if (mn + P_WIN2 < m) m = mn + P_WIN2;
                    if (m > 255) m = 255; /* every candidate saturates at 255 */
                    L[o + d] = (uint8_t)(C[o + d] + m - mn);
                }
            }
        }
        for (size_t i = 0; i < N; i++) {
            const int x = (int)(i % W), y = (int)(i / W);
            const int lo = tile_lo(blk, tiles_x, x, y), hi = tile_hi(blk, tiles_x, x, y);
            for (int d = lo; d < hi; d++)
                S[i * D + d] = (uint16_t)(S[i * D + d] + L[i * D + d]);
        }
    }
    for (size_t i = 0; i < N; i++) {
        const int x = (int)(i % W), y = (int)(i / W);
        const int lo = tile_lo(blk, tiles_x, x, y), hi = tile_hi(blk, tiles_x, x, y);
        int best = lo;
        for (int d = lo + 1; d < hi; d++)
            if (S[i * D + d] < S[i * D + best])
                best = d;
        disp[i] = (uint8_t)best;
    }
    rc = 0;
out:
    free(cl); free(cr); free(C); free(L); free(S);
    return rc;
}
