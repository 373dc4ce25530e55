/* batch_main.c -- sgm_batch_hip: the batched configuration on the GPU.
 *
 *   sgm_batch_hip [-n N] [-g goldens.txt] [-j out.json] [-f flags] [-v]
 *
 * Builds the 24-pair batch -- 8 x 640x480 D=64 4-path, 8 x 1280x720 D=64
 * 4-path, 4 x 1920x800 D=128 two-scale 8-path, 4 x 1920x1080 D=128 4-path, on
 * seeded piecewise-planar scenes (include/sgm_scene.h) -- uploads it once,
 * plans it with sgmh_batch_plan (the autotuner by default; -f pins the
 * SGMH_BATCH_* flags instead), then times N batched runs and N loops of
 * single-pair launches (sgmh_batch_run_loop) after 3 warm-ups of each, with
 * HIP events on the default stream, and reports the medians.  Inputs are
 * resident in device memory before timing starts.
 *
 * Each pair's disparity map is hashed (FNV-1a 64).  -g names a text file of
 * "<pair name> <fnv hex>" lines; the result line then ends GOLDEN OK or
 * GOLDEN FAIL and a mismatch exits 2, as the reference's binaries do
 * (common/harness.c:182-194, multiscale/sgm_ms_ref.c:151-224).  -j appends a
 * JSON record, -v prints every pair's hash.  Exit 1: bad arguments or a
 * runtime error.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <hip/hip_runtime_api.h>
#include "sgm_hip.h"
#include "sgm_scene.h"

#define NPAIRS 24

typedef struct {
    char name[32];
    int W, H, D, mode, paths;
    sgm_scene_spec scene;
} batch_spec;

/* the batched configuration's pairs, in submission order */
static int make_specs(batch_spec *s)
{
    int n = 0;
    for (int i = 0; i < 8; i++, n++) {
        snprintf(s[n].name, sizeof s[n].name, "p640_%d", i);
        s[n].W = 640, s[n].H = 480, s[n].D = 64, s[n].mode = SGMH_PLAIN, s[n].paths = 4;
        s[n].scene = (sgm_scene_spec){640, 480, 63, 200u + i, 8, 12, 0, 2, 0};
    }
    for (int i = 0; i < 8; i++, n++) {
        snprintf(s[n].name, sizeof s[n].name, "p1280_%d", i);
        s[n].W = 1280, s[n].H = 720, s[n].D = 64, s[n].mode = SGMH_PLAIN, s[n].paths = 4;
        s[n].scene = (sgm_scene_spec){1280, 720, 63, 210u + i, 8, 12, 20, 2, 0};
    }
    for (int i = 0; i < 4; i++, n++) {
        snprintf(s[n].name, sizeof s[n].name, "ms1920x800_%d", i);
        s[n].W = 1920, s[n].H = 800, s[n].D = 128, s[n].mode = SGMH_MULTISCALE, s[n].paths = 8;
        s[n].scene = (sgm_scene_spec){1920, 800, 127, 220u + i, 20, 20, 50, 2, 0};
    }
    for (int i = 0; i < 4; i++, n++) {
        snprintf(s[n].name, sizeof s[n].name, "p1080_%d", i);
        s[n].W = 1920, s[n].H = 1080, s[n].D = 128, s[n].mode = SGMH_PLAIN, s[n].paths = 4;
        s[n].scene = (sgm_scene_spec){1920, 1080, 127, 230u + i, 20, 20, 50, 2, 0};
    }
    return n;
}

static int by_value(const void *a, const void *b)
{
    const double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

#define HIPCK(call)                                                                              \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) {                                                                  \
            fprintf(stderr, "batch_hip: %s:%d: %s\n", __FILE__, __LINE__, hipGetErrorString(e_)); \
            return 1;                                                                            \
        }                                                                                        \
    } while (0)

/* median of `reps` timed calls of fn(b, NULL) after 3 warm-ups, in ms */
static int time_runs(int (*fn)(sgmh_batch *, void *), sgmh_batch *b, int reps, double *ms, double *med)
{
    hipEvent_t e0, e1;
    HIPCK(hipEventCreate(&e0));
    HIPCK(hipEventCreate(&e1));
    for (int r = 0; r < reps + 3; r++) {
        HIPCK(hipEventRecord(e0, NULL));
        if (fn(b, NULL)) {
            fprintf(stderr, "batch_hip: %s\n", sgmh_last_error());
            return 1;
        }
        HIPCK(hipEventRecord(e1, NULL));
        HIPCK(hipEventSynchronize(e1));
        float t = 0.f;
        HIPCK(hipEventElapsedTime(&t, e0, e1));
        if (r >= 3)
            ms[r - 3] = t;
    }
    qsort(ms, (size_t)reps, sizeof(double), by_value);
    *med = ms[reps / 2];
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 0;
}

/* 1: every pair's hash is listed and equal; 0: a mismatch or a missing pair */
static int check_goldens(const char *path, const batch_spec *s, const uint64_t *h, int n)
{
    FILE *f = fopen(path, "r");
    if (!f)
        return 0;
    int ok = 0;
    char name[64];
    unsigned long long v;
    int seen[NPAIRS] = {0};
    while (fscanf(f, "%63s %llx", name, &v) == 2)
        for (int i = 0; i < n; i++)
            if (!strcmp(name, s[i].name)) {
                seen[i] = 1;
                if ((uint64_t)v != h[i]) {
                    fprintf(stderr, "batch_hip: %s hash %016llx, golden %016llx\n", s[i].name,
                            (unsigned long long)h[i], v);
                    seen[i] = -1;
                }
            }
    fclose(f);
    ok = 1;
    for (int i = 0; i < n; i++)
        if (seen[i] != 1) {
            if (seen[i] == 0)
                fprintf(stderr, "batch_hip: no golden for %s\n", s[i].name);
            ok = 0;
        }
    return ok;
}

int main(int argc, char **argv)
{
    const char *gp = NULL, *jp = NULL;
    int reps = 20, flags = SGMH_BATCH_AUTOTUNE, verbose = 0;
    for (int i = 1; i < argc; i++) {
        if (!strcmp(argv[i], "-n") && i + 1 < argc) reps = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-g") && i + 1 < argc) gp = argv[++i];
        else if (!strcmp(argv[i], "-j") && i + 1 < argc) jp = argv[++i];
        else if (!strcmp(argv[i], "-f") && i + 1 < argc) flags = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-v")) verbose = 1;
        else {
            fprintf(stderr, "usage: %s [-n N] [-g goldens.txt] [-j json] [-f flags] [-v]\n", argv[0]);
            return 1;
        }
    }
    if (reps < 1)
        reps = 1;
    batch_spec spec[NPAIRS];
    const int n = make_specs(spec);
    sgmh_pair pairs[NPAIRS];
    uint8_t *host_out[NPAIRS];
    double mde = 0.0; /* disparity evaluations per batch, millions (multiscale: both scales) */
    int nplain = 0;
    for (int i = 0; i < n; i++) {
        const batch_spec *s = &spec[i];
        const size_t N = (size_t)s->W * s->H;
        const size_t out = s->mode == SGMH_PLAIN ? N : N / 4;
        uint8_t *L = malloc(N), *R = malloc(N);
        host_out[i] = malloc(out);
        if (!L || !R || !host_out[i] || sgm_scene_planes(&s->scene, L, R, NULL)) {
            fprintf(stderr, "batch_hip: cannot make pair %s\n", s->name);
            return 1;
        }
        uint8_t *dl, *dr, *dd;
        HIPCK(hipMalloc((void **)&dl, N));
        HIPCK(hipMalloc((void **)&dr, N));
        HIPCK(hipMalloc((void **)&dd, out));
        HIPCK(hipMemcpy(dl, L, N, hipMemcpyHostToDevice));
        HIPCK(hipMemcpy(dr, R, N, hipMemcpyHostToDevice));
        free(L);
        free(R);
        pairs[i] = (sgmh_pair){dl, dr, s->W, s->H, s->D, s->mode, s->paths, dd};
        mde += (double)(s->mode == SGMH_PLAIN ? N : N + N / 4) * s->D / 1e6;
        nplain += s->mode == SGMH_PLAIN;
    }
    sgmh_ctx *ctx = NULL;
    sgmh_batch *b = NULL;
    if (sgmh_ctx_create(&ctx, 0) || sgmh_batch_plan(ctx, pairs, n, flags, &b)) {
        fprintf(stderr, "batch_hip: %s\n", sgmh_last_error());
        return 1;
    }
    double *ms = malloc(sizeof(double) * (size_t)reps);
    double med_loop = 0.0, med = 0.0;
    if (!ms || time_runs(sgmh_batch_run_loop, b, reps, ms, &med_loop) || time_runs(sgmh_batch_run, b, reps, ms, &med))
        return 1;
    uint64_t h[NPAIRS];
    for (int i = 0; i < n; i++) {
        const size_t out = (size_t)spec[i].W * spec[i].H / (spec[i].mode == SGMH_PLAIN ? 1 : 4);
        HIPCK(hipMemcpy(host_out[i], pairs[i].d_disp, out, hipMemcpyDeviceToHost));
        h[i] = sgm_fnv1a64(host_out[i], out);
        if (verbose)
            printf("%-14s %dx%d D=%d %s  hash %016llx\n", spec[i].name, spec[i].W, spec[i].H, spec[i].D,
                   spec[i].mode == SGMH_PLAIN ? "plain" : "multiscale", (unsigned long long)h[i]);
    }
    const int gm = gp ? check_goldens(gp, spec, h, n) : -1;
    printf("batch_hip  pairs %d (%d plain, %d multiscale)  median %.3f ms  frames/s %.1f  loop %.3f ms  "
           "speedup %.3f  MDE/s %.0f%s\n",
           n, nplain, n - nplain, med, n * 1000.0 / med, med_loop, med_loop / med, mde * 1000.0 / med,
           gm == 1 ? "  GOLDEN OK" : gm == 0 ? "  GOLDEN FAIL" : "");
    if (jp) {
        FILE *f = fopen(jp, "a");
        if (f) {
            static char plan[32768];
            const int k = sgmh_batch_describe(b, plan, sizeof plan);
            if (k < 0 || (size_t)k >= sizeof plan) /* an error, or truncated */
                strcpy(plan, "null");
            fprintf(f, "{\"impl\":\"batch_hip\",\"pairs\":%d,\"median_ms\":%.4f,\"loop_median_ms\":%.4f,"
                       "\"speedup_vs_loop\":%.4f,\"frames_per_s\":%.1f,\"golden_match\":%d,\"plan\":%s}\n",
                    n, med, med_loop, med_loop / med, n * 1000.0 / med, gm, plan);
            fclose(f);
        }
    }
    sgmh_batch_destroy(b);
    sgmh_ctx_destroy(ctx);
    for (int i = 0; i < n; i++) {
        (void)hipFree((void *)pairs[i].d_left);
        (void)hipFree((void *)pairs[i].d_right);
        (void)hipFree(pairs[i].d_disp);
        free(host_out[i]);
    }
    free(ms);
    return gm == 0 ? 2 : 0;
}
