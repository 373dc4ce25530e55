/* ms_main.c -- sgm_ms_hip: Configurations B/C on the GPU.
 *
 *   sgm_ms_hip L.pgm R.pgm [-o out.pgm] [-g golden.pgm] [-n N] [-j out.json] [-D d]
 *
 * The command line, result line and exit code follow the reference's
 * multiscale binaries (multiscale/sgm_ms_ref.c:151-224, sgm_ms_cuda.cu:160-223)
 * so scripts/record_ms.sh parses it unchanged: 3 warm-up frames, N timed
 * frames (median), FNV-1a of the (W/2)x(H/2) map, "GOLDEN OK/FAIL", exit 2 on
 * a mismatch.  MDE/s(work) counts both scales, (W/2*H/2 + W*H)*D per frame
 * (sgm_ms_ref.c:211-216).  Frames run on host buffers, transfers included,
 * as the reference's CUDA binary times them.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "sgm_hip.h"
#include "sgm_scene.h"

static int by_value(const void *a, const void *b)
{
    const double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

int main(int argc, char **argv)
{
    if (argc < 3) {
        fprintf(stderr, "usage: %s L.pgm R.pgm [-o out] [-g golden] [-n N] [-j json] [-D d]\n", argv[0]);
        return 1;
    }
    const char *op = NULL, *gp = NULL, *jp = NULL;
    int reps = 1, D = 64;
    for (int i = 3; i < argc; i++) {
        if (!strcmp(argv[i], "-o") && i + 1 < argc) op = argv[++i];
        else if (!strcmp(argv[i], "-g") && i + 1 < argc) gp = argv[++i];
        else if (!strcmp(argv[i], "-n") && i + 1 < argc) reps = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-j") && i + 1 < argc) jp = argv[++i];
        else if (!strcmp(argv[i], "-D") && i + 1 < argc) D = atoi(argv[++i]);
    }
    if (reps < 1)
        reps = 1;
    int W, H, W2, H2;
    uint8_t *L = sgm_pgm_read(argv[1], &W, &H), *R = sgm_pgm_read(argv[2], &W2, &H2);
    if (!L || !R || W != W2 || H != H2) {
        fprintf(stderr, "bad input\n");
        return 1;
    }
    const int ow = W / 2, oh = H / 2;
    uint8_t *disp = malloc((size_t)ow * oh);
    double *ms = malloc(sizeof(double) * (size_t)reps);
    sgmh_ctx *ctx = NULL;
    if (!disp || !ms || sgmh_ctx_create(&ctx, 0)) {
        fprintf(stderr, "ms_hip: %s\n", sgmh_last_error());
        return 1;
    }
    for (int r = 0; r < reps + 3; r++) {
        const double t0 = sgm_now_ms();
        if (sgmh_ms(ctx, L, R, W, H, D, disp, NULL, NULL)) {
            fprintf(stderr, "ms_hip: %s\n", sgmh_last_error());
            return 1;
        }
        if (r >= 3)
            ms[r - 3] = sgm_now_ms() - t0;
    }
    qsort(ms, (size_t)reps, sizeof(double), by_value);
    const double med = ms[reps / 2];
    const uint64_t h = sgm_fnv1a64(disp, (size_t)ow * oh);
    int gm = -1;
    if (gp) {
        int gw, gh;
        uint8_t *G = sgm_pgm_read(gp, &gw, &gh);
        gm = (G && gw == ow && gh == oh) ? !memcmp(G, disp, (size_t)ow * oh) : 0;
        free(G);
    }
    const double work = (double)ow * oh + (double)W * H;
    printf("ms_hip  in %dx%d out %dx%d D=%d paths=8  median %.2f ms  fps %.2f  MDE/s(work) %.0f  hash %016llx%s\n",
           W, H, ow, oh, D, med, 1000.0 / med, work * D / med / 1000.0, (unsigned long long)h,
           gm == 1 ? "  GOLDEN OK" : gm == 0 ? "  GOLDEN FAIL" : "");
    if (op)
        sgm_pgm_write(op, disp, ow, oh);
    if (jp) {
        FILE *f = fopen(jp, "a");
        if (f) {
            fprintf(f, "{\"impl\":\"ms_hip\",\"W\":%d,\"H\":%d,\"outW\":%d,\"outH\":%d,\"D\":%d,\"paths\":8,"
                       "\"median_ms\":%.3f,\"hash\":\"%016llx\",\"golden_match\":%d}\n",
                    W, H, ow, oh, D, med, (unsigned long long)h, gm);
            fclose(f);
        }
    }
    sgmh_ctx_destroy(ctx);
    free(L);
    free(R);
    free(disp);
    free(ms);
    return gm == 0 ? 2 : 0;
}
