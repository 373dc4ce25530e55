/* gen_scene.c -- write a synthetic stereo pair as PGM files.
 *
 *   gen_scene reference W H OUTDIR [seed] [D]   the reference's scene
 *                                               (common/gen_synthetic.c, CLI :46-50)
 *   gen_scene planes W H OUTDIR seed dmax planes_lo planes_hi occl_permille noise slanted
 *
 * writes OUTDIR/left.pgm, right.pgm, gt.pgm.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "sgm_scene.h"

int main(int argc, char **argv)
{
    if (argc < 5) {
        fprintf(stderr, "usage: %s reference W H OUTDIR [seed] [D]\n"
                        "       %s planes W H OUTDIR seed dmax lo hi occl_permille noise slanted\n",
                argv[0], argv[0]);
        return 1;
    }
    const int W = atoi(argv[2]), H = atoi(argv[3]);
    const char *out = argv[4];
    uint8_t *l = malloc((size_t)W * H), *r = malloc((size_t)W * H), *g = malloc((size_t)W * H);
    if (!l || !r || !g)
        return 1;
    int rc;
    if (!strcmp(argv[1], "reference")) {
        const uint32_t seed = argc > 5 ? (uint32_t)atoi(argv[5]) : 1u;
        const int D = argc > 6 ? atoi(argv[6]) : 64;
        rc = sgm_scene_reference(W, H, D, seed, l, r, g);
    } else if (!strcmp(argv[1], "planes") && argc >= 12) {
        sgm_scene_spec sp = {W, H, atoi(argv[6]), (uint32_t)atoi(argv[5]), atoi(argv[7]),
                             atoi(argv[8]), atoi(argv[9]), atoi(argv[10]), atoi(argv[11])};
        rc = sgm_scene_planes(&sp, l, r, g);
    } else {
        fprintf(stderr, "unknown generator or missing arguments\n");
        return 1;
    }
    if (rc) {
        fprintf(stderr, "generator rejected the arguments\n");
        return 1;
    }
    char p[4096];
    const char *names[3] = {"left.pgm", "right.pgm", "gt.pgm"};
    uint8_t *bufs[3] = {l, r, g};
    for (int i = 0; i < 3; i++) {
        snprintf(p, sizeof p, "%s/%s", out, names[i]);
        if (sgm_pgm_write(p, bufs[i], W, H))
            return 1;
    }
    printf("wrote %s/{left,right,gt}.pgm %dx%d\n", out, W, H);
    return 0;
}
