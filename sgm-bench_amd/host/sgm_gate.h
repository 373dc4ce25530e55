/* sgm_gate.h -- the per-phase cost gate (right answer, produced too slowly).
 *
 * Restates the behaviour of the reference's common/roofline.h:62-179: each
 * phase has a scalar-equivalent op count fixed by the workload
 *   census     2 images x W x H x census bits
 *   cost       W x H x D
 *   aggregate  paths x W x H x D x 4
 *   argmin     W x H x D
 * and measured ms / ops = ns per op is compared, phase by phase, with a
 * calibration recorded from a trusted build, keyed by implementation, board,
 * resolution and D.  A fused phase (time -1) folds its ops into aggregate; an
 * unreported aggregate fails closed; no calibration leaves the gate unarmed
 * and says so on stderr.
 */
#ifndef SGM_GATE_H
#define SGM_GATE_H
#include <stdio.h>
#include <string.h>
#include "sgm.h"

#define SGM_GATE_DEFAULT_MAX 1.15
#define SGM_GATE_PHASES 4

static const char *const sgm_gate_names[SGM_GATE_PHASES] = {"census", "cost", "aggregate", "argmin"};

static inline void sgm_gate_ops(int W, int H, const sgm_stage_times *t, double ops[SGM_GATE_PHASES],
                                double ms[SGM_GATE_PHASES])
{
    const double px = (double)W * H;
    ops[0] = 2.0 * px * SGM_CENSUS_BITS;
    ops[1] = px * SGM_D;
    ops[2] = (double)SGM_PATHS * px * SGM_D * 4.0;
    ops[3] = px * SGM_D;
    ms[0] = t->census_ms;
    ms[1] = t->cost_ms;
    ms[2] = t->aggregate_ms;
    ms[3] = t->argmin_ms;
    for (int i = 0; i < SGM_GATE_PHASES; i++)
        if (i != 2 && ms[i] < 0) {
            ops[2] += ops[i];
            ops[i] = 0;
        }
}

static inline int sgm_gate_calibrate(const char *path, const char *impl, const char *board, int W,
                                     int H, const sgm_stage_times *t)
{
    double ops[SGM_GATE_PHASES], ms[SGM_GATE_PHASES];
    sgm_gate_ops(W, H, t, ops, ms);
    FILE *f = fopen(path, "a");
    if (!f) {
        fprintf(stderr, "cannot write calibration %s\n", path);
        return 1;
    }
    for (int i = 0; i < SGM_GATE_PHASES; i++)
        if (ms[i] >= 0 && ops[i] > 0)
            fprintf(f, "%s %s %dx%d D%d %s %.6f\n", impl, board, W, H, SGM_D, sgm_gate_names[i],
                    ms[i] * 1e6 / ops[i]);
    fclose(f);
    fprintf(stderr, "cost gate: appended the %s/%s per-phase ns/op lines to %s\n", impl, board, path);
    return 0;
}

/* 0 = pass or not armed, 1 = fail; *armed says whether it ran. */
static inline int sgm_gate_check(const char *path, const char *impl, const char *board, int W, int H,
                                 const sgm_stage_times *t, double max_ratio, int *armed)
{
    double ops[SGM_GATE_PHASES], ms[SGM_GATE_PHASES], cal[SGM_GATE_PHASES];
    sgm_gate_ops(W, H, t, ops, ms);
    *armed = 0;
    if (t->aggregate_ms < 0) {
        fprintf(stderr, "ROOFLINE FAIL: no aggregate time was reported, so the phases folded into "
                        "aggregate cannot be checked (the gate fails closed).\n");
        return 1;
    }
    int any = 0;
    for (int i = 0; i < SGM_GATE_PHASES; i++)
        any |= ms[i] >= 0 && ops[i] > 0;
    if (!any) {
        fprintf(stderr, "ROOFLINE NOT ARMED: no per-stage times reported.\n");
        return 0;
    }
    FILE *f = path ? fopen(path, "r") : NULL;
    if (!f) {
        fprintf(stderr, "ROOFLINE NOT ARMED: no calibration file %s. Calibrate on a build you "
                        "trust (--roofline-calibrate).\n", path ? path : "(none)");
        return 0;
    }
    char key[96];
    snprintf(key, sizeof key, "%dx%d D%d", W, H, SGM_D);
    for (int i = 0; i < SGM_GATE_PHASES; i++)
        cal[i] = -1;
    char fi[64], fb[64], fr[64], fd[64], fp[64];
    double v;
    while (fscanf(f, "%63s %63s %63s %63s %63s %lf", fi, fb, fr, fd, fp, &v) == 6) {
        char got[160];
        snprintf(got, sizeof got, "%s %s", fr, fd);
        if (strcmp(fi, impl) || strcmp(fb, board) || strcmp(got, key))
            continue;
        for (int i = 0; i < SGM_GATE_PHASES; i++)
            if (!strcmp(fp, sgm_gate_names[i]))
                cal[i] = v;
    }
    fclose(f);
    int have = 0;
    for (int i = 0; i < SGM_GATE_PHASES; i++)
        have |= cal[i] > 0;
    if (!have) {
        fprintf(stderr, "ROOFLINE NOT ARMED: no calibration for %s/%s at %s.\n", impl, board, key);
        return 0;
    }
    *armed = 1;
    int fail = 0;
    printf("  roofline:");
    for (int i = 0; i < SGM_GATE_PHASES; i++) {
        if (ms[i] < 0 || ops[i] == 0)
            continue;
        const double nspo = ms[i] * 1e6 / ops[i];
        if (cal[i] <= 0) {
            printf("  %s %.6f (uncal)", sgm_gate_names[i], nspo);
            continue;
        }
        const double r = nspo / cal[i];
        printf("  %s %.6f/%.6f %.2fx", sgm_gate_names[i], nspo, cal[i], r);
        if (r > max_ratio) {
            fprintf(stderr, "\nROOFLINE FAIL: phase '%s' costs %.6f ns/op against a calibrated %.6f "
                            "(%.2fx, limit %.2fx); the output may still be bit-exact.\n",
                    sgm_gate_names[i], nspo, cal[i], r, max_ratio);
            fail = 1;
        }
    }
    printf("\n");
    return fail;
}
#endif
