/* harness.c -- timing / golden / cost-gate driver for one SGM_IMPL.
 *
 * Same command line, output line, JSON record and exit codes as the
 * reference's common/harness.c (:3-23, :197-237), so a run of the HIP build is
 * read by the same scripts:
 *
 *   sgm_hip LEFT.pgm RIGHT.pgm [-o OUT.pgm] [-g GOLDEN.pgm] [-t THREADS]
 *           [-w WARM] [-n TIMED] [-j RESULTS.json] [--board NAME]
 *           [--no-roofline] [--roofline-max X] [--roofline-cal FILE]
 *           [--roofline-calibrate] [--max-spread X]
 *
 * Exit 0 ok, 1 usage/IO/implementation failure, 2 golden mismatch (wrong
 * answer), 3 cost gate failed (right answer, a phase far over its calibrated
 * ns per scalar-equivalent op).
 */
#define _GNU_SOURCE
#include <math.h>
#include <sched.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "sgm.h"
#include "sgm_gate.h"

#ifndef CFLAGS_STR
#define CFLAGS_STR "unknown"
#endif
#ifndef GIT_SHA
#define GIT_SHA "unknown"
#endif

static int by_value(const void *a, const void *b)
{
    const double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

/* CPUs this process may run on (the HIP build has no thread team). */
static void affinity_list(char *buf, size_t n)
{
    cpu_set_t set;
    CPU_ZERO(&set);
    size_t len = 0;
    buf[0] = 0;
    if (sched_getaffinity(0, sizeof set, &set) == 0)
        for (int c = 0; c < CPU_SETSIZE && len + 8 < n; c++)
            if (CPU_ISSET(c, &set))
                len += (size_t)snprintf(buf + len, n - len, "%s%d", len ? "," : "", c);
    if (!len)
        snprintf(buf, n, "?");
}

static void usage(const char *argv0)
{
    fprintf(stderr, "usage: %s LEFT.pgm RIGHT.pgm [-o out.pgm] [-g golden.pgm] [-t threads] "
                    "[-w warm] [-n timed] [-j results.json] [--board name] [--no-roofline] "
                    "[--roofline-max X] [--roofline-cal F] [--roofline-calibrate] "
                    "[--max-spread X]\n", argv0);
}

int main(int argc, char **argv)
{
    if (argc < 3) {
        usage(argv[0]);
        return 1;
    }
    const char *lp = argv[1], *rp = argv[2], *op = NULL, *gp = NULL, *jp = NULL;
    const char *board = "unknown", *cal = "data/golden/roofline.cal";
    int threads = 0, warm = 10, timed = 100, gate = 1, calibrate = 0;
    double gate_max = SGM_GATE_DEFAULT_MAX, max_spread = 1.25;
    for (int i = 3; i < argc; i++) {
        const int more = i + 1 < argc;
        if (!strcmp(argv[i], "-o") && more) op = argv[++i];
        else if (!strcmp(argv[i], "-g") && more) gp = argv[++i];
        else if (!strcmp(argv[i], "-t") && more) threads = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-w") && more) warm = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-n") && more) timed = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-j") && more) jp = argv[++i];
        else if (!strcmp(argv[i], "--board") && more) board = argv[++i];
        else if (!strcmp(argv[i], "--no-roofline")) gate = 0;
        else if (!strcmp(argv[i], "--roofline-max") && more) gate_max = atof(argv[++i]);
        else if (!strcmp(argv[i], "--roofline-cal") && more) cal = argv[++i];
        else if (!strcmp(argv[i], "--roofline-calibrate")) calibrate = 1;
        else if (!strcmp(argv[i], "--max-spread") && more) max_spread = atof(argv[++i]);
        else {
            fprintf(stderr, "unknown arg %s\n", argv[i]);
            return 1;
        }
    }
    if (threads > 0)
        fprintf(stderr, "WARNING: -t %d ignored, the GPU implementation has no thread count\n", threads);
    char mask[512];
    affinity_list(mask, sizeof mask);

    int W, H, W2, H2;
    uint8_t *L = pgm_read(lp, &W, &H), *R = pgm_read(rp, &W2, &H2);
    if (!L || !R)
        return 1;
    if (W != W2 || H != H2) {
        fprintf(stderr, "left/right size mismatch\n");
        return 1;
    }
    uint8_t *disp = calloc((size_t)W * H, 1);
    double *ms = malloc(sizeof(double) * (timed > 0 ? (size_t)timed : 1));
    if (!disp || !ms)
        return 1;
    sgm_stage_times st = {-1, -1, -1, -1, -1, 0};
    for (int i = 0; i < warm; i++)
        if (SGM_IMPL.run(L, R, W, H, disp, threads, &st)) {
            fprintf(stderr, "impl failed\n");
            return 1;
        }
    int n = timed > 0 ? timed : 1;
    for (int i = 0; i < n; i++) {
        const double t0 = now_ms();
        if (SGM_IMPL.run(L, R, W, H, disp, threads, &st)) {
            fprintf(stderr, "impl failed\n");
            return 1;
        }
        ms[i] = now_ms() - t0;
    }
    qsort(ms, (size_t)n, sizeof(double), by_value);
    const double med = ms[n / 2], mn = ms[0];
    int p95i = (n * 95 + 99) / 100 - 1; /* ceil(0.95 n) - 1 */
    p95i = p95i < 0 ? 0 : p95i >= n ? n - 1 : p95i;
    const double p95 = ms[p95i];
    double ss = 0;
    for (int i = 0; i < n; i++)
        ss += (ms[i] - med) * (ms[i] - med);
    const double rmsd = n > 1 ? sqrt(ss / (n - 1)) : 0; /* RMS deviation about the median */
    const double hi = med > 0 ? p95 / med : 0, lo = med > 0 ? mn / med : 0;
    const int unstable = hi > max_spread;
    const double fps = 1000.0 / med, mpix = (double)W * H / 1e6 * fps;

    const uint64_t hash = fnv1a64(disp, (size_t)W * H);
    int golden = -1;
    if (gp) {
        int gw, gh;
        uint8_t *G = pgm_read(gp, &gw, &gh);
        if (!G)
            return 1;
        if (gw != W || gh != H) {
            golden = 0;
            fprintf(stderr, "GOLDEN MISMATCH: golden is %dx%d, output %dx%d\n", gw, gh, W, H);
        } else {
            size_t diff = 0, first = 0;
            for (size_t i = 0; i < (size_t)W * H; i++)
                if (G[i] != disp[i] && diff++ == 0)
                    first = i;
            golden = diff == 0;
            if (diff)
                fprintf(stderr, "GOLDEN MISMATCH: %zu pixels differ, first at (%zu,%zu) got %d want %d\n",
                        diff, first % (size_t)W, first / (size_t)W, disp[first], G[first]);
        }
        free(G);
    }
    if (op && pgm_write(op, disp, W, H))
        return 1;
    printf("%-12s %dx%d D=%d paths=%d th=%d  median %8.2f ms  p95 %8.2f  min %8.2f  sd %6.3f  "
           "spread %.2f/%.2f%s  fps %7.2f  Mpix/s %7.2f  hash %016llx%s\n",
           SGM_IMPL.name, W, H, SGM_D, SGM_PATHS, st.threads_used > 0 ? st.threads_used : 1, med, p95,
           mn, rmsd, hi, lo, unstable ? " UNSTABLE" : "", fps, mpix, (unsigned long long)hash,
           golden == 1 ? "  GOLDEN OK" : golden == 0 ? "  GOLDEN FAIL" : "");

    int gate_fail = 0, armed = 0;
    if (calibrate)
        sgm_gate_calibrate(cal, SGM_IMPL.name, board, W, H, &st);
    else if (gate)
        gate_fail = sgm_gate_check(cal, SGM_IMPL.name, board, W, H, &st, gate_max, &armed);

    if (jp) {
        FILE *f = fopen(jp, "a");
        if (!f) {
            fprintf(stderr, "cannot open %s\n", jp);
            return 1;
        }
        char ts[64];
        const time_t now = time(NULL);
        strftime(ts, sizeof ts, "%Y-%m-%dT%H:%M:%SZ", gmtime(&now));
        fprintf(f,
                "{\"impl\":\"%s\",\"board\":\"%s\",\"W\":%d,\"H\":%d,\"D\":%d,\"paths\":%d,"
                "\"census\":\"%dx%d\",\"P1\":%d,\"P2\":%d,\"threads\":%d,\"threads_requested\":%d,"
                "\"cpu_mask\":\"%s\",\"warm\":%d,\"timed\":%d,\"median_ms\":%.3f,\"p95_ms\":%.3f,"
                "\"min_ms\":%.3f,\"rmsd_ms\":%.4f,\"spread_hi\":%.3f,\"spread_lo\":%.3f,"
                "\"unstable\":%d,\"transfer_ms\":%.3f,\"threads_used\":%d,\"fps\":%.3f,"
                "\"mpix_s\":%.3f,\"census_ms\":%.3f,\"cost_ms\":%.3f,\"aggregate_ms\":%.3f,"
                "\"argmin_ms\":%.3f,\"hash\":\"%016llx\",\"golden_match\":%d,\"roofline_ok\":%d,"
                "\"cflags\":\"%s\",\"git\":\"%s\",\"ts\":\"%s\"}\n",
                SGM_IMPL.name, board, W, H, SGM_D, SGM_PATHS, SGM_CENSUS_W, SGM_CENSUS_H, SGM_P1,
                SGM_P2, st.threads_used > 0 ? st.threads_used : 1, threads, mask, warm, n, med, p95,
                mn, rmsd, hi, lo, unstable, st.transfer_ms, st.threads_used, fps, mpix,
                st.census_ms, st.cost_ms, st.aggregate_ms, st.argmin_ms,
                (unsigned long long)hash, golden, armed ? !gate_fail : -1, CFLAGS_STR, GIT_SHA, ts);
        fclose(f);
    }
    free(L);
    free(R);
    free(disp);
    free(ms);
    if (golden == 0)
        return 2;
    return gate_fail ? 3 : 0;
}
