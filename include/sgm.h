/* sgm.h -- the harness ABI every plain-path implementation exposes.
 *
 * Interface restatement of the reference's common/sgm.h:9-51: the struct
 * layout, typedefs and symbol names are the contract a drop-in must share, so
 * they are kept identical; the implementation in this repository is the HIP
 * adapter sgm-bench_amd/host/sgm_hip_impl.c.
 */
#ifndef SGM_H
#define SGM_H
#include <stddef.h>
#include <stdint.h>
#include "sgm_params.h"

/* Per-stage times in ms, -1 when not applicable.  cost_ms / argmin_ms are -1
 * when fused; transfer_ms is the host<->device copy time kept apart from the
 * phases; threads_used is the team size observed inside a parallel region
 * (0 when not applicable). */
typedef struct {
    double census_ms;
    double cost_ms;
    double aggregate_ms;
    double argmin_ms;
    double transfer_ms;
    int threads_used;
} sgm_stage_times;

/* One frame: left/right W*H uint8, disp W*H uint8 out; threads 0 = default.
 * Returns 0 on success. */
typedef int (*sgm_run_fn)(const uint8_t *left, const uint8_t *right, int W, int H,
                          uint8_t *disp, int threads, sgm_stage_times *t);

typedef struct {
    const char *name;
    sgm_run_fn run;
} sgm_impl;

/* exactly one per binary */
extern const sgm_impl SGM_IMPL;

uint8_t *pgm_read(const char *path, int *W, int *H);
int pgm_write(const char *path, const uint8_t *img, int W, int H);
uint64_t fnv1a64(const void *data, size_t n);
double now_ms(void);

#endif
