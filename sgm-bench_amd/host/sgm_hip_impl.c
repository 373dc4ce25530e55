/* sgm_hip_impl.c -- the drop-in: SGM_IMPL = {"hip", hip_run}.
 *
 * Built against a harness's sgm.h (the reference's common/sgm.h or this
 * repository's restatement in include/), it turns one harness call
 * (common/sgm.h:29-31) into one sgmh_plain() call on GPU 0.  D and the path
 * count come from the harness's compile-time SGM_D / SGM_PATHS
 * (common/sgm_params.h:13-18).  Device state is created on the first call and
 * kept for the life of the process, as the reference's CUDA implementation
 * keeps its buffers in statics (cuda/sgm_cuda_opt.cu:292-319).
 */
#include <stdio.h>
#include "sgm.h"
#include "sgm_hip.h"

static sgmh_ctx *g_ctx;

static int hip_run(const uint8_t *left, const uint8_t *right, int W, int H, uint8_t *disp,
                   int threads, sgm_stage_times *t)
{
    (void)threads; /* the GPU has no thread count to set */
    if (!g_ctx) {
        if (sgmh_ctx_create(&g_ctx, 0) || sgmh_set_timing(g_ctx, 1)) {
            fprintf(stderr, "hip: %s\n", sgmh_last_error());
            return -1;
        }
    }
    sgmh_times ht;
    if (sgmh_plain(g_ctx, left, right, W, H, SGM_D, SGM_PATHS, disp, &ht)) {
        fprintf(stderr, "hip: %s\n", sgmh_last_error());
        return -1;
    }
    if (t) {
        t->census_ms = ht.census_ms;
        t->cost_ms = -1;       /* fused into aggregation */
        t->aggregate_ms = ht.aggregate_ms;
        t->argmin_ms = ht.argmin_ms;
        t->transfer_ms = ht.transfer_ms;
        t->threads_used = 0;
    }
    return 0;
}

/* The 8-path build names itself apart: the cost gate keys its calibration by
 * implementation, board, resolution and D, not by the path count. */
#if SGM_PATHS == 8
const sgm_impl SGM_IMPL = {"hip_p8", hip_run};
#else
const sgm_impl SGM_IMPL = {"hip", hip_run};
#endif
