/* sgm_hip.h -- the C ABI of libsgm_hip.so: Semi-Global Matching on one AMD
 * Instinct MI355X (gfx950) GPU.
 *
 * Plain pointers and sizes only; no HIP or torch types cross this boundary
 * (streams are passed as void*, a hipStream_t).  Every entry point returns 0 on
 * success and -1 on failure; sgmh_last_error() then names the failing call with
 * file:line, the way the reference's CK() macro reports CUDA errors
 * (cuda/sgm_cuda_opt.cu:67-69).
 *
 * The reference exposes exactly one drop-in interface for this path: the
 * harness ABI of common/sgm.h:27-40
 *
 *     int run(const uint8_t *left, const uint8_t *right, int W, int H,
 *             uint8_t *disp, int threads, sgm_stage_times *t);
 *
 * with D and the path count fixed at compile time (common/sgm_params.h:13-18),
 * plus the standalone multiscale binaries (multiscale/sgm_ms_ref.c:151-224).
 * This library is what such an implementation calls; the one-line SGM_IMPL
 * adapter is sgm-bench_amd/host/sgm_hip_impl.c.  D and the path count are
 * runtime arguments here.
 *
 * Semantics (bit-exact to the reference oracles in every call):
 *   plain       common/sgm_params.h:44-59 and common/sgm_ref.c:101-144:
 *               9x7 census, Hamming cost (x-d<0 costs 62), P1=20, P2=192,
 *               paths 4 (L->R, R->L, T->B, B->T) or 8 (+ diagonals), uint16
 *               path sums, argmin with the lowest d winning ties.
 *   multiscale  multiscale/sgm_ms_ref.c:47-202: two scales, Sobel-X prefilter,
 *               invalid 63, cost merge with the half-resolution cost at
 *               (x/2, y/2, d/2), 8 direction-weighted saturating paths
 *               (P1 40/20/10, P2 200), output (W/2)x(H/2) decimated argmin.
 *
 * Constraints: 16 <= W, H <= 16384; D in {16, 32, ..., 256} (a multiple of 16,
 * as common/sgm_params.h:37-39 requires); multiscale also needs even W and H
 * (the reference's merge assumes it, sgm_ms_ref.c:184-188).
 */
#ifndef SGM_HIP_H
#define SGM_HIP_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sgmh_ctx sgmh_ctx;

/* Per-phase times in ms of the most recent call, measured with HIP events on
 * the stream the kernels ran on (-1 = phase not run).  Mirrors the reference's
 * sgm_stage_times (common/sgm.h:9-25); cost is fused into aggregation
 * (cost_ms = -1) on the plain path, as in cuda/sgm_cuda_opt.cu:362. */
typedef struct {
    double census_ms;
    double cost_ms;
    double aggregate_ms;
    double argmin_ms;
    double transfer_ms;
    double total_ms;
} sgmh_times;

/* Create / destroy a context bound to one device.  A context owns a stream and
 * device workspaces that grow to the largest frame seen and are reused (the
 * reference caches its device state in statics, cuda/sgm_cuda_opt.cu:292-319).
 * A context is not thread-safe; calls on one context must be sequential. */
int sgmh_ctx_create(sgmh_ctx **out, int device);
void sgmh_ctx_destroy(sgmh_ctx *ctx);
const char *sgmh_last_error(void);

/* Per-phase HIP-event timing on (1) or off (0, default). */
int sgmh_set_timing(sgmh_ctx *ctx, int on);
/* Wait for the last call and read its phase times. */
int sgmh_get_times(sgmh_ctx *ctx, sgmh_times *t);

/* No counterpart in the reference (the reference's kernels have one layout):
 * how this context's plain path kernel lays scanlines on a wave64.  Every
 * layout gives the same output.
 *   SGMH_LAYOUT_AUTO   (default) picks by frame size and D;
 *   SGMH_LAYOUT_LINES4 four scanlines per wave, 16 lanes each, D/16
 *                      disparities per lane;
 *   SGMH_LAYOUT_WAVE   one scanline per wave, D/64 disparities per lane
 *                      (D = 64, 128 or 256; other D fail at the call);
 *   SGMH_LAYOUT_PLACED four scanlines per wave, one workgroup per compute
 *                      unit whose waves take row or column line groups by
 *                      the SIMD they run on (4 paths, D = 32, 64 or 128;
 *                      other cases take the four-line launch).
 * Returns -1 for an unknown layout. */
#define SGMH_LAYOUT_AUTO 0
#define SGMH_LAYOUT_LINES4 1
#define SGMH_LAYOUT_WAVE 2
#define SGMH_LAYOUT_PLACED 3
int sgmh_set_path_layout(sgmh_ctx *ctx, int layout);

/* ---- plain SGM: replaces the body of sgm_run_fn (common/sgm.h:29-31) ---- */

/* Host buffers, synchronous: disp is complete on return (common/sgm.h:27-31,
 * the harness times the whole call including H2D/D2H, harness.c:154-158). */
int sgmh_plain(sgmh_ctx *ctx, const uint8_t *left, const uint8_t *right,
               int W, int H, int D, int paths, uint8_t *disp, sgmh_times *t);

/* Device buffers, asynchronous on `stream` (a hipStream_t; NULL is the HIP
 * default stream, as in the HIP API). */
int sgmh_plain_dev(sgmh_ctx *ctx, const uint8_t *d_left, const uint8_t *d_right,
                   int W, int H, int D, int paths, uint8_t *d_disp, void *stream);

/* No counterpart in the reference: sgmh_plain_dev for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same semantics and output as
 * sgmh_plain_dev, with no context: the census codes and direction volumes live
 * in one static device array of SGMH_NOALLOC_SCRATCH_BYTES per GPU (each GPU
 * the library runs on holds it until the process ends), resolved on the
 * current device (hipSetDevice) at the first call there or at
 * sgmh_ctx_create.  It needs 2 * align256((W*H + 256) * 8) + paths*W*H*D bytes
 * (640x480 D=64 4-path: 83.6 MB) and returns -1 for frames that need more.
 * It calls no hipMalloc / hipFree.  Every call on a GPU shares that GPU's
 * array, so calls that overlap on one GPU (other streams, other threads)
 * race; calls on one stream are ordered and safe. */
#define SGMH_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
int sgmh_plain_noalloc_dev(const uint8_t *d_left, const uint8_t *d_right, int W, int H, int D,
                           int paths, uint8_t *d_disp, void *stream);

/* ---- multiscale (Configurations B/C): replaces sgm_ms_ref.c:166-202 ---- */

/* disp is (W/2)*(H/2).  disp1 (may be NULL) receives the stage-1 half-resolution
 * argmin that the reference computes and discards (sgm_ms_ref.c:175-180). */
int sgmh_ms(sgmh_ctx *ctx, const uint8_t *left, const uint8_t *right,
            int W, int H, int D, uint8_t *disp, uint8_t *disp1, sgmh_times *t);
int sgmh_ms_dev(sgmh_ctx *ctx, const uint8_t *d_left, const uint8_t *d_right,
                int W, int H, int D, uint8_t *d_disp, uint8_t *d_disp1, void *stream);

/* ---- single phases, device buffers (for descriptor- and volume-level
 *      parity against the oracle, the reference's common/test_phase1.c
 *      pattern) ---- */

/* 9x7 census of one image into W*H uint64 codes (sgm_ref.c:23-41). */
int sgmh_census_dev(sgmh_ctx *ctx, const uint8_t *d_img, int W, int H,
                    uint64_t *d_codes, void *stream);
/* Sobel-X prefilter (sgm_ms_ref.c:47-54) and 2x2 half (sgm_ms_ref.c:145-150). */
int sgmh_sobel_dev(sgmh_ctx *ctx, const uint8_t *d_img, int W, int H,
                   uint8_t *d_out, void *stream);
int sgmh_half_dev(sgmh_ctx *ctx, const uint8_t *d_img, int W, int H,
                  uint8_t *d_out, void *stream);
/* The uint8 path-cost volume L[(y*W+x)*D+d] of ONE direction (dx, dy), from
 * census codes, plain-path rules (P1=20, P2=192, invalid 62, no saturation);
 * the oracle's orc_path_L over orc_cost (sgm_ref.c:44-94). */
int sgmh_path_plain_dev(sgmh_ctx *ctx, const uint64_t *d_cl, const uint64_t *d_cr,
                        int W, int H, int D, int dx, int dy, uint8_t *d_L, void *stream);

/* ---- windowed SGM: the learned path's reduced search (no reference
 *      counterpart).  Plain-path rules (4 paths, P1=20, P2=192, invalid 62)
 *      restricted per 16x16 tile to the 16-disparity blocks
 *      [blk[t] & 15, blk[t] >> 4]; tiles are row-major, ceil(W/16) per row.
 *      With every block active the output equals sgmh_plain's.  D = 32, 64 or
 *      128. ---- */
int sgmh_window(sgmh_ctx *ctx, const uint8_t *left, const uint8_t *right, int W, int H, int D,
                const uint8_t *blk, uint8_t *disp, sgmh_times *t);
int sgmh_window_dev(sgmh_ctx *ctx, const uint8_t *d_left, const uint8_t *d_right, int W, int H,
                    int D, const uint8_t *d_blk, uint8_t *d_disp, void *stream);

/* ---- the learned path's window-proposal network (no reference
 *      counterpart): 4x-downsampled matching features, 6 conv layers
 *      3x3 32->32 (bf16 MFMA, fp32 accumulation), per low-resolution pixel a
 *      softmax over 32 disparity bins of 4 (D = 128); per 16x16 tile the mean
 *      of its pixels' distributions, the 32-disparity window (two 16-blocks)
 *      with the most mass, that mass as the confidence.  Frames need W and H
 *      multiples of 4 and 4-byte aligned images.  A network owns its parameters
 *      (fp32 master copy, bf16 packed copies), Adam state and activation
 *      buffers. ---- */
typedef struct sgmh_cnn sgmh_cnn;
size_t sgmh_cnn_num_params(void);
int sgmh_cnn_create(sgmh_ctx *ctx, sgmh_cnn **out);
void sgmh_cnn_destroy(sgmh_cnn *net);
/* fp32 parameters, per layer W[co][ci][tap] (tap = 3*ky + kx) then b[co];
 * setting them also resets the optimiser state */
int sgmh_cnn_set_params(sgmh_cnn *net, const float *theta, size_t n);
int sgmh_cnn_get_params(sgmh_cnn *net, float *theta, size_t n);
/* Per-tile block ranges for sgmh_window_dev (b0 | b1 << 4, ceil(W/16) per row):
 * the best window where its confidence >= thresh, else all 8 blocks. */
int sgmh_cnn_propose_dev(sgmh_cnn *net, const uint8_t *d_left, const uint8_t *d_right, int W, int H,
                         float thresh, uint8_t *d_blk, float *d_conf, void *stream);
/* The learned frame in one call: proposal (at `thresh`) and the windowed SGM
 * over it at D = 128, the census running on a side stream beside the
 * network.  Same disparities as sgmh_cnn_propose_dev + sgmh_window_dev;
 * d_blk (may be NULL) receives the block table. */
int sgmh_learned_dev(sgmh_cnn *net, const uint8_t *d_left, const uint8_t *d_right, int W, int H,
                     float thresh, uint8_t *d_blk, uint8_t *d_disp, void *stream);
/* One Adam step (lr > 0) on one frame against target disparities (the plain
 * full-D SGM output); lr <= 0 computes the gradients only.  *loss (may be NULL)
 * receives the per-pixel cross-entropy against the histograms of the target's
 * 4x4 blocks, before the step. */
int sgmh_cnn_train_step_dev(sgmh_cnn *net, const uint8_t *d_left, const uint8_t *d_right,
                            const uint8_t *d_target, int W, int H, float lr, float *loss, void *stream);
int sgmh_cnn_get_grads(sgmh_cnn *net, float *grad, size_t n);
/* features (bf16 bits, [H/4][W/4][32]) and logits (fp32, same shape) of the
 * last forward pass; either may be NULL */
int sgmh_cnn_get_activations(sgmh_cnn *net, uint16_t *features, float *logits);

/* ---- batched mode: many pairs of uneven size, D and configuration in one
 *      submission (no reference counterpart; the reference runs one pair per
 *      binary invocation, common/harness.c:150-158) ---- */

#define SGMH_PLAIN 0      /* plain path, sgmh_plain semantics */
#define SGMH_MULTISCALE 1 /* multiscale path, sgmh_ms semantics */

typedef struct {
    const uint8_t *d_left, *d_right; /* device, W*H each */
    int W, H, D;
    int mode;                        /* SGMH_PLAIN or SGMH_MULTISCALE */
    int paths;                       /* plain: 4 or 8; multiscale: 8 */
    uint8_t *d_disp;                 /* device; W*H (plain) or (W/2)*(H/2) */
} sgmh_pair;

typedef struct sgmh_batch sgmh_batch;

#define SGMH_BATCH_AUTOTUNE 1 /* time candidate schedules on this batch, keep the best */
#define SGMH_BATCH_NO_GRAPH 2 /* launch the phase kernels directly, no hipGraph */
/* Without SGMH_BATCH_AUTOTUNE, these pin the schedule (the first two are
 * choices the autotuner also makes; FUSED and PIPELINED exclude each other): */
#define SGMH_BATCH_CENSUS_32X32 4 /* census in 32x32 tiles instead of 64x16 */
#define SGMH_BATCH_XCD_SLICES 8   /* path line groups taken from per-XCD slices of the queue */
#define SGMH_BATCH_FUSED 16       /* path aggregation and sum + argmin in one launch, each pair's
                                     argmin as soon as its directions are done (ignores XCD_SLICES) */
#define SGMH_BATCH_PIPELINED 32   /* three groups of pairs pipelined over three streams: one group's
                                     argmin and the next's cost volumes beside the path launches
                                     (the autotuner also times 3, 4 and 6 groups) */

/* Plan a batch: allocate every pair's workspace, build the work lists of the
 * persistent per-phase kernels (pre-filter / census, cost, path aggregation,
 * sum + argmin), optionally autotune the schedule (lane partition or pipelined
 * groups, path item order, path waves per SIMD, census tile shape, scanline
 * grouping), and capture the phases as one hipGraph.  The device buffers
 * named in `pairs` are bound into the plan: later runs read whatever they
 * then hold. */
int sgmh_batch_plan(sgmh_ctx *ctx, const sgmh_pair *pairs, int n, int flags, sgmh_batch **out);
/* Run the planned batch on `stream` (NULL: the HIP default stream), asynchronous. */
int sgmh_batch_run(sgmh_batch *b, void *stream);
/* The comparison baseline: the same pairs, one sgmh_plain_dev / sgmh_ms_dev
 * call after another, on `stream`. */
int sgmh_batch_run_loop(sgmh_batch *b, void *stream);
/* JSON description of the plan: pairs, items per phase, chosen schedule and
 * the autotuner's timings, written into buf (at most len-1 characters and a
 * terminating NUL).  Returns the description's full length, as snprintf
 * does: a value >= len means it was truncated; -1 on a NULL argument. */
int sgmh_batch_describe(const sgmh_batch *b, char *buf, size_t len);
void sgmh_batch_destroy(sgmh_batch *b);

/* Streaming copy of `bytes` (a multiple of 16) from src to dst on the device:
 * the achievable-bandwidth ceiling and the known-size reference kernel for
 * hardware byte counters (the reference's tools/gpu_bandwidth.cu:12-25). */
int sgmh_copy_probe(sgmh_ctx *ctx, void *d_dst, const void *d_src, size_t bytes, void *stream);

#ifdef __cplusplus
}
#endif
#endif
