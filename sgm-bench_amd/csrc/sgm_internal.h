// sgm_internal.h -- host-side plumbing shared by the single-pair and batch
// entry points of libsgm_hip: error reporting, device workspaces, the context.
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <vector>

#include "sgm_hip.h"

// thread-local last error, read back through sgmh_last_error()
void sgmh_set_error(const char *fmt, ...);
#define set_err sgmh_set_error

#define CK(call)                                                                     \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess) {                                                      \
            set_err("HIP %s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
            return -1;                                                               \
        }                                                                            \
    } while (0)

#define REQUIRE(cond, ...)                                                           \
    do {                                                                             \
        if (!(cond)) {                                                               \
            set_err(__VA_ARGS__);                                                    \
            return -1;                                                               \
        }                                                                            \
    } while (0)

enum Phase { PH_CENSUS = 0, PH_COST, PH_AGG, PH_ARGMIN, PH_XFER, PH_N };

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

struct sgmh_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool timing = false;
    int path_layout = 0;                       // SGMH_LAYOUT_*
    std::vector<hipEvent_t> events;
    size_t ev_used = 0;
    std::vector<std::pair<int, std::pair<int, int>>> marks; // phase, ev begin, ev end
    hipStream_t last_stream = nullptr;
    int chain_ev = -1;            // end event of the last phase, which the next one on
    hipStream_t chain_s = nullptr; // the same stream starts from
    // workspaces
    DevBuf img_l, img_r, disp, disp1;          // host-buffer entry points
    DevBuf cl, cr;                             // census codes
    DevBuf vol;                                // direction volumes
    DevBuf cl1, cr1, cost1, cost2, vol1;       // multiscale stage buffers
    DevBuf queues;                             // work-queue counters of the placed path launch (zeroed once)
    int cus = 0;                               // compute units of the device
};

inline int ensure(DevBuf &b, size_t bytes)
{
    if (b.cap >= bytes)
        return 0;
    if (b.p) {
        CK(hipDeviceSynchronize());
        CK(hipFree(b.p));
        b.p = nullptr;
        b.cap = 0;
    }
    CK(hipMalloc(&b.p, bytes));
    b.cap = bytes;
    return 0;
}

inline void free_buf(DevBuf &b)
{
    if (b.p)
        (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

// Phase timing: events are recorded only when timing is on (and never
// without a context).  A phase that follows another on the same stream starts
// at that one's end event: every entry point's phases are contiguous (no
// untimed work between them on one stream), and each event record costs the
// host-buffer frame about 7 us.
struct PhaseScope {
    sgmh_ctx *ctx;
    hipStream_t s;
    int phase, ev0 = -1;
    PhaseScope(sgmh_ctx *c, hipStream_t st, int ph) : ctx(c), s(st), phase(ph)
    {
        if (ctx && ctx->timing) // no context: nothing is timed
            ev0 = ctx->chain_ev >= 0 && ctx->chain_s == s ? ctx->chain_ev : record();
    }
    int record()
    {
        if (ctx->ev_used == ctx->events.size()) {
            // timing only: no system-scope fence (a cache write-back and
            // invalidate) at each record; the entry points that hand results
            // to the host synchronise their stream, which fences
            hipEvent_t e;
            if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess)
                return -1;
            ctx->events.push_back(e);
        }
        const int i = (int)ctx->ev_used++;
        if (hipEventRecord(ctx->events[i], s) != hipSuccess)
            return -1;
        return i;
    }
    ~PhaseScope()
    {
        if (ctx && ctx->timing && ev0 >= 0) {
            const int ev1 = record();
            if (ev1 >= 0) {
                ctx->marks.push_back({phase, {ev0, ev1}});
                ctx->chain_ev = ev1;
                ctx->chain_s = s;
            }
        }
    }
};

// the windowed path's parts (sgm_hip.hip), for the learned path's entry point
int window_prepare(sgmh_ctx *ctx, int W, int H, int D);
int window_census(sgmh_ctx *ctx, hipStream_t s, const uint8_t *l, const uint8_t *r, int W, int H);
int window_rest(sgmh_ctx *ctx, hipStream_t s, int W, int H, int D, const uint8_t *blk, uint8_t *disp);

inline void begin_call(sgmh_ctx *ctx, hipStream_t s)
{
    ctx->ev_used = 0;
    ctx->marks.clear();
    ctx->last_stream = s;
    ctx->chain_ev = -1;
    ctx->chain_s = nullptr;
}

inline int check_dims(int W, int H, int D)
{
    REQUIRE(W >= 16 && H >= 16 && W <= 16384 && H <= 16384,
            "sgmh: image %dx%d outside 16..16384", W, H);
    REQUIRE(D == 16 || D == 32 || D == 64 || D == 128 || D == 256,
            "sgmh: D=%d unsupported (16, 32, 64, 128 or 256)", D);
    return 0;
}

