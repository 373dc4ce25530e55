// sgm_batch.hip -- batched mode: many stereo pairs of uneven size, D and
// configuration in one submission.
//
// A plan turns the pairs into work lists for four persistent phase kernels,
// each of which serves every pair at once:
//   k_batch_pre   census tiles (plain), and for multiscale pairs the half
//                 images, Sobel-X images and census tiles of both scales
//                 (three launches, the dependency levels of the pre-filters)
//   k_batch_cost  multiscale cost volumes (two launches: half scale, then the
//                 full scale merged with it)
//   k_batch_path  path aggregation: one item is one group of 4 scanlines of
//                 one direction of one pair; waves pull items from a global
//                 counter, longest lines first by default
//   k_batch_wta   sum over directions + argmin, all pairs and both
//                 multiscale stages
// The launches are captured once into a hipGraph; a run is one graph launch.
// The autotuner times candidate schedules (persistent waves per SIMD for the
// path phase, item order) on the batch itself and keeps the fastest.
// SGMH_BATCH_FUSED instead runs path and argmin in one launch, k_batch_fused,
// which starts each pair's argmin as soon as its directions are complete.
//
// Every pair is computed by the same device functions as the single-pair
// entry points (sgm_device.h), so each stays bit-exact to its oracle.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <string>
#include <vector>

#include "sgm_device.h"
#include "sgm_internal.h"

using namespace sgmk;

namespace {

enum PreKind { PRE_HALF = 0, PRE_SOBEL = 1, PRE_CENSUS = 2 };

struct PreTask {
    const uint8_t *src;
    void *dst;       // uint8_t image (half, Sobel) or uint64_t codes (census)
    int W, H;        // source image size
    int kind;
    int ostride;     // census row stride in codes
    int tiles_x;     // tiles per row
    int first_item;  // first item of this task in the launch
};

struct CostTask {
    const uint64_t *cl, *cr;
    const uint8_t *merge; // nullptr: half scale; else the half-scale volume
    uint8_t *C;
    int W, H, D;
    uint32_t invalid;
    int first_item;
};

// Field names match PathArgs so the path device functions accept it.
struct BatchPathTask {
    const uint64_t *cl, *cr;
    const uint8_t *cost;
    int W, H, D;
    uint32_t P2, invalid;
    uint32_t out_bytes;
    DirTask dir;
    int variant; // kVar* below
    int dep;     // the sum + argmin task this direction feeds
};

enum {
    kVarPlain32 = 0, kVarPlain64, kVarPlain128,
    kVarMs32, kVarMs64, kVarMs128,          // multiscale stage 1 (full-layout store)
    kVarMsDec32, kVarMsDec64, kVarMsDec128, // multiscale stage 2 (decimated store)
};

struct WtaTask {
    const uint8_t *L;
    size_t dir_stride;
    int ndirs, npix, D;
    uint8_t *disp;
    int first_item;
    int units; // fused launch: claimable units of kWtaUnit wta_batch calls
};

// The items of a launch are split into contiguous chunks, one per block (or
// wave); a chunk crosses few task boundaries, so the task is found once and
// then advanced as the item index passes the next task's first item.
template <class T>
__device__ __forceinline__ int find_task(const T *__restrict__ tasks, int ntasks, int item)
{
    int lo = 0, hi = ntasks - 1; // last task whose first_item <= item
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tasks[mid].first_item <= item)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

template <class T>
struct TaskCursor {
    const T *tasks;
    int ntasks, t, next; // current task, first item of the one after it
    T cur;
    __device__ void seek(int item)
    {
        t = find_task(tasks, ntasks, item);
        cur = tasks[t];
        next = t + 1 < ntasks ? tasks[t + 1].first_item : 0x7FFFFFFF;
    }
    __device__ void advance(int item)
    {
        while (item >= next) {
            t++;
            cur = tasks[t];
            next = t + 1 < ntasks ? tasks[t + 1].first_item : 0x7FFFFFFF;
        }
    }
};

__device__ __forceinline__ void chunk_of(int unit, int nunits, int nitems, int &begin, int &end)
{
    const int per = (nitems + nunits - 1) / nunits;
    begin = min(nitems, unit * per);
    end = min(nitems, begin + per);
}


// Census tile shapes the autotuner picks between (TX x TY output pixels).
constexpr int kCenShapes = 2;
constexpr int kCenShapeW[kCenShapes] = {64, 32}, kCenShapeH[kCenShapes] = {16, 32};

template <int TX, int TY>
__global__ __launch_bounds__(256) void k_batch_pre(const PreTask *__restrict__ tasks, int ntasks, int nitems)
{
    __shared__ __align__(16) uint32_t tile[kCenTileWords];
    int begin, end;
    chunk_of(blockIdx.x, gridDim.x, nitems, begin, end);
    if (begin >= end)
        return;
    TaskCursor<PreTask> c{tasks, ntasks};
    c.seek(begin);
    for (int item = begin; item < end; item++) {
        c.advance(item);
        const PreTask &tk = c.cur;
        const int local = item - tk.first_item;
        const int tx = local % tk.tiles_x, ty = local / tk.tiles_x;
        if (tk.kind == PRE_CENSUS) {
            census_tile<TX, TY>(tk.src, (uint64_t *)tk.dst, tk.W, tk.H, tk.ostride, tx * TX, ty * TY, tile);
        } else {
            const int x = tx * 64 + (threadIdx.x & 63), y = ty * 4 + (threadIdx.x >> 6);
            if (tk.kind == PRE_HALF) {
                if (x < tk.W / 2 && y < tk.H / 2)
                    half_px(tk.src, tk.W, (uint8_t *)tk.dst, tk.W / 2, x, y);
            } else if (x < tk.W && y < tk.H) {
                sobel_px(tk.src, tk.W, tk.H, (uint8_t *)tk.dst, x, y);
            }
        }
    }
}

// one item = 16 pixels: 4 waves x 4 sixteen-lane rows
__global__ __launch_bounds__(256) void k_batch_cost(const CostTask *__restrict__ tasks, int ntasks, int nitems)
{
    const int lane = threadIdx.x & 63, g = lane >> 4, l = lane & 15;
    int begin, end;
    chunk_of(blockIdx.x, gridDim.x, nitems, begin, end);
    if (begin >= end)
        return;
    TaskCursor<CostTask> c{tasks, ntasks};
    c.seek(begin);
    for (int item = begin; item < end; item++) {
        c.advance(item);
        const CostTask &tk = c.cur;
        const uint32_t p = (uint32_t)(item - tk.first_item) * 16u + (threadIdx.x >> 6) * 4u + (uint32_t)g;
        if (p >= (uint32_t)(tk.W * tk.H))
            continue;
        const int y = (int)(p / (uint32_t)tk.W), x = (int)(p - (uint32_t)y * (uint32_t)tk.W);
        switch (tk.D) {
        case 32: cost_px<2>(tk.cl, tk.cr, tk.W, tk.H, tk.D, tk.invalid, tk.merge, tk.C, x, y, l); break;
        case 64: cost_px<4>(tk.cl, tk.cr, tk.W, tk.H, tk.D, tk.invalid, tk.merge, tk.C, x, y, l); break;
        default: cost_px<8>(tk.cl, tk.cr, tk.W, tk.H, tk.D, tk.invalid, tk.merge, tk.C, x, y, l); break;
        }
    }
}

__device__ __forceinline__ void batch_path_item(const BatchPathTask *__restrict__ tasks, uint32_t code)
{
    const BatchPathTask tk = tasks[code >> 16];
    const int grp = (int)(code & 0xFFFFu);
    switch (tk.variant) {
    case kVarPlain32: path_group<2, true, false, false>(tk, tk.dir, grp); break;
    case kVarPlain64: path_group<4, true, false, false>(tk, tk.dir, grp); break;
    case kVarPlain128: path_group<8, true, false, false>(tk, tk.dir, grp); break;
    case kVarMs32: path_group<2, false, true, false>(tk, tk.dir, grp); break;
    case kVarMs64: path_group<4, false, true, false>(tk, tk.dir, grp); break;
    case kVarMs128: path_group<8, false, true, false>(tk, tk.dir, grp); break;
    case kVarMsDec32: path_group<2, false, true, true>(tk, tk.dir, grp); break;
    case kVarMsDec64: path_group<4, false, true, true>(tk, tk.dir, grp); break;
    default: path_group<8, false, true, true>(tk, tk.dir, grp); break;
    }
    // a row sets its wave's issue priority for the row alone
    __builtin_amdgcn_s_setprio(0);
}

// Scanline groupings of the path work list the autotuner picks between:
//   global   one queue, items taken in queue order by whichever wave is free;
//   per-XCD  the queue cut into kXcdSlices contiguous slices, one per XCD:
//            a wave takes items from its own XCD's slice, then from the
//            others in turn, so neighbouring line groups (whose disparity
//            windows overlap) are mostly processed under one L2.
constexpr int kXcdSlices = 8;
constexpr int kCounterStride = 16; // ints: one 64-byte line per counter
constexpr int kCounters = kXcdSlices + 1;

__global__ __launch_bounds__(64) void k_batch_reset(int *counter)
{
    if (threadIdx.x < kCounters)
        counter[threadIdx.x * kCounterStride] = 0;
}

// One wave per block; items are (task << 16 | line group) in queue order.
__global__ __launch_bounds__(64) void k_batch_path(const BatchPathTask *__restrict__ tasks,
                                                   const uint32_t *__restrict__ items, int nitems,
                                                   int *__restrict__ counter)
{
    for (;;) {
        int it = 0;
        if (threadIdx.x == 0)
            it = atomicAdd(counter, 1);
        it = __builtin_amdgcn_readfirstlane(__shfl(it, 0));
        if (it >= nitems)
            break;
        batch_path_item(tasks, items[it]);
    }
}

__global__ __launch_bounds__(64) void k_batch_path_xcd(const BatchPathTask *__restrict__ tasks,
                                                       const uint32_t *__restrict__ items, int nitems,
                                                       int *__restrict__ counter)
{
    // hwreg(HW_REG_XCC_ID, 0, 4): the XCD this wave runs on
    const int xcd = (int)(__builtin_amdgcn_s_getreg((3 << 11) | 20) & (kXcdSlices - 1));
    for (int v = 0; v < kXcdSlices; v++) {
        const int q = (xcd + v) & (kXcdSlices - 1);
        const int lo = (int)((long)nitems * q / kXcdSlices), hi = (int)((long)nitems * (q + 1) / kXcdSlices);
        for (;;) {
            int it = 0;
            if (threadIdx.x == 0)
                it = atomicAdd(counter + (1 + q) * kCounterStride, 1);
            it = lo + __builtin_amdgcn_readfirstlane(__shfl(it, 0));
            if (it >= hi)
                break;
            batch_path_item(tasks, items[it]);
        }
    }
}

// The lane index through an empty asm statement: lane arithmetic derived from
// it is recomputed where it is used rather than hoisted out of a persistent
// loop, where it would hold registers across everything else the loop runs.
__device__ __forceinline__ int opaque_lane()
{
    int l = threadIdx.x & 63;
    asm volatile("" : "+v"(l));
    return l;
}

// one item = one wta_batch of one task (64/LPP * kWtaBatch pixels); each wave
// takes a contiguous chunk of items
__global__ __launch_bounds__(256) void k_batch_wta(const WtaTask *__restrict__ tasks, int ntasks, int nitems)
{
    const int wave = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    int begin, end;
    chunk_of(wave, (int)gridDim.x * 4, nitems, begin, end);
    if (begin >= end)
        return;
    TaskCursor<WtaTask> c{tasks, ntasks};
    c.seek(begin);
    for (int item = begin; item < end; item++) {
        c.advance(item);
        const WtaTask &tk = c.cur;
        const int lpp = tk.D / 16;
        const int b0 = (item - tk.first_item) * (64 / lpp) * kWtaBatch;
        const bool eight = tk.ndirs == 8;
        const int lane = opaque_lane();
        switch (lpp) {
        case 2: eight ? wta_batch<2, 8>(tk.L, tk.dir_stride, tk.npix, tk.disp, b0, lane)
                      : wta_batch<2, 4>(tk.L, tk.dir_stride, tk.npix, tk.disp, b0, lane); break;
        case 4: eight ? wta_batch<4, 8>(tk.L, tk.dir_stride, tk.npix, tk.disp, b0, lane)
                      : wta_batch<4, 4>(tk.L, tk.dir_stride, tk.npix, tk.disp, b0, lane); break;
        default: eight ? wta_batch<8, 8>(tk.L, tk.dir_stride, tk.npix, tk.disp, b0, lane)
                       : wta_batch<8, 4>(tk.L, tk.dir_stride, tk.npix, tk.disp, b0, lane); break;
        }
    }
}

// ---- path aggregation and sum + argmin in one persistent launch ----
//
// A sum + argmin task (one plain pair, or one multiscale stage) depends on the
// path items of its directions only.  In the fused launch a wave that
// completes a path item counts its task's remaining items down (a release at
// device scope: its volume stores are visible before the count); the wave
// that takes the count to zero publishes the task in the next ready slot.
// Waves then take argmin units from ready slots, so the memory-bound argmin
// of early tasks overlaps the arithmetic-bound aggregation of later ones
// instead of running after all of it.  Waves whose block index is a multiple
// of 4 (a quarter of them) look for argmin units first, the others for path
// items first; both take the other kind when theirs has run out, and a wave
// with neither sleeps briefly while path items are still in flight.  Every
// claimed path item is held by a running wave, so every task is published
// and the launch drains.
//
// Synchronisation block (ints, one 64-byte line per counter):
//   [0] path item head, [1] publish counter,
//   [2 + t] remaining path items of task t,
//   [2 + nwta + s] 1 + the task published in slot s (0 = not yet),
//   [2 + 2 nwta + s] argmin units claimed from slot s.
constexpr int kWtaUnit = 16; // wta_batch calls per claimed argmin unit
// pixel batches per wta_batch call in the fused launch: eight-direction tasks
// load half as many at a time, so every task keeps 256 bytes per lane in
// flight and the kernel stays within 168 registers (three waves per SIMD)
template <int NDIRS>
constexpr int kFusedWtaBatch = NDIRS == 8 ? kWtaBatch / 2 : kWtaBatch;

__device__ __forceinline__ int wave_fetch_add(int *p, int v)
{
    int r = 0;
    if (threadIdx.x == 0)
        r = atomicAdd(p, v);
    return __builtin_amdgcn_readfirstlane(__shfl(r, 0));
}

template <int LPP, int NDIRS>
__device__ __forceinline__ void wta_unit(const WtaTask &tk, int u)
{
    constexpr int U = kFusedWtaBatch<NDIRS>, per = (64 / LPP) * U; // pixels per wta_batch call
    const int nb = (tk.npix + per - 1) / per;
    const int j1 = min(nb, (u + 1) * kWtaUnit);
    for (int j = u * kWtaUnit; j < j1; j++)
        wta_batch<LPP, NDIRS, U>(tk.L, tk.dir_stride, tk.npix, tk.disp, j * per, opaque_lane());
}

__device__ __forceinline__ void fused_wta_unit(const WtaTask &tk, int u)
{
    const bool eight = tk.ndirs == 8;
    switch (tk.D) {
    case 32: eight ? wta_unit<2, 8>(tk, u) : wta_unit<2, 4>(tk, u); break;
    case 64: eight ? wta_unit<4, 8>(tk, u) : wta_unit<4, 4>(tk, u); break;
    default: eight ? wta_unit<8, 8>(tk, u) : wta_unit<8, 4>(tk, u); break;
    }
}

__global__ __launch_bounds__(256) void k_fused_reset(int *__restrict__ sync, const int *__restrict__ left0, int nwta)
{
    const int n = 2 + 3 * nwta;
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        sync[i * kCounterStride] = (i >= 2 && i < 2 + nwta) ? left0[i - 2] : 0;
}

__global__ __launch_bounds__(64) void k_batch_fused(const BatchPathTask *__restrict__ tasks,
                                                    const uint32_t *__restrict__ items, int nitems,
                                                    const WtaTask *__restrict__ wtas, int nwta,
                                                    int *__restrict__ sync)
{
    int *head = sync, *pub = sync + kCounterStride;
    int *left = sync + 2 * kCounterStride;
    int *ready = left + nwta * kCounterStride, *taken = ready + nwta * kCounterStride;
    const bool wta_first = (blockIdx.x & 3) == 0;
    bool paths_open = true;
    int s = 0, acquired = -1; // this wave's slot cursor; the last slot it acquired

    auto try_wta = [&]() -> bool {
        while (s < nwta) {
            int t = 0;
            if (threadIdx.x == 0)
                t = __hip_atomic_load(ready + s * kCounterStride, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            t = __builtin_amdgcn_readfirstlane(__shfl(t, 0));
            if (t == 0)
                return false; // slot s not yet published
            if (acquired < s) {
                // pairs with the publisher's release: the task's volumes are
                // complete (the barrier extends lane 0's acquire to the wave)
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                __syncthreads();
                acquired = s;
            }
            const WtaTask tk = wtas[t - 1];
            const int u = wave_fetch_add(taken + s * kCounterStride, 1);
            if (u < tk.units) {
                fused_wta_unit(tk, u);
                return true;
            }
            s++;
        }
        return false;
    };
    auto try_path = [&]() -> bool {
        if (!paths_open)
            return false;
        const int it = wave_fetch_add(head, 1);
        if (it >= nitems) {
            paths_open = false;
            return false;
        }
        const uint32_t code = items[it];
        batch_path_item(tasks, code);
        const int dep = tasks[code >> 16].dep;
        __syncthreads(); // every lane's volume stores happen before lane 0's release
        int old = 0;
        if (threadIdx.x == 0)
            old = __hip_atomic_fetch_add(left + dep * kCounterStride, -1, __ATOMIC_RELEASE,
                                         __HIP_MEMORY_SCOPE_AGENT);
        old = __builtin_amdgcn_readfirstlane(__shfl(old, 0));
        if (old == 1) {
            // the last item of task dep: every other item's release is ordered before this
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            const int slot = wave_fetch_add(pub, 1);
            if (threadIdx.x == 0)
                __hip_atomic_store(ready + slot * kCounterStride, dep + 1, __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
        return true;
    };
    for (;;) {
        // one call site per kind keeps both inlined
        bool got = false;
        bool wta_now = wta_first;
#pragma nounroll
        for (int k = 0; k < 2 && !got; k++, wta_now = !wta_now)
            got = wta_now ? try_wta() : try_path();
        if (got)
            continue;
        if (s >= nwta)
            break; // every slot published and every unit claimed
        __builtin_amdgcn_s_sleep(8);
    }
}

template <class T>
struct DevArray {
    std::vector<T> host;
    T *dev = nullptr;
};

struct Schedule {
    int path_waves_per_simd = 2;
    int order = 0;        // 0: longest lines first; 1: pair by pair, directions in order
    int census_shape = 0; // census tile kCenShapeW[i] x kCenShapeH[i]
    int xcd_slices = 0;   // 0: one global path queue; 1: per-XCD slices
    int fused = 0;        // 1: path and argmin in one launch (k_batch_fused; order and slices unused)
};

} // namespace

// One lane: a subset of the pairs and their phase work lists, run as one
// chain of phase kernels on its own stream.
struct BatchLane {
    DevArray<PreTask> pre[kCenShapes][3]; // per census tile shape
    DevArray<CostTask> cost[2];
    DevArray<BatchPathTask> path;
    std::vector<uint32_t> items_lpt, items_pair;
    // fused launch: tasks by decreasing work, each task's items longest first
    std::vector<uint32_t> items_task;
    uint32_t *d_items = nullptr; // the three orders, back to back
    DevArray<WtaTask> wta;
    int pre_n[kCenShapes][3] = {}, cost_n[2] = {0, 0}, wta_n = 0; // items per launch
    int *d_counter = nullptr;
    DevArray<int> left0; // path items per argmin task
    int *d_sync = nullptr; // fused launch synchronisation block
    bool empty() const { return path.host.empty(); }
};

// A partition of the pairs into lanes.  Lanes share no data, so their chains
// run concurrently and one lane's memory-bound phases (census, cost volumes,
// argmin) overlap another's arithmetic-bound path aggregation.
//
// A pipelined set orders its lanes as groups instead: one stream runs every
// group's path launch in turn, a second every group's census and cost
// volumes ahead of it, and a third every group's argmin behind it, so group
// k's argmin and group k+1's cost volumes overlap path aggregation, and
// each hand-over is a launch boundary.
struct LaneSet {
    std::string name;
    std::vector<BatchLane> lanes;
    bool pipelined = false;
};

constexpr int kMaxLanes = 4;  // concurrent lane chains (one stream each)
constexpr int kMaxGroups = 8; // lanes of a pipelined set

struct sgmh_batch {
    sgmh_ctx *ctx = nullptr;
    std::vector<sgmh_pair> pairs;
    void *arena = nullptr;
    size_t arena_bytes = 0;
    std::vector<LaneSet> sets; // candidate partitions; sets[set] runs
    int set = 0;
    int cus = 256;
    Schedule sched;
    std::string tuning; // JSON fragment
    bool graph_ok = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t cap = nullptr;
    hipStream_t side[kMaxLanes - 1] = {};
    hipEvent_t fork = nullptr, join[kMaxLanes - 1] = {};
    hipEvent_t ev_pre[kMaxGroups] = {}, ev_path[kMaxGroups] = {}; // pipelined hand-overs
};

namespace {

template <class T>
int upload(DevArray<T> &a)
{
    if (a.host.empty())
        return 0;
    CK(hipMalloc(&a.dev, sizeof(T) * a.host.size()));
    CK(hipMemcpy(a.dev, a.host.data(), sizeof(T) * a.host.size(), hipMemcpyHostToDevice));
    return 0;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

int plain_variant(int D) { return D == 32 ? kVarPlain32 : D == 64 ? kVarPlain64 : kVarPlain128; }
int ms_variant(int D, bool dec)
{
    const int base = D == 32 ? 0 : D == 64 ? 1 : 2;
    return (dec ? kVarMsDec32 : kVarMs32) + base;
}

// total pixel-steps of a line group, the LPT key
long group_work(const BatchPathTask &t, int grp)
{
    const int W = t.W, H = t.H, dx = t.dir.dx, dy = t.dir.dy;
    if (dy == 0)
        return (long)W * t.D;
    if (dx == 0)
        return (long)H * t.D;
    long best = 0;
    for (int q = 0; q < kLinesPerWave; q++) {
        const int m = grp * kLinesPerWave + q;
        if (m >= t.dir.nlines)
            break;
        const int i = t.dir.dec ? dec_line(m, dx, dy, W) : m;
        best = std::max<long>(best, line_geom(i, dx, dy, W, H).steps);
    }
    return best * t.D;
}

// census / half / Sobel and the multiscale cost volumes of one lane
int enqueue_precost(sgmh_batch *b, BatchLane &ln, hipStream_t s, const Schedule &sc)
{
    const int pre_blocks = b->cus * 8;
    const int sh = sc.census_shape;
    for (int k = 0; k < 3; k++)
        if (ln.pre_n[sh][k]) {
            const dim3 g(std::min(ln.pre_n[sh][k], pre_blocks));
            if (sh == 0)
                hipLaunchKernelGGL((k_batch_pre<64, 16>), g, dim3(256), 0, s, ln.pre[sh][k].dev,
                                   (int)ln.pre[sh][k].host.size(), ln.pre_n[sh][k]);
            else
                hipLaunchKernelGGL((k_batch_pre<32, 32>), g, dim3(256), 0, s, ln.pre[sh][k].dev,
                                   (int)ln.pre[sh][k].host.size(), ln.pre_n[sh][k]);
        }
    for (int k = 0; k < 2; k++)
        if (ln.cost_n[k])
            hipLaunchKernelGGL(k_batch_cost, dim3(std::min(ln.cost_n[k], pre_blocks)), dim3(256), 0, s,
                               ln.cost[k].dev, (int)ln.cost[k].host.size(), ln.cost_n[k]);
    CK(hipGetLastError());
    return 0;
}

int enqueue_path(sgmh_batch *b, BatchLane &ln, hipStream_t s, const Schedule &sc)
{
    if (ln.path.host.empty())
        return 0;
    const int n = (int)ln.items_lpt.size();
    const uint32_t *items = ln.d_items + (sc.order == 0 ? 0 : n);
    const int waves = std::min(n, b->cus * 4 * sc.path_waves_per_simd);
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, ln.d_counter);
    if (sc.xcd_slices)
        hipLaunchKernelGGL(k_batch_path_xcd, dim3(waves), dim3(64), 0, s, ln.path.dev, items, n, ln.d_counter);
    else
        hipLaunchKernelGGL(k_batch_path, dim3(waves), dim3(64), 0, s, ln.path.dev, items, n, ln.d_counter);
    CK(hipGetLastError());
    return 0;
}

int enqueue_wta(sgmh_batch *b, BatchLane &ln, hipStream_t s)
{
    if (ln.wta_n)
        hipLaunchKernelGGL(k_batch_wta, dim3(std::min((ln.wta_n + 3) / 4, b->cus * 8)), dim3(256), 0, s,
                           ln.wta.dev, (int)ln.wta.host.size(), ln.wta_n);
    CK(hipGetLastError());
    return 0;
}

int enqueue_lane(sgmh_batch *b, BatchLane &ln, hipStream_t s, const Schedule &sc)
{
    if (enqueue_precost(b, ln, s, sc))
        return -1;
    if (sc.fused && !ln.path.host.empty()) {
        const int n = (int)ln.items_lpt.size(), nwta = (int)ln.wta.host.size();
        const int waves = std::min(n, b->cus * 4 * sc.path_waves_per_simd);
        hipLaunchKernelGGL(k_fused_reset, dim3(1), dim3(256), 0, s, ln.d_sync, ln.left0.dev, nwta);
        hipLaunchKernelGGL(k_batch_fused, dim3(waves), dim3(64), 0, s, ln.path.dev, ln.d_items + 2 * n, n,
                           ln.wta.dev, nwta, ln.d_sync);
        CK(hipGetLastError());
        return 0;
    }
    if (enqueue_path(b, ln, s, sc))
        return -1;
    return enqueue_wta(b, ln, s);
}

// A pipelined set: path launches in group order on `s`; census and cost
// volumes on side[0] ahead of them, argmin on side[1] behind them, ordered by
// events; both side streams join `s` at the end.
int enqueue_pipelined(sgmh_batch *b, hipStream_t s, std::vector<BatchLane> &groups, const Schedule &sc)
{
    hipStream_t pre = b->side[0], arg = b->side[1];
    const int G = (int)groups.size();
    CK(hipEventRecord(b->fork, s));
    CK(hipStreamWaitEvent(pre, b->fork, 0));
    CK(hipStreamWaitEvent(arg, b->fork, 0));
    for (int k = 0; k < G; k++) {
        if (enqueue_precost(b, groups[k], pre, sc))
            return -1;
        CK(hipEventRecord(b->ev_pre[k], pre));
    }
    for (int k = 0; k < G; k++) {
        CK(hipStreamWaitEvent(s, b->ev_pre[k], 0));
        if (enqueue_path(b, groups[k], s, sc))
            return -1;
        CK(hipEventRecord(b->ev_path[k], s));
        CK(hipStreamWaitEvent(arg, b->ev_path[k], 0));
        if (enqueue_wta(b, groups[k], arg))
            return -1;
    }
    CK(hipEventRecord(b->join[0], pre));
    CK(hipEventRecord(b->join[1], arg));
    CK(hipStreamWaitEvent(s, b->join[0], 0));
    CK(hipStreamWaitEvent(s, b->join[1], 0));
    return 0;
}

// Every lane of the chosen set: lane 0 on `s`, the others on side streams
// forked from it and joined back (or the pipelined order above).
int enqueue_phases(sgmh_batch *b, hipStream_t s, int set, const Schedule &sc)
{
    std::vector<BatchLane> &lanes = b->sets[set].lanes;
    if (b->sets[set].pipelined)
        return enqueue_pipelined(b, s, lanes, sc);
    const int K = (int)lanes.size();
    if (K == 1)
        return enqueue_lane(b, lanes[0], s, sc);
    CK(hipEventRecord(b->fork, s));
    for (int k = 1; k < K; k++) {
        CK(hipStreamWaitEvent(b->side[k - 1], b->fork, 0));
        if (enqueue_lane(b, lanes[k], b->side[k - 1], sc))
            return -1;
        CK(hipEventRecord(b->join[k - 1], b->side[k - 1]));
    }
    if (enqueue_lane(b, lanes[0], s, sc))
        return -1;
    for (int k = 1; k < K; k++)
        CK(hipStreamWaitEvent(s, b->join[k - 1], 0));
    return 0;
}

int capture(sgmh_batch *b)
{
    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
    if (b->graph) {
        (void)hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
    CK(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_phases(b, b->cap, b->set, b->sched);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->cap, &g);
    if (rc)
        return -1;
    CK(e);
    b->graph = g;
    CK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->graph_ok = true;
    return 0;
}

// Time candidate schedules (median of 3 runs after one warm-up) and keep the
// fastest, in two stages: every lane partition x path item order x path waves
// per SIMD with the default census tile and one global path queue; then, on
// the best of those, the other census tile shape and the per-XCD scanline
// grouping, alone and together.
const char *order_name(const Schedule &sc)
{
    return sc.fused ? "task-by-task" : sc.order == 0 ? "longest-first" : "pair-major";
}

int autotune(sgmh_batch *b)
{
    struct Cand {
        int set;
        Schedule sc;
    };
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    hipStream_t s = b->cap;
    double best = 1e30;
    std::string js = "[";
    int ntimed = 0;
    auto time_cands = [&](const std::vector<Cand> &cands) -> int {
        for (const Cand &c : cands) {
            float t[3];
            if (enqueue_phases(b, s, c.set, c.sc))
                return -1;
            for (int r = 0; r < 3; r++) {
                CK(hipEventRecord(e0, s));
                if (enqueue_phases(b, s, c.set, c.sc))
                    return -1;
                CK(hipEventRecord(e1, s));
                CK(hipEventSynchronize(e1));
                CK(hipEventElapsedTime(&t[r], e0, e1));
            }
            std::sort(t, t + 3);
            char buf[320];
            snprintf(buf, sizeof buf,
                     "%s{\"lanes\":\"%s\",\"order\":\"%s\",\"path_waves_per_simd\":%d,\"census_tile\":\"%dx%d\","
                     "\"scanline_grouping\":\"%s\",\"fused\":%s,\"ms\":%.4f}",
                     ntimed++ ? "," : "", b->sets[c.set].name.c_str(), order_name(c.sc),
                     c.sc.path_waves_per_simd, kCenShapeW[c.sc.census_shape], kCenShapeH[c.sc.census_shape],
                     c.sc.xcd_slices ? "per-XCD" : "global", c.sc.fused ? "true" : "false",
                     t[1]);
            js += buf;
            if (t[1] < best) {
                best = t[1];
                b->set = c.set;
                b->sched = c.sc;
            }
        }
        return 0;
    };
    // The fused path + argmin launch is not a candidate: it measured slower
    // than separate launches on every partition (its per-item device-scope
    // release is the cost), so SGMH_BATCH_FUSED selects it explicitly.
    std::vector<Cand> stage1;
    for (int set = 0; set < (int)b->sets.size(); set++)
        for (int order = 0; order < 2; order++)
            for (int pw : {1, 2, 4}) {
                Cand c;
                c.set = set;
                c.sc.order = order;
                c.sc.path_waves_per_simd = pw;
                stage1.push_back(c);
            }
    if (time_cands(stage1))
        return -1;
    std::vector<Cand> stage2;
    for (int sh = 0; sh < kCenShapes; sh++)
        for (int xs = 0; xs < 2; xs++) {
            if (sh == b->sched.census_shape && xs == b->sched.xcd_slices)
                continue;
            Cand c{b->set, b->sched};
            c.sc.census_shape = sh;
            c.sc.xcd_slices = xs;
            stage2.push_back(c);
        }
    if (time_cands(stage2))
        return -1;
    js += "]";
    b->tuning = js;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 0;
}

} // namespace

extern "C" {

int sgmh_batch_plan(sgmh_ctx *ctx, const sgmh_pair *pairs, int n, int flags, sgmh_batch **out)
{
    REQUIRE(ctx && pairs && out && n > 0, "sgmh_batch_plan: bad arguments");
    REQUIRE(!((flags & SGMH_BATCH_FUSED) && (flags & SGMH_BATCH_PIPELINED)),
            "sgmh_batch_plan: SGMH_BATCH_FUSED and SGMH_BATCH_PIPELINED exclude each other");
    *out = nullptr;
    for (int i = 0; i < n; i++) {
        const sgmh_pair &p = pairs[i];
        REQUIRE(p.d_left && p.d_right && p.d_disp, "sgmh_batch_plan: pair %d has a NULL buffer", i);
        if (check_dims(p.W, p.H, p.D))
            return -1;
        REQUIRE(p.D == 32 || p.D == 64 || p.D == 128,
                "sgmh_batch_plan: pair %d D=%d (batched mode runs D = 32, 64, 128)", i, p.D);
        REQUIRE(p.mode == SGMH_PLAIN || p.mode == SGMH_MULTISCALE, "sgmh_batch_plan: pair %d mode", i);
        if (p.mode == SGMH_PLAIN)
            REQUIRE(p.paths == 4 || p.paths == 8, "sgmh_batch_plan: pair %d paths=%d", i, p.paths);
        else
            REQUIRE(p.W % 2 == 0 && p.H % 2 == 0 && p.W >= 32 && p.H >= 32,
                    "sgmh_batch_plan: pair %d multiscale needs even W, H >= 32", i);
        const size_t N = (size_t)p.W * p.H;
        REQUIRE(N * p.D < (1ull << 31) && (N + kLead) * 8 < (1ull << 31),
                "sgmh_batch_plan: pair %d exceeds the 2 GiB per-volume limit", i);
    }
    CK(hipSetDevice(ctx->device));
    sgmh_batch *b = new sgmh_batch();
    b->ctx = ctx;
    b->pairs.assign(pairs, pairs + n);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess && prop.multiProcessorCount > 0)
        b->cus = prop.multiProcessorCount;

    // ---- workspace layout ----
    struct Ws {
        size_t cl, cr, vol;                                  // plain
        size_t hl, hr, sl, sr, shl, shr, cl1, cr1, cl2, cr2; // multiscale
        size_t c1, c2, v1, v2, d1;
    };
    std::vector<Ws> ws(n);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += align256(bytes);
        return o;
    };
    for (int i = 0; i < n; i++) {
        const sgmh_pair &p = pairs[i];
        const size_t N = (size_t)p.W * p.H, Nh = N / 4, D = (size_t)p.D;
        Ws &w = ws[i];
        if (p.mode == SGMH_PLAIN) {
            w.cl = take((N + kLead) * 8);
            w.cr = take((N + kLead) * 8);
            w.vol = take(N * D * p.paths);
        } else {
            w.hl = take(Nh);
            w.hr = take(Nh);
            w.sl = take(N);
            w.sr = take(N);
            w.shl = take(Nh);
            w.shr = take(Nh);
            w.cl1 = take(Nh * 8);
            w.cr1 = take(Nh * 8);
            w.cl2 = take(N * 8);
            w.cr2 = take(N * 8);
            w.c1 = take(Nh * D);
            w.c2 = take(N * D);
            w.v1 = take(Nh * D * 8);
            w.v2 = take(Nh * D * 8);
            w.d1 = take(Nh);
        }
    }
    b->arena_bytes = off;
    if (hipMalloc(&b->arena, off) != hipSuccess) {
        set_err("sgmh_batch_plan: cannot allocate %zu bytes of workspace", off);
        delete b;
        return -1;
    }
    // zeroed once: the census lead slack of each plain pair must hold valid codes
    if (hipMemset(b->arena, 0, off) != hipSuccess) {
        set_err("sgmh_batch_plan: cannot clear the workspace");
        (void)hipFree(b->arena);
        delete b;
        return -1;
    }
    uint8_t *A = (uint8_t *)b->arena;

    // ---- phase work lists, one set per candidate partition of the pairs ----
    auto build = [&](const std::vector<int> &lane_of, int nl, const char *name) -> int {
        LaneSet ls;
        ls.name = name;
        ls.lanes.resize(nl);
        int pre_items[kMaxGroups][kCenShapes][3] = {}, cost_items[kMaxGroups][2] = {}, wta_items[kMaxGroups] = {};
        int li = 0;
        auto add_pre = [&](int level, int kind, const uint8_t *src, void *dst, int W, int H, int ostride) {
            for (int sh = 0; sh < kCenShapes; sh++) {
                PreTask t;
                t.src = src;
                t.dst = dst;
                t.W = W;
                t.H = H;
                t.kind = kind;
                t.ostride = ostride;
                int ty;
                if (kind == PRE_CENSUS) {
                    t.tiles_x = (W + kCenShapeW[sh] - 1) / kCenShapeW[sh];
                    ty = (H + kCenShapeH[sh] - 1) / kCenShapeH[sh];
                } else {
                    const int ow = kind == PRE_HALF ? W / 2 : W, oh = kind == PRE_HALF ? H / 2 : H;
                    t.tiles_x = (ow + 63) / 64;
                    ty = (oh + 3) / 4;
                }
                t.first_item = pre_items[li][sh][level];
                pre_items[li][sh][level] += t.tiles_x * ty;
                ls.lanes[li].pre[sh][level].host.push_back(t);
            }
        };
        auto add_cost = [&](int level, const uint64_t *cl, const uint64_t *cr, const uint8_t *merge, uint8_t *C,
                            int W, int H, int D) {
            CostTask t{cl, cr, merge, C, W, H, D, kMsInvalid, cost_items[li][level]};
            cost_items[li][level] += (int)(((size_t)W * H + 15) / 16);
            ls.lanes[li].cost[level].host.push_back(t);
        };
        auto add_wta = [&](const uint8_t *L, size_t stride, int ndirs, int npix, int D, uint8_t *disp) {
            const int per = (64 / (D / 16)) * kWtaBatch, nb = (npix + per - 1) / per;
            const int fper = (64 / (D / 16)) * (ndirs == 8 ? kFusedWtaBatch<8> : kFusedWtaBatch<4>);
            const int fnb = (npix + fper - 1) / fper;
            WtaTask t{L, stride, ndirs, npix, D, disp, wta_items[li], (fnb + kWtaUnit - 1) / kWtaUnit};
            wta_items[li] += nb;
            ls.lanes[li].wta.host.push_back(t);
        };
        // index the next add_wta call of the current lane gets (+ k)
        auto next_wta = [&](int k) { return (int)ls.lanes[li].wta.host.size() + k; };
        auto add_path = [&](const uint64_t *cl, const uint64_t *cr, const uint8_t *cost, int W, int H, int D,
                            uint32_t P2, uint32_t inval, uint32_t out_bytes, int dx, int dy, int P1, uint8_t *L,
                            int variant, int dep) {
            BatchPathTask t;
            t.cl = cl;
            t.cr = cr;
            t.cost = cost;
            t.W = W;
            t.H = H;
            t.D = D;
            t.P2 = P2;
            t.invalid = inval;
            t.out_bytes = out_bytes;
            t.dir.dx = dx;
            t.dir.dy = dy;
            t.dir.P1 = P1;
            // the decimated stage aggregates only the lines through even (x, y)
            t.dir.dec = variant >= kVarMsDec32 ? 1 : 0;
            t.dir.nlines = t.dir.dec ? dec_lines(dx, dy, W, H) : (dy == 0 ? H : (dx == 0 ? W : W + H - 1));
            t.dir.first_block = 0;
            t.dir.L = L;
            t.variant = variant;
            t.dep = dep;
            ls.lanes[li].path.host.push_back(t);
        };
        for (int i = 0; i < n; i++) {
            const sgmh_pair &p = pairs[i];
            const Ws &w = ws[i];
            const int W = p.W, H = p.H, D = p.D, hw = W / 2, hh = H / 2;
            const size_t N = (size_t)W * H, Nh = (size_t)hw * hh;
            li = lane_of[i];
            if (p.mode == SGMH_PLAIN) {
                uint64_t *cl = (uint64_t *)(A + w.cl), *cr = (uint64_t *)(A + w.cr);
                add_pre(0, PRE_CENSUS, p.d_left, cl + kLead, W, H, W);
                add_pre(0, PRE_CENSUS, p.d_right, cr + kLead, W, H, W);
                for (int k = 0; k < p.paths; k++)
                    add_path(cl, cr, nullptr, W, H, D, kPlainP2, kPlainInvalid, (uint32_t)(N * D),
                             kDirsPlain[k][0], kDirsPlain[k][1], (int)kPlainP1, A + w.vol + (size_t)k * N * D,
                             plain_variant(D), next_wta(0));
                add_wta(A + w.vol, N * D, p.paths, (int)N, D, p.d_disp);
            } else {
                uint8_t *hl = A + w.hl, *hr = A + w.hr, *sl = A + w.sl, *sr = A + w.sr;
                uint8_t *shl = A + w.shl, *shr = A + w.shr;
                uint64_t *cl1 = (uint64_t *)(A + w.cl1), *cr1 = (uint64_t *)(A + w.cr1);
                uint64_t *cl2 = (uint64_t *)(A + w.cl2), *cr2 = (uint64_t *)(A + w.cr2);
                uint8_t *C1 = A + w.c1, *C2 = A + w.c2, *V1 = A + w.v1, *V2 = A + w.v2;
                add_pre(0, PRE_HALF, p.d_left, hl, W, H, 0);
                add_pre(0, PRE_HALF, p.d_right, hr, W, H, 0);
                add_pre(0, PRE_SOBEL, p.d_left, sl, W, H, 0);
                add_pre(0, PRE_SOBEL, p.d_right, sr, W, H, 0);
                add_pre(1, PRE_SOBEL, hl, shl, hw, hh, 0);
                add_pre(1, PRE_SOBEL, hr, shr, hw, hh, 0);
                add_pre(1, PRE_CENSUS, sl, cl2, W, H, W);
                add_pre(1, PRE_CENSUS, sr, cr2, W, H, W);
                add_pre(2, PRE_CENSUS, shl, cl1, hw, hh, hw);
                add_pre(2, PRE_CENSUS, shr, cr1, hw, hh, hw);
                add_cost(0, cl1, cr1, nullptr, C1, hw, hh, D);
                add_cost(1, cl2, cr2, C1, C2, W, H, D);
                for (int k = 0; k < 8; k++) {
                    add_path(nullptr, nullptr, C1, hw, hh, D, kMsP2, kMsInvalid, (uint32_t)(Nh * D),
                             kDirsMs[k][0], kDirsMs[k][1], kDirsMs[k][2], V1 + (size_t)k * Nh * D,
                             ms_variant(D, false), next_wta(0));
                    add_path(nullptr, nullptr, C2, W, H, D, kMsP2, kMsInvalid, (uint32_t)(Nh * D),
                             kDirsMs[k][0], kDirsMs[k][1], kDirsMs[k][2], V2 + (size_t)k * Nh * D,
                             ms_variant(D, true), next_wta(1));
                }
                add_wta(V1, Nh * D, 8, (int)Nh, D, A + w.d1);
                add_wta(V2, Nh * D, 8, (int)Nh, D, p.d_disp);
            }
        }
        // drop empty lanes
        std::vector<BatchLane> kept;
        int kept_idx[kMaxGroups];
        for (int L = 0; L < nl; L++)
            if (!ls.lanes[L].path.host.empty()) {
                kept_idx[kept.size()] = L;
                kept.push_back(ls.lanes[L]);
            }
        for (size_t q = 0; q < kept.size(); q++) {
            BatchLane &ln = kept[q];
            const int L = kept_idx[q];
            for (int sh = 0; sh < kCenShapes; sh++)
                for (int k = 0; k < 3; k++)
                    ln.pre_n[sh][k] = pre_items[L][sh][k];
            for (int k = 0; k < 2; k++)
                ln.cost_n[k] = cost_items[L][k];
            ln.wta_n = wta_items[L];
            // path items in three orders: longest lines first; pair by pair;
            // and (fused launch) argmin task by task, the task with the most
            // work first and each task's longest lines first, so tasks become
            // ready one after another and the last ones are short
            std::vector<std::pair<long, uint32_t>> all;
            const int nwta = (int)ln.wta.host.size();
            std::vector<long> task_work(nwta, 0);
            ln.left0.host.assign(nwta, 0);
            for (size_t t = 0; t < ln.path.host.size(); t++) {
                const BatchPathTask &tk = ln.path.host[t];
                const int groups = (tk.dir.nlines + kLinesPerWave - 1) / kLinesPerWave;
                for (int g = 0; g < groups; g++) {
                    const uint32_t code = ((uint32_t)t << 16) | (uint32_t)g;
                    ln.items_pair.push_back(code);
                    all.push_back({group_work(tk, g), code});
                    task_work[tk.dep] += all.back().first;
                }
                ln.left0.host[tk.dep] += groups;
            }
            REQUIRE(ln.path.host.size() < 65536, "sgmh_batch_plan: too many directions");
            std::stable_sort(all.begin(), all.end(),
                             [](const auto &x, const auto &y) { return x.first > y.first; });
            for (auto &x : all)
                ln.items_lpt.push_back(x.second);
            std::vector<std::pair<long, uint32_t>> by_task = all; // already longest first
            std::stable_sort(by_task.begin(), by_task.end(), [&](const auto &x, const auto &y) {
                const int dx = ln.path.host[x.second >> 16].dep, dy = ln.path.host[y.second >> 16].dep;
                return task_work[dx] != task_work[dy] ? task_work[dx] > task_work[dy] : dx < dy;
            });
            for (auto &x : by_task)
                ln.items_task.push_back(x.second);
        }
        ls.lanes = kept;
        b->sets.push_back(ls);
        return 0;
    };
    // Candidate partitions: by configuration (plain / multiscale), and
    // 2, 3, 4 lanes balanced by estimated work (longest pair first to the
    // least-loaded lane).
    std::vector<double> work(n);
    for (int i = 0; i < n; i++) {
        const sgmh_pair &p = pairs[i];
        const double nd = (double)p.W * p.H * p.D;
        work[i] = p.mode == SGMH_PLAIN ? nd * p.paths * 1.5 : nd * 8 * 1.25 * 2.0;
    }
    std::vector<int> by_work(n);
    for (int i = 0; i < n; i++)
        by_work[i] = i;
    std::stable_sort(by_work.begin(), by_work.end(), [&](int x, int y) { return work[x] > work[y]; });
    int rc = 0;
    {
        std::vector<int> lane_of(n);
        for (int i = 0; i < n; i++)
            lane_of[i] = pairs[i].mode == SGMH_PLAIN ? 0 : 1;
        rc = build(lane_of, 2, "by-configuration");
    }
    const bool tune = (flags & SGMH_BATCH_AUTOTUNE) != 0;
    for (int K = 2; K <= kMaxLanes && tune && !rc && K <= n; K++) {
        std::vector<int> lane_of(n);
        std::vector<double> load(K, 0.0);
        for (int i : by_work) {
            const int L = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            lane_of[i] = L;
            load[L] += work[i];
        }
        char name[32];
        snprintf(name, sizeof name, "balanced-%d", K);
        rc = build(lane_of, K, name);
    }
    // Pipelined sets: the pairs in decreasing path work (line steps x D:
    // paths * N * D plain; 6 * N * D multiscale, a quarter-size stage and a
    // stage that aggregates half the full-size lines, 8 directions each), cut
    // into G contiguous groups of about equal work (a pair joins the group its
    // work's midpoint falls in).  Multiscale pairs in the first group trade
    // places with the next plain pairs, so the first path launch waits only
    // for census.
    const bool pipe_flag = (flags & SGMH_BATCH_PIPELINED) != 0;
    int first_pipelined = -1;
    if ((tune || pipe_flag) && !rc) {
        std::vector<double> pwork(n);
        double total = 0.0;
        for (int i = 0; i < n; i++) {
            const sgmh_pair &p = pairs[i];
            const double nd = (double)p.W * p.H * p.D;
            pwork[i] = p.mode == SGMH_PLAIN ? nd * p.paths : nd * 6.0;
            total += pwork[i];
        }
        std::vector<int> order(n);
        for (int i = 0; i < n; i++)
            order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return pwork[x] > pwork[y]; });
        for (int G : {3, 4, 6}) {
            if (G > n || rc)
                break;
            std::vector<int> lane_of(n);
            double cum = 0.0;
            for (int i : order) {
                lane_of[i] = std::min(G - 1, (int)(G * (cum + 0.5 * pwork[i]) / total));
                cum += pwork[i];
            }
            size_t next_plain = 0;
            for (int i : order) {
                if (lane_of[i] != 0 || pairs[i].mode == SGMH_PLAIN)
                    continue;
                while (next_plain < order.size() &&
                       (lane_of[order[next_plain]] == 0 || pairs[order[next_plain]].mode != SGMH_PLAIN))
                    next_plain++;
                if (next_plain == order.size())
                    break;
                std::swap(lane_of[i], lane_of[order[next_plain]]);
            }
            char name[32];
            snprintf(name, sizeof name, "pipelined-%d", G);
            rc = build(lane_of, G, name);
            if (!rc) {
                b->sets.back().pipelined = true;
                if (first_pipelined < 0)
                    first_pipelined = (int)b->sets.size() - 1;
            }
        }
    }
    for (LaneSet &ls : b->sets)
        for (BatchLane &ln : ls.lanes) {
            for (int sh = 0; sh < kCenShapes; sh++)
                for (int k = 0; k < 3 && !rc; k++)
                    rc = upload(ln.pre[sh][k]);
            for (int k = 0; k < 2 && !rc; k++)
                rc = upload(ln.cost[k]);
            if (!rc)
                rc = upload(ln.path);
            if (!rc)
                rc = upload(ln.wta);
            if (!rc)
                rc = upload(ln.left0);
            if (!rc && !ln.items_lpt.empty()) {
                const size_t m = ln.items_lpt.size();
                if (hipMalloc(&ln.d_items, 3 * m * sizeof(uint32_t)) != hipSuccess ||
                    hipMemcpy(ln.d_items, ln.items_lpt.data(), m * 4, hipMemcpyHostToDevice) != hipSuccess ||
                    hipMemcpy(ln.d_items + m, ln.items_pair.data(), m * 4, hipMemcpyHostToDevice) != hipSuccess ||
                    hipMemcpy(ln.d_items + 2 * m, ln.items_task.data(), m * 4, hipMemcpyHostToDevice) !=
                        hipSuccess) {
                    set_err("sgmh_batch_plan: cannot upload the path work list");
                    rc = -1;
                }
            }
            if (!rc && (hipMalloc(&ln.d_counter, sizeof(int) * kCounterStride * kCounters) != hipSuccess ||
                        hipMalloc(&ln.d_sync, sizeof(int) * kCounterStride * (2 + 3 * ln.wta.host.size())) !=
                            hipSuccess)) {
                set_err("sgmh_batch_plan: cannot allocate the work counters");
                rc = -1;
            }
        }
    if (!rc && (hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess)) {
        set_err("sgmh_batch_plan: cannot create streams / events");
        rc = -1;
    }
    for (int k = 0; k < kMaxLanes - 1 && !rc; k++)
        if (hipStreamCreateWithFlags(&b->side[k], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&b->join[k], hipEventDisableTiming) != hipSuccess) {
            set_err("sgmh_batch_plan: cannot create streams / events");
            rc = -1;
        }
    for (int k = 0; k < kMaxGroups && !rc; k++)
        if (hipEventCreateWithFlags(&b->ev_pre[k], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&b->ev_path[k], hipEventDisableTiming) != hipSuccess) {
            set_err("sgmh_batch_plan: cannot create events");
            rc = -1;
        }
    if (!rc && (flags & SGMH_BATCH_AUTOTUNE)) {
        rc = autotune(b);
    } else {
        b->sched.census_shape = (flags & SGMH_BATCH_CENSUS_32X32) ? 1 : 0;
        b->sched.fused = (flags & SGMH_BATCH_FUSED) ? 1 : 0;
        b->sched.xcd_slices = (flags & SGMH_BATCH_XCD_SLICES) && !b->sched.fused ? 1 : 0;
        if (pipe_flag && first_pipelined >= 0)
            b->set = first_pipelined;
    }
    if (!rc && !(flags & SGMH_BATCH_NO_GRAPH))
        rc = capture(b);
    if (!rc && hipStreamSynchronize(b->cap) != hipSuccess) {
        set_err("sgmh_batch_plan: stream synchronisation failed");
        rc = -1;
    }
    if (rc) {
        sgmh_batch_destroy(b);
        return -1;
    }
    *out = b;
    return 0;
}

int sgmh_batch_run(sgmh_batch *b, void *stream)
{
    REQUIRE(b, "sgmh_batch_run: NULL batch");
    CK(hipSetDevice(b->ctx->device));
    hipStream_t s = (hipStream_t)stream;
    if (b->graph_ok) {
        CK(hipGraphLaunch(b->exec, s));
        return 0;
    }
    return enqueue_phases(b, s, b->set, b->sched);
}

int sgmh_batch_run_loop(sgmh_batch *b, void *stream)
{
    REQUIRE(b, "sgmh_batch_run_loop: NULL batch");
    for (const sgmh_pair &p : b->pairs) {
        const int rc = p.mode == SGMH_PLAIN
                           ? sgmh_plain_dev(b->ctx, p.d_left, p.d_right, p.W, p.H, p.D, p.paths, p.d_disp, stream)
                           : sgmh_ms_dev(b->ctx, p.d_left, p.d_right, p.W, p.H, p.D, p.d_disp, nullptr, stream);
        if (rc)
            return -1;
    }
    return 0;
}

int sgmh_batch_describe(const sgmh_batch *b, char *buf, size_t len)
{
    if (!b || !buf || len == 0)
        return -1;
    int plain = 0, ms = 0;
    for (const sgmh_pair &p : b->pairs)
        (p.mode == SGMH_PLAIN ? plain : ms)++;
    size_t items = 0, tasks = 0, wtas = 0;
    const LaneSet &ls = b->sets[b->set];
    for (const BatchLane &ln : ls.lanes) {
        items += ln.items_lpt.size();
        tasks += ln.path.host.size();
        wtas += ln.wta.host.size();
    }
    const int n = snprintf(
        buf, len,
        "{\"pairs\":%zu,\"plain\":%d,\"multiscale\":%d,\"workspace_bytes\":%zu,\"path_items\":%zu,"
        "\"path_tasks\":%zu,\"wta_tasks\":%zu,\"graph\":%s,\"schedule\":{\"lanes\":\"%s\",\"lane_count\":%zu,"
        "\"pipelined\":%s,\"order\":\"%s\",\"path_waves_per_simd\":%d,\"census_tile\":\"%dx%d\",\"scanline_grouping\":\"%s\","
        "\"fused\":%s},\"tuning\":%s}",
        b->pairs.size(), plain, ms, b->arena_bytes, items, tasks, wtas, b->graph_ok ? "true" : "false",
        ls.name.c_str(), ls.lanes.size(), ls.pipelined ? "true" : "false", order_name(b->sched),
        b->sched.path_waves_per_simd, kCenShapeW[b->sched.census_shape], kCenShapeH[b->sched.census_shape],
        b->sched.xcd_slices ? "per-XCD" : "global", b->sched.fused ? "true" : "false",
        b->tuning.empty() ? "null" : b->tuning.c_str());
    return n;
}

void sgmh_batch_destroy(sgmh_batch *b)
{
    if (!b)
        return;
    (void)hipSetDevice(b->ctx->device);
    (void)hipDeviceSynchronize();
    if (b->exec)
        (void)hipGraphExecDestroy(b->exec);
    if (b->graph)
        (void)hipGraphDestroy(b->graph);
    for (LaneSet &ls : b->sets)
        for (BatchLane &ln : ls.lanes) {
        for (auto &shape : ln.pre)
            for (auto &a : shape)
                if (a.dev)
                    (void)hipFree(a.dev);
        for (auto &a : ln.cost)
            if (a.dev)
                (void)hipFree(a.dev);
        if (ln.path.dev)
            (void)hipFree(ln.path.dev);
        if (ln.wta.dev)
            (void)hipFree(ln.wta.dev);
        if (ln.d_items)
            (void)hipFree(ln.d_items);
        if (ln.d_counter)
            (void)hipFree(ln.d_counter);
        if (ln.left0.dev)
            (void)hipFree(ln.left0.dev);
        if (ln.d_sync)
            (void)hipFree(ln.d_sync);
        }
    if (b->arena)
        (void)hipFree(b->arena);
    if (b->cap)
        (void)hipStreamDestroy(b->cap);
    for (int k = 0; k < kMaxGroups; k++) {
        if (b->ev_pre[k])
            (void)hipEventDestroy(b->ev_pre[k]);
        if (b->ev_path[k])
            (void)hipEventDestroy(b->ev_path[k]);
    }
    for (int k = 0; k < kMaxLanes - 1; k++) {
        if (b->side[k])
            (void)hipStreamDestroy(b->side[k]);
        if (b->join[k])
            (void)hipEventDestroy(b->join[k]);
    }
    if (b->fork)
        (void)hipEventDestroy(b->fork);
    delete b;
}

} // extern "C"
