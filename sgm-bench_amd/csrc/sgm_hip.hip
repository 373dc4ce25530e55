// sgm_hip.hip -- single-pair SGM on gfx950 and the C ABI of include/sgm_hip.h.
//
// Pipeline (plain path, common/sgm_ref.c:101-144):
//   k_census        9x7 census of both images from an LDS-staged tile
//   k_path<DPL,..>  every direction of the frame in ONE launch: one scanline
//                   per 16-lane DPP row, the Hamming cost computed on the fly
//                   from the census codes, the path cost L of that direction
//                   written as uint8 to its own volume
//   k_wta<LPP,..>   S = sum of the direction volumes (uint16 semantics) and
//                   argmin with the lowest d winning ties, fused
// Multiscale (multiscale/sgm_ms_ref.c:166-202) makes the census of the
// Sobel images at both scales in one launch (k_census_ms, half-scale and Sobel
// filters fused into the census tile), adds k_cost (materialised cost volume,
// merged across scales) and runs both stages' paths in one launch
// (k_path_ms) in the cost-volume mode with saturation and direction-weighted
// P1.  k_half and k_sobel remain as the filters' own entry points.
// The device code itself is in sgm_device.h; the persistent batch kernels
// that reuse it are in sgm_batch.hip.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <mutex>
#include <vector>

#include "sgm_device.h"
#include "sgm_internal.h"

using namespace sgmk;

static thread_local char g_err[512] = "";

void sgmh_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

// ---------------------------------------------------------------------------
// kernels: one grid point per tile / pixel group / line group
// ---------------------------------------------------------------------------
// The census of a stereo pair in one launch: blockIdx.z = 0 the left image, 1 the right.
__global__ __launch_bounds__(256) void k_census(const uint8_t *__restrict__ img0, const uint8_t *__restrict__ img1,
                                                uint64_t *__restrict__ out0, uint64_t *__restrict__ out1, int W,
                                                int H, int ostride)
{
    __shared__ __align__(16) uint32_t tile[kCenRows * kCenRowWords];
    const bool second = blockIdx.z != 0;
    census_tile(second ? img1 : img0, second ? out1 : out0, W, H, ostride, blockIdx.x * kCenTX,
                blockIdx.y * kCenTY, tile);
}

// The multiscale census of both images at both scales in one launch
// (sobel_census_tile): z = 0, 1 the half scale of left, right; z = 2, 3 the
// full scale.  The grid covers the full scale; the half scale's extra blocks
// return at once.
__global__ __launch_bounds__(256) void k_census_ms(const uint8_t *__restrict__ l, const uint8_t *__restrict__ r,
                                                   int W, int H, uint64_t *__restrict__ cl1,
                                                   uint64_t *__restrict__ cr1, uint64_t *__restrict__ cl2,
                                                   uint64_t *__restrict__ cr2)
{
    __shared__ __align__(16) uint32_t tile[kCenRows * kCenRowWords];
    __shared__ uint8_t win[kSobWinX * kSobWinY];
    const int z = (int)blockIdx.z;
    const bool half = z < 2, right = (z & 1) != 0;
    const int x0 = (int)blockIdx.x * kCenTX, y0 = (int)blockIdx.y * kCenTY;
    if (x0 >= (half ? W / 2 : W) || y0 >= (half ? H / 2 : H))
        return;
    const uint8_t *img = right ? r : l;
    if (half)
        sobel_census_tile<true>(img, W, H, right ? cr1 : cl1, x0, y0, tile, win);
    else
        sobel_census_tile<false>(img, W, H, right ? cr2 : cl2, x0, y0, tile, win);
}

__global__ __launch_bounds__(256) void k_half(const uint8_t *__restrict__ img, int W, int H,
                                              uint8_t *__restrict__ out)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x < W / 2 && y < H / 2)
        half_px(img, W, out, W / 2, x, y);
}

__global__ __launch_bounds__(256) void k_sobel(const uint8_t *__restrict__ img, int W, int H,
                                               uint8_t *__restrict__ out)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x < W && y < H)
        sobel_px(img, W, H, out, x, y);
}

constexpr int kCostBatch = 4; // pixel groups per wave of the cost kernel

// Materialised cost volume (multiscale), one pixel per 16-lane row, 64
// consecutive pixels of one row per workgroup; a wave takes 4 pixels 16 apart
// at a time, kCostBatch times, with every load issued before any result:
//   merge == nullptr: C = popcount(cl ^ cr[x-d]) or `invalid`      (sgm_ms_ref.c:66-72)
//   merge != nullptr: C = (C + merge[(y/2)*hw + x/2][d/2] + 1) >> 1 (sgm_ms_ref.c:184-188)
template <int DPL>
__global__ __launch_bounds__(256) void k_cost(const uint64_t *__restrict__ cl,
                                              const uint64_t *__restrict__ cr, int W, int H, int D,
                                              uint32_t invalid, const uint8_t *__restrict__ merge,
                                              uint8_t *__restrict__ C)
{
    const int lane = threadIdx.x & 63, g = lane >> 4, l = lane & 15;
    const int x0 = (int)blockIdx.x * (16 * kCostBatch) + (int)(threadIdx.x >> 6) * 4 + g, y = (int)blockIdx.y;
    const rsrc_t rr = make_rsrc(cr, (uint32_t)(W * H) * 8u);
    CostIn<DPL> in[kCostBatch];
#pragma unroll
    for (int u = 0; u < kCostBatch; u++) {
        const int x = x0 + 16 * u;
        if (x < W)
            in[u].load(cl, cr, rr, W, D, merge, x, y, l);
    }
#pragma unroll
    for (int u = 0; u < kCostBatch; u++) {
        const int x = x0 + 16 * u;
        if (x < W)
            in[u].finish(W, D, invalid, merge != nullptr, C, x, y, l);
    }
}

// Workgroups are dealt to the 8 XCDs round-robin by id, and each XCD has its
// own L2.  Every direction's block range starts at a multiple of 8 and is
// padded to one.  For columns, block i serves line group
// (i % 8) * (n8 / 8) + i / 8: each XCD owns one contiguous slice of the
// columns, so neighbouring columns -- whose 64-code disparity windows overlap
// -- share an L2.  Rows gain nothing from it, and diagonals must stay
// interleaved (their lengths vary along the range), so they keep block order.
constexpr int kXcds = 8;
constexpr int kRowBcastMinH = 512; // 2 x 512 / 4 row waves = 256, one per CU
constexpr long kWaveLinesMaxPixelsD128 = 1280L * 720; // D = 128 frames up to this size take one line per wave

// LPL: lanes per scanline, 16 (four lines per wave) or 64 (the wave-line
// layout, one line per wave, plain census-fed paths only).
// Block `block` of a launch over the direction table of a: its direction, its
// line group, then the group's scanlines.
template <int DPL, bool FROM_CENSUS, bool SAT, bool DECIM, bool BCAST = false, int LPL = kRowLanes>
__device__ __forceinline__ void path_block(const PathArgs &a, int block)
{
    constexpr int LPW = 64 / LPL;
    int k = 0;
    while (k + 1 < a.ndirs && block >= a.dir[k + 1].first_block)
        k++;
    const DirTask t = a.dir[k];
    const int i = block - t.first_block;
    const int n8 = (t.nlines + LPW * kXcds - 1) / (LPW * kXcds) * kXcds;
    const int grp = t.dx == 0 ? (i % kXcds) * (n8 / kXcds) + i / kXcds : i;
    if (grp * LPW >= t.nlines)
        return; // padding block
    path_group<DPL, FROM_CENSUS, SAT, DECIM, PathArgs, BCAST, LPL>(a, t, grp);
}

template <int DPL, bool FROM_CENSUS, bool SAT, bool DECIM, bool BCAST = false, int LPL = kRowLanes>
__global__ __launch_bounds__(64) void k_path(PathArgs a)
{
    path_block<DPL, FROM_CENSUS, SAT, DECIM, BCAST, LPL>(a, (int)blockIdx.x);
}

// ---------------------------------------------------------------------------
// Placed plain paths (census-fed, 4 directions, four lines per wave).
//
// A row line group is 1920 steps at 1080p against a column group's 1080, and
// a SIMD issues one wave64 VALU instruction per 4 cycles whichever wave it
// comes from.  The hardware places the one-wave workgroups of k_path by wave
// count (about 1.5 per SIMD at 1080p), so about half the SIMDs hold a row wave
// and a column wave: 0.28 ms of issue at 1080p D=64 against the 0.19 ms a
// balanced placement would need, and the launch takes 0.26 ms.  Here one
// 8-wave workgroup holds each CU (its LDS reservation keeps a second one off
// it), and each wave reads which SIMD it runs on: on the CU's first two SIMDs
// (three on the first `extra` CUs) the first wave takes one row group and the
// other waves leave; on the other SIMDs two waves take column groups.  Groups
// come from two queues (rows, columns); every wave that has done its share
// drains what is left of both, so the result does not depend on placement.
// The last wave out resets the queue counters for the next launch.
constexpr int kPlacedThreads = 512;          // 8 waves: two per SIMD
constexpr int kPlacedLds = 81 * 1024;        // dynamic LDS: a second one would not fit in 160 KB
constexpr int kQueueStride = 16;             // ints: one 64-byte line per counter

struct PlacedQ {
    int kr[2], kc[2]; // direction-table indices of the row and the column directions
    int gr, gc;       // line groups per row / column direction
    int nrow, ncol;   // row / column groups in all
    int extra;        // CUs with three row SIMDs
    int total_waves;  // waves in the launch
};

// The placed launch's wave loop: role from the SIMD, then groups from the two
// queues (run(k, grp) takes line group grp of direction-table entry k), then
// the count-out and the reset.  `run` is called from one site, so the path
// code it holds inlines.
template <class Run>
__device__ __forceinline__ void placed_loop(const PlacedQ &q, int *__restrict__ ctr, Run &&run)
{
    __shared__ int simd_waves[4];
    if (threadIdx.x < 4)
        simd_waves[threadIdx.x] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    // hwreg(HW_REG_HW_ID): SIMD_ID is bits 5:4
    const int simd = (int)((__builtin_amdgcn_s_getreg((31 << 11) | 4) >> 4) & 3);
    int slot = 0;
    if (lane == 0)
        slot = atomicAdd(&simd_waves[simd], 1);
    slot = __builtin_amdgcn_readfirstlane(__shfl(slot, 0));
    const int row_simds = 2 + ((int)blockIdx.x < q.extra ? 1 : 0);
    const bool row_role = simd < row_simds && slot == 0;
    const bool col_role = simd >= row_simds && slot < 2;
    // phases: 0 the row role's one row group, 1 column groups until none is
    // left, 2 row groups until none is left, 3 done
    int phase = row_role ? 0 : (col_role ? 1 : 3);
    while (phase < 3) {
        const int which = phase == 1 ? 1 : 0;
        int it = 0;
        if (lane == 0)
            it = atomicAdd(ctr + which * kQueueStride, 1);
        it = __builtin_amdgcn_readfirstlane(__shfl(it, 0));
        if (it >= (which == 0 ? q.nrow : q.ncol)) {
            phase++;
            continue;
        }
        if (phase == 0)
            phase = 1;
        const int g = which == 0 ? q.gr : q.gc, h = it >= g ? 1 : 0;
        run(which == 0 ? q.kr[h] : q.kc[h], it - h * g);
    }
    // Every wave counts itself out after its last pop (release: the pops are
    // done first); the last one resets the three counters, which no other wave
    // of this launch touches any more.
    if (lane == 0) {
        const int done = __hip_atomic_fetch_add(ctr + 2 * kQueueStride, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (done == q.total_waves - 1) {
            __hip_atomic_store(ctr, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(ctr + kQueueStride, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(ctr + 2 * kQueueStride, 0, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

template <int DPL, bool BCAST>
__global__ __launch_bounds__(kPlacedThreads) void k_path_placed(PathArgs a, PlacedQ q, int *__restrict__ ctr)
{
    placed_loop(q, ctr, [&](int k, int grp) { path_group<DPL, true, false, false, PathArgs, BCAST>(a, a.dir[k], grp); });
}

// The two multiscale stages' 8-path aggregations in one launch: they read
// different cost volumes (C2 already holds the merge with C1) and write
// different volumes, so nothing orders them.  Blocks [0, nhalf) run the
// half-scale stage, the rest the full-scale stage (decimated store).  The
// half-scale stage alone leaves the GPU short of waves and is bound by its
// lines' step chains; dispatched first, those chains run beside the
// full-scale work (measured at 1920x1080: aggregation 0.507 ms, against 0.568
// with the full-scale blocks first and 0.597 as two launches).  nhalf is a
// multiple of the XCD count, so each stage's blocks keep their XCD dealing.
template <int DPL>
__global__ __launch_bounds__(64) void k_path_ms(PathArgs full, PathArgs half, int nhalf)
{
    if ((int)blockIdx.x < nhalf)
        path_block<DPL, false, true, false>(half, (int)blockIdx.x);
    else
        path_block<DPL, false, true, true>(full, (int)blockIdx.x - nhalf);
}

template <int LPP, int NDIRS>
__global__ __launch_bounds__(256) void k_wta(const uint8_t *__restrict__ L, size_t dir_stride,
                                             int npix, uint8_t *__restrict__ disp)
{
    constexpr int PPW = 64 / LPP;
    const int wave = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    const int nwaves = (int)gridDim.x * 4;
    for (int b0 = wave * PPW * kWtaBatch; b0 < npix; b0 += nwaves * PPW * kWtaBatch)
        wta_batch<LPP, NDIRS>(L, dir_stride, npix, disp, b0, threadIdx.x & 63);
}

// windowed SGM: every direction of the frame in one launch, one line group
// (4 rows or 4 columns) per wave
template <int NB>
__device__ __forceinline__ void window_group(const WinArgs &a, int k, int grp, uint8_t *wlds)
{
    const DirTask &t = a.dir[k];
    if (t.dy == 0) {
        if (t.dx > 0)
            path_window_rows<NB, 1>(a, t, a.Ln[k], grp, wlds);
        else
            path_window_rows<NB, -1>(a, t, a.Ln[k], grp, wlds);
    } else {
        path_window_cols<NB>(a, t, a.Ln[k], grp, wlds);
    }
}

template <int NB>
__global__ __launch_bounds__(64) void k_path_window(WinArgs a)
{
    __shared__ __align__(16) uint8_t wlds[win_wave_lds<NB>()];
    int k = 0;
    while (k + 1 < a.ndirs && (int)blockIdx.x >= a.dir[k + 1].first_block)
        k++;
    window_group<NB>(a, k, (int)blockIdx.x - a.dir[k].first_block, wlds);
}

template <int NB>
__global__ __launch_bounds__(256) void k_wta_window(const uint8_t *__restrict__ L, size_t dir_stride,
                                                    const uint8_t *__restrict__ Ln, size_t nstride, int ndirs,
                                                    const uint8_t *__restrict__ blk, int tiles_x, int ntiles, int W,
                                                    int H, uint8_t *__restrict__ disp)
{
    // one workgroup per tile, wave w takes the w-th quarter of its passes
    for (int t = (int)blockIdx.x; t < ntiles; t += (int)gridDim.x)
        wta_window_tile<NB>(L, dir_stride, Ln, nstride, ndirs, blk, tiles_x, W, H, disp, t, (int)(threadIdx.x >> 6));
}

// ---------------------------------------------------------------------------
// streaming-copy probe: the achievable HBM ceiling and the reference size for
// scaling hardware byte counters (the pattern of the reference's
// tools/gpu_bandwidth.cu:12-25: a float4 copy, one element per thread)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                size_t n)
{
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride)
        dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// census of one image, or of a pair (img1 != nullptr) in one launch
static int launch_census(hipStream_t s, const uint8_t *img0, uint64_t *out0, int W, int H, int ostride,
                         const uint8_t *img1 = nullptr, uint64_t *out1 = nullptr)
{
    dim3 grid((W + kCenTX - 1) / kCenTX, (H + kCenTY - 1) / kCenTY, img1 ? 2 : 1);
    hipLaunchKernelGGL(k_census, grid, dim3(256), 0, s, img0, img1, out0, out1, W, H, ostride);
    CK(hipGetLastError());
    return 0;
}

template <bool FROM_CENSUS, bool SAT, bool DECIM, bool BCAST = false>
static int launch_path_t(hipStream_t s, const PathArgs &a, int blocks)
{
    switch (a.D / 16) {
    case 1: hipLaunchKernelGGL((k_path<1, FROM_CENSUS, SAT, DECIM, BCAST>), dim3(blocks), dim3(64), 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_path<2, FROM_CENSUS, SAT, DECIM, BCAST>), dim3(blocks), dim3(64), 0, s, a); break;
    case 4: hipLaunchKernelGGL((k_path<4, FROM_CENSUS, SAT, DECIM, BCAST>), dim3(blocks), dim3(64), 0, s, a); break;
    case 8: hipLaunchKernelGGL((k_path<8, FROM_CENSUS, SAT, DECIM, BCAST>), dim3(blocks), dim3(64), 0, s, a); break;
    case 16: hipLaunchKernelGGL((k_path<16, FROM_CENSUS, SAT, DECIM, BCAST>), dim3(blocks), dim3(64), 0, s, a); break;
    default: set_err("sgmh: bad D %d", a.D); return -1;
    }
    CK(hipGetLastError());
    return 0;
}

// the wave-line layout: plain census-fed paths, DPL = D/64
static int launch_path_wave(hipStream_t s, const PathArgs &a, int blocks)
{
    switch (a.D / 64) {
    case 1: hipLaunchKernelGGL((k_path<1, true, false, false, false, 64>), dim3(blocks), dim3(64), 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_path<2, true, false, false, false, 64>), dim3(blocks), dim3(64), 0, s, a); break;
    case 4: hipLaunchKernelGGL((k_path<4, true, false, false, false, 64>), dim3(blocks), dim3(64), 0, s, a); break;
    default: set_err("sgmh: the wave-line layout needs D = 64, 128 or 256 (D = %d)", a.D); return -1;
    }
    CK(hipGetLastError());
    return 0;
}

// Fill the direction table and launch: census-fed 4-path launches with the
// row directions last (see below), the others with the longest lines first,
// so the first blocks the dispatcher places are the longest-running ones.
// Two variants are instantiated: census-fed plain paths and cost-volume-fed
// saturating multiscale paths (full or decimated store).
// fill_paths builds a launch's direction table in a and returns its block
// count (-1 on a size error); launch_paths launches it.
static int fill_paths(PathArgs &a, const int (*dirs)[3], int ndirs, bool from_census, bool sat,
                      uint8_t *vol, size_t dir_stride, int decimate, bool wave_lines)
{
    REQUIRE(!wave_lines || (from_census && !sat && !decimate), "sgmh: the wave-line layout is for plain paths");
    const int lpw = wave_lines ? 1 : kLinesPerWave; // scanlines per wave
    // every byte offset the kernels form must fit the 32-bit buffer offsets
    const size_t in_bytes = (size_t)a.W * a.H * (from_census ? 8 : (size_t)a.D);
    REQUIRE(dir_stride < (1ull << 31) && in_bytes < (1ull << 31),
            "sgmh: %dx%d D=%d exceeds the 2 GiB per-volume limit", a.W, a.H, a.D);
    a.out_bytes = (uint32_t)dir_stride;
    int order[8];
    for (int k = 0; k < ndirs; k++)
        order[k] = k;
    auto len = [&](int k) { return dirs[k][1] == 0 ? a.W : (dirs[k][0] == 0 ? a.H : std::min(a.W, a.H)); };
    // Census-fed 4-path launches put the row directions last.  The hardware
    // places single-wave workgroups on SIMDs by wave count, not by role: with
    // the rows first it put two row waves on one SIMD in some frames (from a
    // few percent of them to nearly all, depending on the kernel that ran
    // before), and such a frame's paths take 0.345 instead of 0.26 ms at 1080p.  Placed after the columns,
    // the rows never shared a SIMD in the frames measured, at the same median
    // time.  With 8 paths the launch holds more waves than the SIMDs take at
    // once, and rows placed last would wait for a slot: those keep the rows
    // first.
    if (from_census && !sat && !decimate && ndirs <= 4)
        std::stable_sort(order, order + ndirs, [&](int p, int q) { return (dirs[p][1] == 0) < (dirs[q][1] == 0); });
    else
        std::stable_sort(order, order + ndirs, [&](int p, int q) { return len(p) > len(q); });
    int blocks = 0;
    a.ndirs = ndirs;
    for (int i = 0; i < ndirs; i++) {
        const int k = order[i];
        DirTask &t = a.dir[i];
        t.dx = dirs[k][0];
        t.dy = dirs[k][1];
        t.P1 = dirs[k][2];
        t.dec = decimate ? 1 : 0;
        t.nlines = decimate ? dec_lines(t.dx, t.dy, a.W, a.H)
                            : (t.dy == 0 ? a.H : (t.dx == 0 ? a.W : a.W + a.H - 1));
        t.first_block = blocks;
        t.L = vol + (size_t)k * dir_stride;
        // line groups of this direction, padded to a multiple of the XCD count
        const int groups = (t.nlines + lpw - 1) / lpw;
        blocks += (groups + kXcds - 1) / kXcds * kXcds;
    }
    return blocks;
}

static int launch_paths(hipStream_t s, PathArgs a, const int (*dirs)[3], int ndirs,
                        bool from_census, bool sat, uint8_t *vol, size_t dir_stride, int decimate,
                        bool wave_lines = false)
{
    const int blocks = fill_paths(a, dirs, ndirs, from_census, sat, vol, dir_stride, decimate, wave_lines);
    if (blocks < 0)
        return -1;
    if (wave_lines)
        return launch_path_wave(s, a, blocks);
    // the broadcast row variant pays where row waves are at least one per CU,
    // at D = 64 (measured; at D = 128 it is 1 % slower)
    if (from_census && !sat && !decimate)
        return a.H >= kRowBcastMinH && a.D == 64 ? launch_path_t<true, false, false, true>(s, a, blocks)
                                                 : launch_path_t<true, false, false>(s, a, blocks);
    if (!from_census && sat && !decimate)
        return launch_path_t<false, true, false>(s, a, blocks);
    if (!from_census && sat && decimate)
        return launch_path_t<false, true, true>(s, a, blocks);
    set_err("sgmh: unsupported path-kernel variant");
    return -1;
}

// Both multiscale stages' paths (k_path_ms): `full` over C2 with the
// decimated store into V2, `half` over C1 into V1.
static int launch_paths_ms(hipStream_t s, PathArgs full, PathArgs half, const int (*dirs)[3], uint8_t *V2,
                           uint8_t *V1, size_t dir_stride)
{
    const int nfull = fill_paths(full, dirs, 8, false, true, V2, dir_stride, 1, false);
    const int nhalf = nfull < 0 ? -1 : fill_paths(half, dirs, 8, false, true, V1, dir_stride, 0, false);
    if (nhalf < 0)
        return -1;
    const dim3 grid(nfull + nhalf);
    switch (full.D / 16) {
    case 1: hipLaunchKernelGGL(k_path_ms<1>, grid, dim3(64), 0, s, full, half, nhalf); break;
    case 2: hipLaunchKernelGGL(k_path_ms<2>, grid, dim3(64), 0, s, full, half, nhalf); break;
    case 4: hipLaunchKernelGGL(k_path_ms<4>, grid, dim3(64), 0, s, full, half, nhalf); break;
    case 8: hipLaunchKernelGGL(k_path_ms<8>, grid, dim3(64), 0, s, full, half, nhalf); break;
    case 16: hipLaunchKernelGGL(k_path_ms<16>, grid, dim3(64), 0, s, full, half, nhalf); break;
    default: set_err("sgmh: bad D %d", full.D); return -1;
    }
    CK(hipGetLastError());
    return 0;
}

template <int LPP>
static void launch_wta_t(hipStream_t s, int blocks, const uint8_t *vol, size_t dir_stride, int ndirs,
                         int n, uint8_t *disp)
{
    if (ndirs == 4)
        hipLaunchKernelGGL((k_wta<LPP, 4>), dim3(blocks), dim3(256), 0, s, vol, dir_stride, n, disp);
    else
        hipLaunchKernelGGL((k_wta<LPP, 8>), dim3(blocks), dim3(256), 0, s, vol, dir_stride, n, disp);
}

static int launch_wta(hipStream_t s, const uint8_t *vol, size_t dir_stride, int ndirs, size_t npix,
                      int D, uint8_t *disp)
{
    REQUIRE(ndirs >= 1 && ndirs <= 8 && npix < (1ull << 31), "sgmh: bad argmin shape");
    const int ppw = 64 / (D / 16);                    // pixels per wave per batch
    const size_t waves = (npix + (size_t)ppw * kWtaBatch - 1) / ((size_t)ppw * kWtaBatch);
    const int blocks = (int)std::min<size_t>((waves + 3) / 4, 8192);
    const int n = (int)npix;
    REQUIRE(ndirs == 4 || ndirs == 8, "sgmh: argmin over %d directions", ndirs);
    switch (D / 16) {
    case 1: launch_wta_t<1>(s, blocks, vol, dir_stride, ndirs, n, disp); break;
    case 2: launch_wta_t<2>(s, blocks, vol, dir_stride, ndirs, n, disp); break;
    case 4: launch_wta_t<4>(s, blocks, vol, dir_stride, ndirs, n, disp); break;
    case 8: launch_wta_t<8>(s, blocks, vol, dir_stride, ndirs, n, disp); break;
    case 16: launch_wta_t<16>(s, blocks, vol, dir_stride, ndirs, n, disp); break;
    default: set_err("sgmh: bad D %d", D); return -1;
    }
    CK(hipGetLastError());
    return 0;
}

static int launch_cost(hipStream_t s, const uint64_t *cl, const uint64_t *cr, int W, int H, int D,
                       uint32_t invalid, const uint8_t *merge, uint8_t *C)
{
    // the codes' byte offsets must fit the 32-bit buffer offsets
    REQUIRE((size_t)W * H * 8 < (1ull << 31), "sgmh: %dx%d exceeds the cost kernel's 2 GiB code limit", W, H);
    const dim3 blocks((W + 16 * kCostBatch - 1) / (16 * kCostBatch), H);
    switch (D / 16) {
    case 1: hipLaunchKernelGGL(k_cost<1>, blocks, dim3(256), 0, s, cl, cr, W, H, D, invalid, merge, C); break;
    case 2: hipLaunchKernelGGL(k_cost<2>, blocks, dim3(256), 0, s, cl, cr, W, H, D, invalid, merge, C); break;
    case 4: hipLaunchKernelGGL(k_cost<4>, blocks, dim3(256), 0, s, cl, cr, W, H, D, invalid, merge, C); break;
    case 8: hipLaunchKernelGGL(k_cost<8>, blocks, dim3(256), 0, s, cl, cr, W, H, D, invalid, merge, C); break;
    case 16: hipLaunchKernelGGL(k_cost<16>, blocks, dim3(256), 0, s, cl, cr, W, H, D, invalid, merge, C); break;
    default: set_err("sgmh: bad D %d", D); return -1;
    }
    CK(hipGetLastError());
    return 0;
}

// Device-buffer entry points run on the caller's stream; NULL is the HIP
// default stream, as in the HIP API itself.
static hipStream_t pick(sgmh_ctx *ctx, void *stream)
{
    (void)ctx;
    return (hipStream_t)stream;
}

// The automatic layout choice for plain paths, from a sweep of both layouts
// over 640x480 .. 1920x1080 and D = 64, 128, 256 on one MI355X (4-path
// kernel, ms, four lines / one line per wave):
//   D = 64:  0.076 / 0.097 at 640x480, 0.262 / 0.575 at 1080p -- never;
//   D = 128: 0.122 / 0.107 at 640x480, 0.316 / 0.300 at 1280x720,
//            0.473 / 0.664 at 1080p -- up to 1280x720;
//   D = 256: 0.242 / 0.167 at 640x480, 1.122 / 1.045 at 1080p -- always.
// Four lines per wave amortise the per-step work (the reduction over D, the
// loads, the shift) over four lines; one line per wave shortens each line's
// chain of steps, which pays where the lanes own many disparities (DPL = D/16
// of 8 or 16) or the frame holds too few lines to fill the GPU.
static bool wave_lines_pay(int W, int H, int D, int ndirs)
{
    if (ndirs > 4)
        return false;
    if (D == 256)
        return true;
    return D == 128 && (long)W * H <= kWaveLinesMaxPixelsD128;
}

// The placed launch pays where the four-line groups outnumber the SIMDs (so
// the one-wave-per-workgroup launch doubles some of them up) and a row step
// is long (D = 128).  Measured at 1920x1080, aggregate ms, four-line / placed:
// D = 128 0.477 / 0.434; D = 64 0.269 / 0.279 (the column SIMDs, two column
// waves each, then finish last); D = 32 0.166 / 0.221; 1280x720 D = 64 0.176
// / 0.202.  With two frames in flight its CU-wide LDS reservation also keeps
// the other frame's kernels off the CU: D = 64 0.317 / 0.380 ms per frame.
static bool placed_pays(sgmh_ctx *ctx, int W, int H, int D, int ndirs)
{
    if (!ctx || ndirs != 4 || ctx->cus <= 0 || D != 128)
        return false;
    const long groups = 2L * ((H + kLinesPerWave - 1) / kLinesPerWave) + 2L * ((W + kLinesPerWave - 1) / kLinesPerWave);
    return groups > 4L * ctx->cus && 2L * ((H + kLinesPerWave - 1) / kLinesPerWave) <= 3L * ctx->cus;
}

// The placed launches' queues over a 4-direction table (two row and two
// column directions).
static_assert(sizeof(int) * 3 * kQueueStride <= 256, "queue counters");
static int placed_queues(const sgmh_ctx *ctx, const DirTask *dir, int ndirs, int W, int H, PlacedQ &q)
{
    q = PlacedQ{};
    int nr = 0, nc = 0;
    for (int k = 0; k < ndirs; k++) {
        if (dir[k].dy == 0 && nr < 2)
            q.kr[nr++] = k;
        else if (dir[k].dx == 0 && nc < 2)
            q.kc[nc++] = k;
    }
    REQUIRE(ndirs == 4 && nr == 2 && nc == 2, "sgmh: the placed launch needs two row and two column directions");
    q.gr = (H + kLinesPerWave - 1) / kLinesPerWave;
    q.gc = (W + kLinesPerWave - 1) / kLinesPerWave;
    q.nrow = 2 * q.gr;
    q.ncol = 2 * q.gc;
    q.extra = std::max(0, std::min(ctx->cus, q.nrow - 2 * ctx->cus));
    q.total_waves = ctx->cus * (kPlacedThreads / 64);
    return 0;
}

static int launch_paths_placed(sgmh_ctx *ctx, hipStream_t s, PathArgs a, const int (*dirs)[3], uint8_t *vol,
                               size_t dir_stride)
{
    if (fill_paths(a, dirs, 4, true, false, vol, dir_stride, 0, false) < 0)
        return -1;
    PlacedQ q;
    if (placed_queues(ctx, a.dir, a.ndirs, a.W, a.H, q))
        return -1;
    int *ctr = (int *)ctx->queues.p;
    const dim3 grid(ctx->cus), block(kPlacedThreads);
    const bool bcast = a.H >= kRowBcastMinH && a.D == 64;
    switch (a.D / 16) {
    case 2: hipLaunchKernelGGL((k_path_placed<2, false>), grid, block, kPlacedLds, s, a, q, ctr); break;
    case 4:
        if (bcast)
            hipLaunchKernelGGL((k_path_placed<4, true>), grid, block, kPlacedLds, s, a, q, ctr);
        else
            hipLaunchKernelGGL((k_path_placed<4, false>), grid, block, kPlacedLds, s, a, q, ctr);
        break;
    case 8: hipLaunchKernelGGL((k_path_placed<8, false>), grid, block, kPlacedLds, s, a, q, ctr); break;
    default: set_err("sgmh: the placed path launch takes D = 32, 64 or 128 (D = %d)", a.D); return -1;
    }
    CK(hipGetLastError());
    return 0;
}

static int plain_paths(sgmh_ctx *ctx, hipStream_t s, const uint64_t *cl, const uint64_t *cr, int W,
                       int H, int D, const int (*dirs)[3], int ndirs, uint8_t *vol, size_t dir_stride)
{
    const int layout = ctx ? ctx->path_layout : SGMH_LAYOUT_AUTO;
    const bool wave = layout == SGMH_LAYOUT_WAVE || (layout == SGMH_LAYOUT_AUTO && wave_lines_pay(W, H, D, ndirs));
    PathArgs a{};
    a.cl = cl;
    a.cr = cr;
    a.W = W;
    a.H = H;
    a.D = D;
    a.P2 = kPlainP2;
    a.invalid = kPlainInvalid;
    const bool placed = ctx && ndirs == 4 && (D == 32 || D == 64 || D == 128) && layout != SGMH_LAYOUT_WAVE &&
                        (layout == SGMH_LAYOUT_PLACED || (layout == SGMH_LAYOUT_AUTO && !wave && placed_pays(ctx, W, H, D, ndirs)));
    if (placed)
        return launch_paths_placed(ctx, s, a, dirs, vol, dir_stride);
    return launch_paths(s, a, dirs, ndirs, true, false, vol, dir_stride, 0, wave);
}

// Census code buffers: (N + kLead) codes.  The kLead slack before pixel 0 is
// read by disparity windows that start left of x = 0 on row 0; it must hold
// valid (62-bit) codes there, so a new allocation gets it zeroed.
static int ensure_codes(DevBuf &b, size_t N)
{
    const bool fresh = b.cap < (N + kLead) * 8;
    if (ensure(b, (N + kLead) * 8))
        return -1;
    if (fresh)
        CK(hipMemset(b.p, 0, (size_t)kLead * 8));
    return 0;
}

// The plain path's device workspace: the two census code buffers, each
// kLead + N codes (the first kLead zero), and the direction volumes,
// paths * N * D bytes.
struct PlainWork {
    uint64_t *cl, *cr; // padded bases
    uint8_t *vol;
};

// The plain pipeline on explicit workspaces (ctx may be NULL: then nothing is timed).
static int plain_launch(sgmh_ctx *ctx, hipStream_t s, const uint8_t *l, const uint8_t *r, int W, int H,
                        int D, int paths, const PlainWork &w, uint8_t *disp)
{
    const size_t N = (size_t)W * H;
    {
        PhaseScope ps(ctx, s, PH_CENSUS);
        if (launch_census(s, l, w.cl + kLead, W, H, W, r, w.cr + kLead))
            return -1;
    }
    {
        PhaseScope ps(ctx, s, PH_AGG);
        int dirs[8][3];
        for (int k = 0; k < paths; k++) {
            dirs[k][0] = kDirsPlain[k][0];
            dirs[k][1] = kDirsPlain[k][1];
            dirs[k][2] = (int)kPlainP1;
        }
        if (plain_paths(ctx, s, w.cl, w.cr, W, H, D, dirs, paths, w.vol, N * D))
            return -1;
    }
    {
        PhaseScope ps(ctx, s, PH_ARGMIN);
        if (launch_wta(s, w.vol, N * D, paths, N, D, disp))
            return -1;
    }
    return 0;
}

static int plain_enqueue(sgmh_ctx *ctx, hipStream_t s, const uint8_t *l, const uint8_t *r, int W,
                         int H, int D, int paths, uint8_t *disp)
{
    const size_t N = (size_t)W * H;
    if (ensure_codes(ctx->cl, N) || ensure_codes(ctx->cr, N) ||
        ensure(ctx->vol, N * D * paths))
        return -1;
    const PlainWork w = {(uint64_t *)ctx->cl.p, (uint64_t *)ctx->cr.p, (uint8_t *)ctx->vol.p};
    return plain_launch(ctx, s, l, r, W, H, D, paths, w, disp);
}

// ---------------------------------------------------------------------------
// the plain path without allocation: its workspace is carved out of one
// static device array per GPU (for callers that cannot allocate)
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = SGMH_NOALLOC_SCRATCH_BYTES; // 128 MiB per GPU, held until the process ends
constexpr int kScratchDevices = 64;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

// The device address of `scratch` on each GPU (its name in host code is only a
// stand-in), resolved once per device with that device current.
static uint8_t *g_scratch_ptr[kScratchDevices];
static std::mutex g_scratch_mu;

static int scratch_on(int dev, uint8_t **out)
{
    REQUIRE(dev >= 0 && dev < kScratchDevices, "sgmh: device %d beyond the %d the scratch table holds", dev,
            kScratchDevices);
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        CK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        REQUIRE(p != nullptr, "sgmh: hipGetSymbolAddress gave no address for the scratch array");
        g_scratch_ptr[dev] = (uint8_t *)p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// Scratch layout of one no-allocation call: cl codes, cr codes, volumes.
static size_t noalloc_bytes(size_t N, int D, int paths)
{
    return 2 * align256((N + kLead) * 8) + (size_t)paths * N * D;
}

// zeroes the kLead codes before pixel 0 of both code buffers (a fill kernel,
// not a memset node, so that a captured graph replays it correctly)
__global__ __launch_bounds__(256) void k_zero_lead(uint64_t *__restrict__ a, uint64_t *__restrict__ b)
{
    const int i = (int)threadIdx.x;
    a[i] = 0;
    b[i] = 0;
}
static_assert(kLead == 256, "k_zero_lead clears kLead codes with 256 threads");

// The windowed path in two parts, so that the learned path can run the census
// beside the proposal network: window_prepare sizes the workspaces (it may
// allocate and synchronise, so it runs before any fork), window_census fills
// the context's codes, window_rest runs the paths and the argmin over them.
int window_prepare(sgmh_ctx *ctx, int W, int H, int D)
{
    const size_t N = (size_t)W * H;
    REQUIRE(N * D < (1ull << 31) && N * 32 < (1ull << 31), "sgmh_window: %dx%d D=%d exceeds the 2 GiB per-volume limit", W, H, D);
    if (ensure_codes(ctx->cl, N) || ensure_codes(ctx->cr, N) || ensure(ctx->vol, N * (D + 32) * 4))
        return -1;
    return 0;
}

int window_census(sgmh_ctx *ctx, hipStream_t s, const uint8_t *l, const uint8_t *r, int W, int H)
{
    uint64_t *cl = (uint64_t *)ctx->cl.p, *cr = (uint64_t *)ctx->cr.p;
    PhaseScope ps(ctx, s, PH_CENSUS);
    if (launch_census(s, l, cl + kLead, W, H, W, r, cr + kLead))
        return -1;
    return 0;
}

static int window_enqueue(sgmh_ctx *ctx, hipStream_t s, const uint8_t *l, const uint8_t *r, int W, int H,
                          int D, const uint8_t *blk, uint8_t *disp)
{
    if (window_prepare(ctx, W, H, D) || window_census(ctx, s, l, r, W, H))
        return -1;
    return window_rest(ctx, s, W, H, D, blk, disp);
}

int window_rest(sgmh_ctx *ctx, hipStream_t s, int W, int H, int D, const uint8_t *blk, uint8_t *disp)
{
    const size_t N = (size_t)W * H;
    uint64_t *cl = (uint64_t *)ctx->cl.p, *cr = (uint64_t *)ctx->cr.p;
    uint8_t *vol = (uint8_t *)ctx->vol.p;
    uint8_t *nvol = vol + N * D * 4; // narrow volumes, [N][32] per direction
    WinArgs a{};
    a.cl = cl;
    a.cr = cr;
    a.blk = blk;
    a.W = W;
    a.H = H;
    a.D = D;
    a.tiles_x = (W + 15) / 16;
    a.out_bytes = (uint32_t)(N * D);
    a.ndirs = 4;
    int blocks = 0;
    // columns first, rows last, as for the plain paths (measured: 1 % faster
    // here); the argmin sums the four volumes, so their order is free
    for (int k = 0; k < 4; k++) {
        const int o = (k + 2) & 3;
        DirTask &t = a.dir[k];
        t.dx = kDirsPlain[o][0];
        t.dy = kDirsPlain[o][1];
        t.P1 = (int)kPlainP1;
        t.nlines = t.dy == 0 ? H : W;
        t.dec = 0;
        t.first_block = blocks;
        t.L = vol + (size_t)k * N * D;
        a.Ln[k] = nvol + (size_t)k * N * 32;
        blocks += (t.nlines + kLinesPerWave - 1) / kLinesPerWave;
    }
    // One wave per workgroup.  A placed launch as for the plain D=128 paths
    // (rows alone on their SIMDs, two column waves on each other SIMD) was
    // measured slower here: aggregate ms at 1080p, one-wave / placed, the
    // network's proposal 0.333 / 0.377, every tile NARROW 0.191 / 0.251, every
    // tile FULL 0.546 / 0.503.  Two column waves on one SIMD issue at
    // about 1.65x one alone, and the NARROW steps are short.
    {
        PhaseScope ps(ctx, s, PH_AGG);
        switch (D / 16) {
        case 2: hipLaunchKernelGGL(k_path_window<2>, dim3(blocks), dim3(64), 0, s, a); break;
        case 4: hipLaunchKernelGGL(k_path_window<4>, dim3(blocks), dim3(64), 0, s, a); break;
        case 8: hipLaunchKernelGGL(k_path_window<8>, dim3(blocks), dim3(64), 0, s, a); break;
        default: set_err("sgmh_window: D=%d (32, 64 or 128)", D); return -1;
        }
        CK(hipGetLastError());
    }
    {
        PhaseScope ps(ctx, s, PH_ARGMIN);
        const int ntiles = a.tiles_x * ((H + 15) / 16);
        const int wb = ntiles;
        switch (D / 16) {
        case 2: hipLaunchKernelGGL(k_wta_window<2>, dim3(wb), dim3(256), 0, s, vol, N * D, nvol, N * 32, 4, blk, a.tiles_x, ntiles, W, H, disp); break;
        case 4: hipLaunchKernelGGL(k_wta_window<4>, dim3(wb), dim3(256), 0, s, vol, N * D, nvol, N * 32, 4, blk, a.tiles_x, ntiles, W, H, disp); break;
        default: hipLaunchKernelGGL(k_wta_window<8>, dim3(wb), dim3(256), 0, s, vol, N * D, nvol, N * 32, 4, blk, a.tiles_x, ntiles, W, H, disp); break;
        }
        CK(hipGetLastError());
    }
    return 0;
}

static int ms_enqueue(sgmh_ctx *ctx, hipStream_t s, const uint8_t *l, const uint8_t *r, int W, int H,
                      int D, uint8_t *disp, uint8_t *disp1)
{
    const int hw = W / 2, hh = H / 2;
    const size_t N = (size_t)W * H, Nh = (size_t)hw * hh;
    if (ensure(ctx->cl1, Nh * 8) || ensure(ctx->cr1, Nh * 8) ||
        ensure(ctx->cl, N * 8) || ensure(ctx->cr, N * 8) || ensure(ctx->cost1, Nh * D) ||
        ensure(ctx->cost2, N * D) || ensure(ctx->vol1, Nh * D * 8) || ensure(ctx->vol, Nh * D * 8) ||
        (disp1 == nullptr && ensure(ctx->disp1, Nh)))
        return -1;
    uint64_t *cl1 = (uint64_t *)ctx->cl1.p, *cr1 = (uint64_t *)ctx->cr1.p;
    uint64_t *cl2 = (uint64_t *)ctx->cl.p, *cr2 = (uint64_t *)ctx->cr.p;
    uint8_t *C1 = (uint8_t *)ctx->cost1.p, *C2 = (uint8_t *)ctx->cost2.p;
    uint8_t *V1 = (uint8_t *)ctx->vol1.p, *V2 = (uint8_t *)ctx->vol.p;
    uint8_t *d1 = disp1 ? disp1 : (uint8_t *)ctx->disp1.p;
    int dirs[8][3];
    memcpy(dirs, kDirsMs, sizeof dirs);
    {
        PhaseScope ps(ctx, s, PH_CENSUS);
        const dim3 grid((W + kCenTX - 1) / kCenTX, (H + kCenTY - 1) / kCenTY, 4);
        hipLaunchKernelGGL(k_census_ms, grid, dim3(256), 0, s, l, r, W, H, cl1, cr1, cl2, cr2);
        CK(hipGetLastError());
    }
    {
        PhaseScope ps(ctx, s, PH_COST);
        if (launch_cost(s, cl1, cr1, hw, hh, D, kMsInvalid, nullptr, C1) ||
            launch_cost(s, cl2, cr2, W, H, D, kMsInvalid, C1, C2))
            return -1;
    }
    PathArgs a{};
    a.D = D;
    a.P2 = kMsP2;
    a.invalid = kMsInvalid;
    {
        PhaseScope ps(ctx, s, PH_AGG);
        PathArgs half = a;
        half.cost = C1;
        half.W = hw;
        half.H = hh;
        a.cost = C2;
        a.W = W;
        a.H = H;
        if (launch_paths_ms(s, a, half, dirs, V2, V1, Nh * D))
            return -1;
    }
    {
        PhaseScope ps(ctx, s, PH_ARGMIN);
        if (launch_wta(s, V1, Nh * D, 8, Nh, D, d1) || launch_wta(s, V2, Nh * D, 8, Nh, D, disp))
            return -1;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char *sgmh_last_error(void) { return g_err; }

int sgmh_ctx_create(sgmh_ctx **out, int device)
{
    REQUIRE(out != nullptr, "sgmh_ctx_create: out is NULL");
    *out = nullptr;
    int n = 0;
    CK(hipGetDeviceCount(&n));
    REQUIRE(device >= 0 && device < n, "sgmh_ctx_create: device %d of %d", device, n);
    CK(hipSetDevice(device));
    // the no-allocation entry point's scratch address on this device, resolved
    // here at set-up (its first call on a device resolves it otherwise)
    uint8_t *scr = nullptr;
    if (scratch_on(device, &scr))
        return -1;
    sgmh_ctx *c = new sgmh_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        set_err("sgmh_ctx_create: hipStreamCreateWithFlags failed");
        delete c;
        return -1;
    }
    // the placed path launch's queue counters start at zero; each launch's
    // last wave resets them
    if (hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        ensure(c->queues, 256) || hipMemset(c->queues.p, 0, 256) != hipSuccess) {
        set_err("sgmh_ctx_create: queue counters");
        sgmh_ctx_destroy(c);
        return -1;
    }
    *out = c;
    return 0;
}

void sgmh_ctx_destroy(sgmh_ctx *c)
{
    if (!c)
        return;
    (void)hipSetDevice(c->device);
    (void)hipDeviceSynchronize();
    DevBuf *bufs[] = {&c->img_l, &c->img_r, &c->disp, &c->disp1, &c->cl, &c->cr, &c->vol,
                      &c->cl1, &c->cr1, &c->cost1, &c->cost2, &c->vol1, &c->queues};
    for (DevBuf *b : bufs)
        free_buf(*b);
    for (hipEvent_t e : c->events)
        (void)hipEventDestroy(e);
    if (c->stream)
        (void)hipStreamDestroy(c->stream);
    delete c;
}

int sgmh_set_timing(sgmh_ctx *c, int on)
{
    REQUIRE(c, "sgmh_set_timing: NULL context");
    c->timing = on != 0;
    return 0;
}

int sgmh_set_path_layout(sgmh_ctx *c, int layout)
{
    REQUIRE(c, "sgmh_set_path_layout: NULL context");
    REQUIRE(layout == SGMH_LAYOUT_AUTO || layout == SGMH_LAYOUT_LINES4 || layout == SGMH_LAYOUT_WAVE ||
                layout == SGMH_LAYOUT_PLACED,
            "sgmh_set_path_layout: unknown layout %d", layout);
    c->path_layout = layout;
    return 0;
}

int sgmh_get_times(sgmh_ctx *c, sgmh_times *t)
{
    REQUIRE(c && t, "sgmh_get_times: NULL argument");
    double ph[PH_N];
    for (int i = 0; i < PH_N; i++)
        ph[i] = -1.0;
    double total = 0.0;
    if (!c->marks.empty()) {
        CK(hipEventSynchronize(c->events[c->marks.back().second.second]));
        for (auto &m : c->marks) {
            float ms = 0.f;
            CK(hipEventElapsedTime(&ms, c->events[m.second.first], c->events[m.second.second]));
            ph[m.first] = (ph[m.first] < 0 ? 0.0 : ph[m.first]) + ms;
        }
        float all = 0.f;
        CK(hipEventElapsedTime(&all, c->events[c->marks.front().second.first],
                               c->events[c->marks.back().second.second]));
        total = all;
    }
    t->census_ms = ph[PH_CENSUS];
    t->cost_ms = ph[PH_COST];
    t->aggregate_ms = ph[PH_AGG];
    t->argmin_ms = ph[PH_ARGMIN];
    t->transfer_ms = ph[PH_XFER];
    t->total_ms = total;
    return 0;
}

int sgmh_plain_dev(sgmh_ctx *c, const uint8_t *l, const uint8_t *r, int W, int H, int D, int paths,
                   uint8_t *disp, void *stream)
{
    REQUIRE(c && l && r && disp, "sgmh_plain_dev: NULL argument");
    if (check_dims(W, H, D))
        return -1;
    REQUIRE(paths == 4 || paths == 8, "sgmh_plain_dev: paths=%d (4 or 8)", paths);
    CK(hipSetDevice(c->device));
    hipStream_t s = pick(c, stream);
    begin_call(c, s);
    return plain_enqueue(c, s, l, r, W, H, D, paths, disp);
}

int sgmh_plain(sgmh_ctx *c, const uint8_t *left, const uint8_t *right, int W, int H, int D,
               int paths, uint8_t *disp, sgmh_times *t)
{
    REQUIRE(c && left && right && disp, "sgmh_plain: NULL argument");
    if (check_dims(W, H, D))
        return -1;
    REQUIRE(paths == 4 || paths == 8, "sgmh_plain: paths=%d (4 or 8)", paths);
    CK(hipSetDevice(c->device));
    const size_t N = (size_t)W * H;
    if (ensure(c->img_l, N) || ensure(c->img_r, N) || ensure(c->disp, N))
        return -1;
    hipStream_t s = c->stream;
    begin_call(c, s);
    {
        PhaseScope ps(c, s, PH_XFER);
        CK(hipMemcpyAsync(c->img_l.p, left, N, hipMemcpyHostToDevice, s));
        CK(hipMemcpyAsync(c->img_r.p, right, N, hipMemcpyHostToDevice, s));
    }
    if (plain_enqueue(c, s, (const uint8_t *)c->img_l.p, (const uint8_t *)c->img_r.p, W, H, D,
                      paths, (uint8_t *)c->disp.p))
        return -1;
    {
        PhaseScope ps(c, s, PH_XFER);
        CK(hipMemcpyAsync(disp, c->disp.p, N, hipMemcpyDeviceToHost, s));
    }
    CK(hipStreamSynchronize(s));
    if (t)
        return sgmh_get_times(c, t);
    return 0;
}

int sgmh_plain_noalloc_dev(const uint8_t *l, const uint8_t *r, int W, int H, int D, int paths,
                           uint8_t *disp, void *stream)
{
    REQUIRE(l && r && disp, "sgmh_plain_noalloc_dev: NULL argument");
    if (check_dims(W, H, D))
        return -1;
    REQUIRE(paths == 4 || paths == 8, "sgmh_plain_noalloc_dev: paths=%d (4 or 8)", paths);
    const size_t N = (size_t)W * H, need = noalloc_bytes(N, D, paths);
    REQUIRE(need <= kScratchBytes,
            "sgmh_plain_noalloc_dev: %dx%d D=%d %d-path needs %zu bytes of scratch, more than the %zu "
            "the library holds per GPU",
            W, H, D, paths, need, kScratchBytes);
    int dev = 0;
    CK(hipGetDevice(&dev));
    uint8_t *base = nullptr;
    if (scratch_on(dev, &base))
        return -1;
    hipStream_t s = (hipStream_t)stream;
    PlainWork w;
    w.cl = (uint64_t *)base;
    w.cr = (uint64_t *)(base + align256((N + kLead) * 8));
    w.vol = base + 2 * align256((N + kLead) * 8);
    hipLaunchKernelGGL(k_zero_lead, dim3(1), dim3(256), 0, s, w.cl, w.cr);
    CK(hipGetLastError());
    return plain_launch(nullptr, s, l, r, W, H, D, paths, w, disp);
}

int sgmh_ms_dev(sgmh_ctx *c, const uint8_t *l, const uint8_t *r, int W, int H, int D, uint8_t *disp,
                uint8_t *disp1, void *stream)
{
    REQUIRE(c && l && r && disp, "sgmh_ms_dev: NULL argument");
    if (check_dims(W, H, D))
        return -1;
    REQUIRE(W % 2 == 0 && H % 2 == 0 && W >= 32 && H >= 32,
            "sgmh_ms_dev: multiscale needs even W, H >= 32 (got %dx%d)", W, H);
    CK(hipSetDevice(c->device));
    hipStream_t s = pick(c, stream);
    begin_call(c, s);
    return ms_enqueue(c, s, l, r, W, H, D, disp, disp1);
}

int sgmh_ms(sgmh_ctx *c, const uint8_t *left, const uint8_t *right, int W, int H, int D,
            uint8_t *disp, uint8_t *disp1, sgmh_times *t)
{
    REQUIRE(c && left && right && disp, "sgmh_ms: NULL argument");
    if (check_dims(W, H, D))
        return -1;
    REQUIRE(W % 2 == 0 && H % 2 == 0 && W >= 32 && H >= 32,
            "sgmh_ms: multiscale needs even W, H >= 32 (got %dx%d)", W, H);
    CK(hipSetDevice(c->device));
    const size_t N = (size_t)W * H, Nh = N / 4;
    if (ensure(c->img_l, N) || ensure(c->img_r, N) || ensure(c->disp, N) ||
        (disp1 && ensure(c->disp1, Nh)))
        return -1;
    hipStream_t s = c->stream;
    begin_call(c, s);
    {
        PhaseScope ps(c, s, PH_XFER);
        CK(hipMemcpyAsync(c->img_l.p, left, N, hipMemcpyHostToDevice, s));
        CK(hipMemcpyAsync(c->img_r.p, right, N, hipMemcpyHostToDevice, s));
    }
    if (ms_enqueue(c, s, (const uint8_t *)c->img_l.p, (const uint8_t *)c->img_r.p, W, H, D,
                   (uint8_t *)c->disp.p, disp1 ? (uint8_t *)c->disp1.p : nullptr))
        return -1;
    {
        PhaseScope ps(c, s, PH_XFER);
        CK(hipMemcpyAsync(disp, c->disp.p, Nh, hipMemcpyDeviceToHost, s));
        if (disp1)
            CK(hipMemcpyAsync(disp1, c->disp1.p, Nh, hipMemcpyDeviceToHost, s));
    }
    CK(hipStreamSynchronize(s));
    if (t)
        return sgmh_get_times(c, t);
    return 0;
}

int sgmh_census_dev(sgmh_ctx *c, const uint8_t *img, int W, int H, uint64_t *codes, void *stream)
{
    REQUIRE(c && img && codes, "sgmh_census_dev: NULL argument");
    REQUIRE(W >= 1 && H >= 1 && W <= 16384 && H <= 16384, "sgmh_census_dev: bad size %dx%d", W, H);
    CK(hipSetDevice(c->device));
    hipStream_t s = pick(c, stream);
    begin_call(c, s);
    return launch_census(s, img, codes, W, H, W);
}

int sgmh_sobel_dev(sgmh_ctx *c, const uint8_t *img, int W, int H, uint8_t *out, void *stream)
{
    REQUIRE(c && img && out, "sgmh_sobel_dev: NULL argument");
    REQUIRE(W >= 1 && H >= 1 && W <= 16384 && H <= 16384, "sgmh_sobel_dev: bad size %dx%d", W, H);
    CK(hipSetDevice(c->device));
    hipStream_t s = pick(c, stream);
    begin_call(c, s);
    hipLaunchKernelGGL(k_sobel, dim3((W + 63) / 64, (H + 3) / 4), dim3(256), 0, s, img, W, H, out);
    CK(hipGetLastError());
    return 0;
}

int sgmh_half_dev(sgmh_ctx *c, const uint8_t *img, int W, int H, uint8_t *out, void *stream)
{
    REQUIRE(c && img && out, "sgmh_half_dev: NULL argument");
    REQUIRE(W >= 2 && H >= 2 && W <= 16384 && H <= 16384, "sgmh_half_dev: bad size %dx%d", W, H);
    CK(hipSetDevice(c->device));
    hipStream_t s = pick(c, stream);
    begin_call(c, s);
    hipLaunchKernelGGL(k_half, dim3((W / 2 + 63) / 64, (H / 2 + 3) / 4), dim3(256), 0, s, img, W, H, out);
    CK(hipGetLastError());
    return 0;
}

int sgmh_path_plain_dev(sgmh_ctx *c, const uint64_t *cl, const uint64_t *cr, int W, int H, int D,
                        int dx, int dy, uint8_t *L, void *stream)
{
    REQUIRE(c && cl && cr && L, "sgmh_path_plain_dev: NULL argument");
    if (check_dims(W, H, D))
        return -1;
    REQUIRE(dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && (dx | dy) != 0,
            "sgmh_path_plain_dev: bad direction (%d,%d)", dx, dy);
    CK(hipSetDevice(c->device));
    hipStream_t s = pick(c, stream);
    begin_call(c, s);
    // copy the caller's dense codes behind the lead pad the kernels expect
    const size_t N = (size_t)W * H;
    if (ensure_codes(c->cl, N) || ensure_codes(c->cr, N))
        return -1;
    uint64_t *pl = (uint64_t *)c->cl.p, *pr = (uint64_t *)c->cr.p;
    CK(hipMemcpyAsync(pl + kLead, cl, N * 8, hipMemcpyDeviceToDevice, s));
    CK(hipMemcpyAsync(pr + kLead, cr, N * 8, hipMemcpyDeviceToDevice, s));
    const int dirs[1][3] = {{dx, dy, (int)kPlainP1}};
    return plain_paths(c, s, pl, pr, W, H, D, dirs, 1, L, N * D);
}

int sgmh_window_dev(sgmh_ctx *c, const uint8_t *l, const uint8_t *r, int W, int H, int D,
                    const uint8_t *blk, uint8_t *disp, void *stream)
{
    REQUIRE(c && l && r && blk && disp, "sgmh_window_dev: NULL argument");
    if (check_dims(W, H, D))
        return -1;
    REQUIRE(D == 32 || D == 64 || D == 128, "sgmh_window_dev: D=%d (32, 64 or 128)", D);
    CK(hipSetDevice(c->device));
    hipStream_t s = pick(c, stream);
    begin_call(c, s);
    return window_enqueue(c, s, l, r, W, H, D, blk, disp);
}

int sgmh_window(sgmh_ctx *c, const uint8_t *left, const uint8_t *right, int W, int H, int D,
                const uint8_t *blk, uint8_t *disp, sgmh_times *t)
{
    REQUIRE(c && left && right && blk && disp, "sgmh_window: NULL argument");
    if (check_dims(W, H, D))
        return -1;
    REQUIRE(D == 32 || D == 64 || D == 128, "sgmh_window: D=%d (32, 64 or 128)", D);
    const int ntiles = ((W + 15) / 16) * ((H + 15) / 16);
    for (int i = 0; i < ntiles; i++)
        REQUIRE((blk[i] & 15) <= (blk[i] >> 4) && (blk[i] >> 4) < D / 16,
                "sgmh_window: tile %d has block range %d..%d outside 0..%d", i, blk[i] & 15, blk[i] >> 4,
                D / 16 - 1);
    CK(hipSetDevice(c->device));
    const size_t N = (size_t)W * H;
    if (ensure(c->img_l, N) || ensure(c->img_r, N) || ensure(c->disp, N) || ensure(c->disp1, (size_t)ntiles))
        return -1;
    hipStream_t s = c->stream;
    begin_call(c, s);
    {
        PhaseScope ps(c, s, PH_XFER);
        CK(hipMemcpyAsync(c->img_l.p, left, N, hipMemcpyHostToDevice, s));
        CK(hipMemcpyAsync(c->img_r.p, right, N, hipMemcpyHostToDevice, s));
        CK(hipMemcpyAsync(c->disp1.p, blk, (size_t)ntiles, hipMemcpyHostToDevice, s));
    }
    if (window_enqueue(c, s, (const uint8_t *)c->img_l.p, (const uint8_t *)c->img_r.p, W, H, D,
                       (const uint8_t *)c->disp1.p, (uint8_t *)c->disp.p))
        return -1;
    {
        PhaseScope ps(c, s, PH_XFER);
        CK(hipMemcpyAsync(disp, c->disp.p, N, hipMemcpyDeviceToHost, s));
    }
    CK(hipStreamSynchronize(s));
    if (t)
        return sgmh_get_times(c, t);
    return 0;
}

int sgmh_copy_probe(sgmh_ctx *c, void *dst, const void *src, size_t bytes, void *stream)
{
    REQUIRE(c && dst && src && bytes % 16 == 0, "sgmh_copy_probe: bad arguments");
    CK(hipSetDevice(c->device));
    hipStream_t s = pick(c, stream);
    const size_t n = bytes / 16;
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 1u << 20);
    hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, s, (const uint4 *)src, (uint4 *)dst, n);
    CK(hipGetLastError());
    return 0;
}

} // extern "C"
