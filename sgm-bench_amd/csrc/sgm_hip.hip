// sgm_hip.hip -- SGM kernels for gfx950 and the C ABI declared in
// include/sgm_hip.h.
//
// Pipeline (plain path, common/sgm_ref.c:101-144):
//   k_census        9x7 census of both images from an LDS-staged tile
//   k_path<DPL,..>  every direction of the frame in ONE launch: one scanline
//                   per 16-lane DPP row, the Hamming cost computed on the fly
//                   from the census codes, the path cost L of that direction
//                   written as uint8 to its own volume
//   k_wta<DPL>      S = sum of the direction volumes (uint16 semantics) and
//                   argmin with the lowest d winning ties, fused
// Multiscale (multiscale/sgm_ms_ref.c:166-202) adds k_half, k_sobel and
// k_cost (materialised cost volume, merged across scales) and runs k_path in
// its cost-volume mode with saturation and direction-weighted P1.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <type_traits>
#include <vector>

#include "sgm_hip.h"
#include "sgm_kernels.h"

using namespace sgmk;

// ---------------------------------------------------------------------------
// error reporting
// ---------------------------------------------------------------------------
static thread_local char g_err[512] = "";

static void set_err(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

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

// ---------------------------------------------------------------------------
// parameters fixed by the reference
// ---------------------------------------------------------------------------
constexpr uint32_t kPlainP1 = 20;      // common/sgm_params.h:23
constexpr uint32_t kPlainP2 = 192;     // common/sgm_params.h:24
constexpr uint32_t kPlainInvalid = 62; // common/sgm_params.h:27-29
constexpr uint32_t kMsP2 = 200;        // multiscale/sgm_ms_ref.c:39
constexpr uint32_t kMsInvalid = 63;    // multiscale/sgm_ms_ref.c:41

// common/sgm_ref.c:96-99
static const int kDirsPlain[8][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1},
                                     {1, 1}, {-1, 1}, {1, -1}, {-1, -1}};
// multiscale/sgm_ms_ref.c:85-87: dx, dy, P1
static const int kDirsMs[8][3] = {{1, 0, 40}, {-1, 0, 40}, {0, 1, 20}, {0, -1, 20},
                                  {1, 1, 10}, {-1, -1, 10}, {1, -1, 10}, {-1, 1, 10}};

// ---------------------------------------------------------------------------
// census: 64x16 output tile per 256-thread block, each thread 4 consecutive x
// ---------------------------------------------------------------------------
constexpr int kCenTX = 64, kCenTY = 16;
constexpr int kCenRowBytes = kCenTX + 8;         // 72: x0-4 .. x0+67
constexpr int kCenRows = kCenTY + 6;             // 22: y0-3 .. y0+18
constexpr int kCenRowWords = kCenRowBytes / 4;   // 18

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__global__ __launch_bounds__(256) void k_census(const uint8_t *__restrict__ img,
                                                uint64_t *__restrict__ out, int W, int H,
                                                int ostride)
{
    __shared__ uint32_t tile[kCenRows * kCenRowWords];
    uint8_t *t8 = reinterpret_cast<uint8_t *>(tile);
    const int x0 = blockIdx.x * kCenTX, y0 = blockIdx.y * kCenTY;
    for (int i = threadIdx.x; i < kCenRows * kCenRowBytes; i += 256) {
        const int r = i / kCenRowBytes, c = i - r * kCenRowBytes;
        const int yy = clampi(y0 - 3 + r, 0, H - 1), xx = clampi(x0 - 4 + c, 0, W - 1);
        t8[i] = img[(size_t)yy * W + xx];
    }
    __syncthreads();
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int x = x0 + 4 * tx, y = y0 + ty;
    if (y >= H || x >= W)
        return;
    // bytes 4tx .. 4tx+11 of the 7 tile rows around this thread's row
    uint32_t win[7][3];
#pragma unroll
    for (int r = 0; r < 7; r++)
#pragma unroll
        for (int k = 0; k < 3; k++)
            win[r][k] = tile[(ty + r) * kCenRowWords + tx + k];
    uint64_t code[4];
#pragma unroll
    for (int p = 0; p < 4; p++) {
        const uint32_t centre = (win[3][(p + 4) >> 2] >> (8 * ((p + 4) & 3))) & 0xFFu;
        uint32_t lo = 0, hi = 0;
        int n = 0;
        // bit n: n-th neighbour in (dy outer -3..3, dx inner -4..4) order,
        // centre skipped (common/sgm_params.h:45-47)
#pragma unroll
        for (int r = 0; r < 7; r++)
#pragma unroll
            for (int c = 0; c < 9; c++) {
                if (r == 3 && c == 4)
                    continue;
                const int b = p + c; // byte index of (x+p + c-4) in the window
                const uint32_t nb = (win[r][b >> 2] >> (8 * (b & 3))) & 0xFFu;
                const uint32_t bit = nb < centre ? 1u : 0u;
                if (n < 32)
                    lo |= bit << n;
                else
                    hi |= bit << (n - 32);
                n++;
            }
        code[p] = ((uint64_t)hi << 32) | lo;
    }
    uint64_t *o = out + (size_t)y * ostride + x;
    if (x + 3 < W) {
        o[0] = code[0];
        o[1] = code[1];
        o[2] = code[2];
        o[3] = code[3];
    } else {
#pragma unroll
        for (int p = 0; p < 4; p++)
            if (x + p < W)
                o[p] = code[p];
    }
}

// ---------------------------------------------------------------------------
// multiscale pre-filters
// ---------------------------------------------------------------------------
// 2x2 box mean with +2 rounding, multiscale/sgm_ms_ref.c:145-150
__global__ __launch_bounds__(256) void k_half(const uint8_t *__restrict__ img, int W, int H,
                                              uint8_t *__restrict__ out)
{
    const int w = W / 2, h = H / 2;
    const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= w || y >= h)
        return;
    const uint8_t *a = img + (size_t)(2 * y) * W + 2 * x;
    out[(size_t)y * w + x] = (uint8_t)((a[0] + a[1] + a[W] + a[W + 1] + 2) >> 2);
}

// Sobel-X, out = clamp(128 + (gx >> 3)), multiscale/sgm_ms_ref.c:47-54
__global__ __launch_bounds__(256) void k_sobel(const uint8_t *__restrict__ img, int W, int H,
                                               uint8_t *__restrict__ out)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= W || y >= H)
        return;
    const int xl = clampi(x - 1, 0, W - 1), xr = clampi(x + 1, 0, W - 1);
    const uint8_t *up = img + (size_t)clampi(y - 1, 0, H - 1) * W;
    const uint8_t *md = img + (size_t)y * W;
    const uint8_t *dn = img + (size_t)clampi(y + 1, 0, H - 1) * W;
    const int gx = (up[xr] - up[xl]) + 2 * (md[xr] - md[xl]) + (dn[xr] - dn[xl]);
    out[(size_t)y * W + x] = (uint8_t)clampi(128 + (gx >> 3), 0, 255);
}

// Materialised cost volume (multiscale): one pixel per 16-lane row.
//   merge == nullptr: C = popcount(cl ^ cr[x-d]) or `invalid`      (sgm_ms_ref.c:66-72)
//   merge != nullptr: C = (C + merge[(y/2)*hw + x/2][d/2] + 1) >> 1 (sgm_ms_ref.c:184-188)
template <int DPL>
__global__ __launch_bounds__(256) void k_cost(const uint64_t *__restrict__ cl,
                                              const uint64_t *__restrict__ cr, int W, int H, int D,
                                              uint32_t invalid, const uint8_t *__restrict__ merge,
                                              uint8_t *__restrict__ C)
{
    const int lane = threadIdx.x & 63, g = lane >> 4, l = lane & 15;
    const size_t p = ((size_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 4 + g;
    const size_t N = (size_t)W * H;
    if (p >= N)
        return;
    const int y = (int)(p / W), x = (int)(p - (size_t)y * W);
    const uint64_t a = cl[p];
    const uint64_t *crow = cr + (size_t)y * W;
    uint32_t c[DPL];
#pragma unroll
    for (int j = 0; j < DPL; j++) {
        const int xr = x - (l * DPL + j);
        const uint64_t b = crow[xr < 0 ? 0 : xr];
        c[j] = xr < 0 ? invalid : (uint32_t)__popcll(a ^ b);
    }
    if (merge) {
        const int hw = W / 2;
        const uint8_t *m = merge + ((size_t)(y / 2) * hw + x / 2) * D;
#pragma unroll
        for (int j = 0; j < DPL; j++)
            c[j] = (c[j] + m[(l * DPL + j) >> 1] + 1) >> 1;
    }
    store_bytes<DPL>(C + p * D + l * DPL, c);
}

// ---------------------------------------------------------------------------
// path aggregation: all directions of a frame in one launch
// ---------------------------------------------------------------------------
// Memory goes through raw buffer resources: a load outside [0, num_records)
// returns 0 and a store there is dropped by the hardware.  The step loops are
// therefore straight-line code -- steps past the end of a scanline store to
// offset 0xFFFFFFFF, census reads left of x=0 come back as harmless values the
// invalid-cost select discards -- which lets the compiler count outstanding
// memory operations exactly and keep the prefetch ring in flight.
// Every volume a kernel addresses must stay below 2^31 bytes (checked on the
// host).

typedef __amdgpu_buffer_rsrc_t rsrc_t;
constexpr int kRsrcWord3 = 0x00020000; // raw, untyped buffer (gfx9 family)
constexpr uint32_t kDrop = 0xFFFFFFFFu;  // an offset no resource covers

__device__ __forceinline__ rsrc_t make_rsrc(const void *p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, kRsrcWord3);
}

__device__ __forceinline__ uint64_t ld64(rsrc_t r, uint32_t off)
{
    const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
    return ((uint64_t)v[1] << 32) | v[0];
}

// Store DPL packed bytes at byte offset `off` (kDrop: no store).
template <int DPL>
__device__ __forceinline__ void st_bytes(rsrc_t r, uint32_t off, const uint32_t (&v)[DPL])
{
    if constexpr (DPL == 1) {
        __builtin_amdgcn_raw_buffer_store_b8((uint8_t)v[0], r, off, 0, 0);
    } else if constexpr (DPL == 2) {
        __builtin_amdgcn_raw_buffer_store_b16((uint16_t)(v[0] | (v[1] << 8)), r, off, 0, 0);
    } else {
        uint32_t w[DPL / 4];
#pragma unroll
        for (int k = 0; k < DPL / 4; k++)
            w[k] = v[4 * k] | (v[4 * k + 1] << 8) | (v[4 * k + 2] << 16) | (v[4 * k + 3] << 24);
        if constexpr (DPL == 4) {
            __builtin_amdgcn_raw_buffer_store_b32(w[0], r, off, 0, 0);
        } else if constexpr (DPL == 8) {
            typedef uint32_t v2u __attribute__((ext_vector_type(2)));
            const v2u q = {w[0], w[1]};
            __builtin_amdgcn_raw_buffer_store_b64(q, r, off, 0, 0);
        } else {
#pragma unroll
            for (int k = 0; k < DPL / 4; k += 4) {
                typedef uint32_t v4u __attribute__((ext_vector_type(4)));
                const v4u q = {w[k], w[k + 1], w[k + 2], w[k + 3]};
                __builtin_amdgcn_raw_buffer_store_b128(q, r, off == kDrop ? kDrop : off + 4 * k, 0, 0);
            }
        }
    }
}

// Load DPL bytes at byte offset `off` as NW = max(1, DPL/4) words.
template <int DPL>
struct CostWords {
    static constexpr int NW = DPL < 4 ? 1 : DPL / 4;
    uint32_t w[NW];
};
template <int DPL>
__device__ __forceinline__ void ld_cost(rsrc_t r, uint32_t off, CostWords<DPL> &c)
{
    if constexpr (DPL == 1) {
        c.w[0] = __builtin_amdgcn_raw_buffer_load_b8(r, off, 0, 0);
    } else if constexpr (DPL == 2) {
        c.w[0] = __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0);
    } else if constexpr (DPL == 4) {
        c.w[0] = __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
    } else if constexpr (DPL == 8) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
        c.w[0] = v[0];
        c.w[1] = v[1];
    } else {
#pragma unroll
        for (int k = 0; k < DPL / 4; k += 4) {
            const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, off + 4 * k, 0, 0);
            c.w[k] = v[0];
            c.w[k + 1] = v[1];
            c.w[k + 2] = v[2];
            c.w[k + 3] = v[3];
        }
    }
}

struct DirTask {
    int dx, dy, P1;
    int nlines;      // scanlines in this direction
    int first_block; // first block of this direction in the grid
    uint8_t *L;      // output volume of this direction
};

// Census buffers read by the path kernels carry kLead codes of slack before
// pixel 0, so a disparity window that starts left of x = 0 on row 0 still has
// a non-negative offset (a load straddling offset 0 would otherwise be out of
// range as a whole and lose its valid half).
constexpr int kLead = 256;

struct PathArgs {
    const uint64_t *cl, *cr; // census codes; pixel i at index kLead + i
    const uint8_t *cost;     // or a materialised cost volume
    int W, H, D;
    uint32_t P2, invalid;
    uint32_t out_bytes;      // bytes of each direction's output volume
    int ndirs;
    DirTask dir[8];
};

struct LineGeom {
    int x, y, steps;
};

// Start pixel and length of scanline i of direction (dx, dy).  Lines of an
// axis-aligned direction are the rows / columns; a diagonal direction has
// W+H-1 lines starting on the edge row its predecessors fall off, then on the
// edge column (multiscale/sgm_ms_cuda.cu:91-98 enumerates the same set).
__device__ __forceinline__ LineGeom line_geom(int i, int dx, int dy, int W, int H)
{
    if (dy == 0)
        return {dx > 0 ? 0 : W - 1, i, W};
    if (dx == 0)
        return {i, dy > 0 ? 0 : H - 1, H};
    int x, y;
    if (i < W) {
        x = dx > 0 ? i : W - 1 - i;
        y = dy > 0 ? 0 : H - 1;
    } else {
        const int j = i - W + 1;
        x = dx > 0 ? 0 : W - 1;
        y = dy > 0 ? j : H - 1 - j;
    }
    const int sx = dx > 0 ? W - x : x + 1, sy = dy > 0 ? H - y : y + 1;
    return {x, y, sx < sy ? sx : sy};
}

__device__ __forceinline__ uint64_t dpp64_shr1(uint64_t old, uint64_t v)
{
    const uint32_t lo = dpp_or<kRowShr1>((uint32_t)old, (uint32_t)v);
    const uint32_t hi = dpp_or<kRowShr1>((uint32_t)(old >> 32), (uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t dpp64_shl1(uint64_t old, uint64_t v)
{
    const uint32_t lo = dpp_or<kRowShl1>((uint32_t)old, (uint32_t)v);
    const uint32_t hi = dpp_or<kRowShl1>((uint32_t)(old >> 32), (uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Horizontal direction DX (+1: L->R, -1: R->L), cost from census codes.
// Lane l keeps the window w[j] = cr[y][x - l*DPL - j] in registers; one step
// along the row shifts it by one disparity: the code leaving lane l's slice
// enters lane l+1's (L->R) or lane l-1's (R->L) through one row_shr:1 /
// row_shl:1 DPP move, and the end lane takes the one new code of the step.
// The window is a register ring indexed by the step modulo DPL (PF is a
// multiple of DPL), so the shift itself moves nothing.
// The path starts from Lp = 0, mn = 0: with P1 > 0 that step yields L = C
// exactly, the reference's first-pixel rule (sgm_ref.c:90-91), branch-free.
template <int DPL, bool SAT, int DX>
__device__ __forceinline__ void path_horizontal(const PathArgs &a, const DirTask &t, LineGeom ln,
                                                int l, int maxsteps)
{
    constexpr int PF = DPL > 8 ? DPL : 8;
    // Rows are the longest scanlines of the frame and set its critical path:
    // give their waves issue priority over the column / diagonal waves that
    // share their SIMD.
    __builtin_amdgcn_s_setprio(2);
    const int W = a.W, D = a.D;
    const uint32_t P1 = (uint32_t)t.P1, P2 = a.P2, inval = a.invalid;
    const uint32_t nb = (uint32_t)(a.W * a.H + kLead) * 8u;
    const rsrc_t rl = make_rsrc(a.cl, nb), rr = make_rsrc(a.cr, nb);
    const rsrc_t ro = make_rsrc(t.L, a.out_bytes);
    const int lbase = l * DPL;
    const int row = ln.y * W + kLead;           // code index of (0, y)
    const int orow = ln.y * W;                  // pixel index of (0, y)
    const int steps = ln.steps;
    int x = ln.x;

    // window one step before the start
    uint64_t phys[DPL];
#pragma unroll
    for (int j = 0; j < DPL; j++) {
        const int slot = DX > 0 ? (j + 1) % DPL : (j - 1 + DPL) % DPL;
        phys[slot] = ld64(rr, (uint32_t)(row + x - DX - lbase - j) * 8u);
    }
    // prefetched cl[x_s] and the window code entering at step s
    const int enter = DX > 0 ? 0 : -(D - 1);
    uint64_t pl[PF], pn[PF];
#pragma unroll
    for (int k = 0; k < PF; k++) {
        pl[k] = ld64(rl, (uint32_t)(row + x + k * DX) * 8u);
        pn[k] = ld64(rr, (uint32_t)(row + x + k * DX + enter) * 8u);
        __builtin_amdgcn_sched_barrier(0); // slot 0 first: its wait is the shortest
    }
    uint32_t L[DPL], mn = 0;
#pragma unroll
    for (int j = 0; j < DPL; j++)
        L[j] = 0;
    // a group past the last line stores nothing: its offset stays at kDrop
    uint32_t off = steps > 0 ? (uint32_t)(orow + x) * (uint32_t)D + (uint32_t)lbase : kDrop;
    const uint32_t ostep = steps > 0 ? (uint32_t)(DX * D) : 0u;

    // One block of PF steps.  CHECK: some step of the block may have x-d < 0
    // for a lane's disparities, or run past the end of the line.  Blocks
    // without either (all but ~D/PF + 1 of them) skip both selects.
    auto block = [&](int s0, auto check_tag) {
        constexpr bool CHECK = decltype(check_tag)::value;
#pragma unroll
        for (int k = 0; k < PF; k++) {
            const int s = s0 + k;
            // shift the window to step s
            if constexpr (DX > 0) {
                const int slot = ((-k) % DPL + DPL) % DPL;        // logical 0 at step s
                phys[slot] = dpp64_shr1(pn[k], phys[slot]);      // was logical DPL-1
            } else {
                const int slot = ((k - 1) % DPL + DPL) % DPL;     // logical DPL-1 at step s
                phys[slot] = dpp64_shl1(pn[k], phys[slot]);      // was logical 0
            }
            const uint64_t cL = pl[k];
            uint32_t C[DPL];
            const int rel = x - lbase;
#pragma unroll
            for (int j = 0; j < DPL; j++) {
                const int slot = DX > 0 ? (((j - k) % DPL) + DPL) % DPL : (j + k) % DPL;
                const uint32_t h = (uint32_t)__popcll(cL ^ phys[slot]);
                C[j] = CHECK ? (rel >= j ? h : inval) : h;
            }
            // ring slot k is dead now: refill it for step s + PF in place
            __builtin_amdgcn_sched_barrier(0);
            const int xp = x + PF * DX;
            pl[k] = ld64(rl, (uint32_t)(row + xp) * 8u);
            pn[k] = ld64(rr, (uint32_t)(row + xp + enter) * 8u);
            sgm_step<DPL, SAT>(L, C, mn, P1, P2);
            st_bytes<DPL>(ro, (!CHECK || s < steps) ? off : kDrop, L);
            off += ostep;
            x += DX;
            // keep each step's uses of the ring in program order, so a
            // prefetched value is consumed PF steps after its load was issued
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    // blocks [0, lo) and [hi, maxsteps) need the selects (lo, hi multiples of PF)
    // (every live row has maxsteps = W steps, so the bounds are wave-uniform)
    int lo, hi;
    if (DX > 0) {
        lo = ((D - 1) + PF - 1) / PF * PF;  // x = s < D-1 somewhere in the block
        hi = maxsteps / PF * PF;            // a block reaching past the line end
    } else {
        lo = 0;
        hi = std::min(maxsteps / PF, (W - D + 1) / PF) * PF; // x = W-1-s >= D-1 throughout
    }
    int s0 = 0;
    for (; s0 < lo && s0 < maxsteps; s0 += PF)
        block(s0, std::true_type{});
    for (; s0 < hi && s0 < maxsteps; s0 += PF)
        block(s0, std::false_type{});
    for (; s0 < maxsteps; s0 += PF)
        block(s0, std::true_type{});
}

// Any direction: cost from census codes read per lane (FROM_CENSUS), or from a
// materialised uint8 cost volume, with a prefetch ring PF steps deep.  DECIM
// stores L only at even (x, y), into a (W/2)x(H/2) layout (the multiscale
// argmin reads S only there, sgm_ms_ref.c:194-199).
template <int DPL, bool FROM_CENSUS, bool SAT, bool DECIM, bool VERT>
__device__ __forceinline__ void path_generic(const PathArgs &a, const DirTask &t, LineGeom ln, int l,
                                             int maxsteps)
{
    constexpr int PF = DPL <= 4 ? 8 : (DPL <= 8 ? 4 : 2);
    const int W = a.W, D = a.D, dx = t.dx, dy = t.dy;
    const uint32_t P1 = (uint32_t)t.P1, P2 = a.P2, inval = a.invalid;
    const int lbase = l * DPL;
    const int hw = W >> 1;
    const int steps = ln.steps;
    const rsrc_t ro = make_rsrc(t.L, a.out_bytes);
    rsrc_t rl, rr, rc;
    if constexpr (FROM_CENSUS) {
        const uint32_t nb = (uint32_t)(W * a.H + kLead) * 8u;
        rl = make_rsrc(a.cl, nb);
        rr = make_rsrc(a.cr, nb);
    } else {
        rc = make_rsrc(a.cost, (uint32_t)(W * a.H) * (uint32_t)D);
    }
    const int pstep = dy * W + dx;
    int x = ln.x;
    int pix = ln.y * W + ln.x;
    int y = ln.y;
    // a group past the last line stores nothing
    const uint32_t dead = steps > 0 ? 0u : kDrop;
    // On a column x is fixed, so which disparities have x-d < 0 is fixed too:
    // cost = max(hamming, inv[j]) with inv = invalid there and 0 elsewhere
    // (the invalid cost 62 is at least any Hamming distance of 62-bit codes).
    uint32_t inv[DPL];
#pragma unroll
    for (int j = 0; j < DPL; j++)
        inv[j] = x - lbase >= j ? 0u : inval;

    // ring[k] holds the inputs of step s with s % PF == k
    uint64_t rq[PF][FROM_CENSUS ? DPL + 1 : 1];
    CostWords<DPL> rw[PF];
    auto fetch = [&](int p, int k) {
        if constexpr (FROM_CENSUS) {
            rq[k][DPL] = ld64(rl, (uint32_t)(p + kLead) * 8u);
            // the lane's DPL window codes cr[p - lbase - j] are contiguous and
            // descending in j: fetch them 16 bytes at a time
            const uint32_t lo = (uint32_t)(p + kLead - lbase - (DPL - 1)) * 8u;
            if constexpr (DPL == 1) {
                rq[k][0] = ld64(rr, lo);
            } else {
#pragma unroll
                for (int i = 0; i < DPL; i += 2) {
                    const auto q = __builtin_amdgcn_raw_buffer_load_b128(rr, lo + 8u * i, 0, 0);
                    rq[k][DPL - 1 - i] = ((uint64_t)q[1] << 32) | q[0];
                    rq[k][DPL - 2 - i] = ((uint64_t)q[3] << 32) | q[2];
                }
            }
        } else {
            ld_cost<DPL>(rc, (uint32_t)p * (uint32_t)D + (uint32_t)lbase, rw[k]);
        }
    };
#pragma unroll
    for (int k = 0; k < PF; k++) {
        fetch(pix + k * pstep, k);
        __builtin_amdgcn_sched_barrier(0); // slot 0 first: its wait is the shortest
    }
    uint32_t L[DPL], mn = 0;
#pragma unroll
    for (int j = 0; j < DPL; j++)
        L[j] = 0;

    // One block of PF steps; CHECK: a step of the block may lie past the end
    // of its line (always possible on diagonals, whose lines differ in length).
    auto block = [&](int s0, auto check_tag) {
        constexpr bool CHECK = decltype(check_tag)::value;
#pragma unroll
        for (int k = 0; k < PF; k++) {
            const int s = s0 + k;
            uint32_t C[DPL];
            if constexpr (FROM_CENSUS) {
                const uint64_t cL = rq[k][DPL];
                const int rel = x - lbase;
#pragma unroll
                for (int j = 0; j < DPL; j++) {
                    const uint32_t h = (uint32_t)__popcll(cL ^ rq[k][j]);
                    if constexpr (VERT)
                        C[j] = umin(255u, h > inv[j] ? h : inv[j]);
                    else
                        C[j] = rel >= j ? h : inval;
                }
            } else {
#pragma unroll
                for (int j = 0; j < DPL; j++)
                    C[j] = (rw[k].w[j / 4] >> (8 * (j % 4))) & 0xFFu;
            }
            // ring slot k is dead now: refill it for step s + PF in place
            __builtin_amdgcn_sched_barrier(0);
            fetch(pix + PF * pstep, k);
            sgm_step<DPL, SAT>(L, C, mn, P1, P2);
            uint32_t off;
            if constexpr (!DECIM) {
                off = (uint32_t)pix * (uint32_t)D + (uint32_t)lbase;
            } else {
                off = (((x | y) & 1) == 0)
                          ? (uint32_t)((y >> 1) * hw + (x >> 1)) * (uint32_t)D + (uint32_t)lbase
                          : kDrop;
            }
            if constexpr (CHECK)
                off = s < steps ? off : kDrop;
            st_bytes<DPL>(ro, off | dead, L);
            pix += pstep;
            x += dx;
            y += dy;
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    // axis-aligned lines all have maxsteps steps: only a final partial block
    // needs the end-of-line select
    const int full = (dx == 0 || dy == 0) ? maxsteps / PF * PF : 0;
    int s0 = 0;
    for (; s0 < full; s0 += PF)
        block(s0, std::false_type{});
    for (; s0 < maxsteps; s0 += PF)
        block(s0, std::true_type{});
}

// Workgroups are dealt to the 8 XCDs round-robin by id, and each XCD has its
// own L2.  Every direction's block range starts at a multiple of 8 and is
// padded to one.  For columns, block i serves line group
// (i % 8) * (n8 / 8) + i / 8: each XCD owns one contiguous slice of the
// columns, so neighbouring columns -- whose 64-code disparity windows overlap
// -- share an L2.  Rows gain nothing from it, and diagonals must stay
// interleaved (their lengths vary along the range), so they keep block order.
constexpr int kXcds = 8;

template <int DPL, bool FROM_CENSUS, bool SAT, bool DECIM>
__global__ __launch_bounds__(64) void k_path(PathArgs a)
{
    const int lane = threadIdx.x, g = lane >> 4, l = lane & 15;
    int k = 0;
    while (k + 1 < a.ndirs && (int)blockIdx.x >= a.dir[k + 1].first_block)
        k++;
    const DirTask t = a.dir[k];
    const int i = (int)blockIdx.x - t.first_block;
    const int n8 = (t.nlines + kLinesPerWave * kXcds - 1) / (kLinesPerWave * kXcds) * kXcds;
    const int grp = t.dx == 0 ? (i % kXcds) * (n8 / kXcds) + i / kXcds : i;
    const int base = grp * kLinesPerWave;
    if (base >= t.nlines)
        return; // padding block
    int maxsteps = 0;
#pragma unroll
    for (int q = 0; q < kLinesPerWave; q++)
        if (base + q < t.nlines)
            maxsteps = max(maxsteps, line_geom(base + q, t.dx, t.dy, a.W, a.H).steps);
    const int line = base + g;
    LineGeom ln = {0, 0, 0};
    if (line < t.nlines)
        ln = line_geom(line, t.dx, t.dy, a.W, a.H);
    if constexpr (FROM_CENSUS && !DECIM) {
        if (t.dy == 0) {
            if (t.dx > 0)
                path_horizontal<DPL, SAT, 1>(a, t, ln, l, maxsteps);
            else
                path_horizontal<DPL, SAT, -1>(a, t, ln, l, maxsteps);
            return;
        }
    }
    if (FROM_CENSUS && t.dx == 0)
        path_generic<DPL, FROM_CENSUS, SAT, DECIM, true>(a, t, ln, l, maxsteps);
    else
        path_generic<DPL, FROM_CENSUS, SAT, DECIM, false>(a, t, ln, l, maxsteps);
}

// ---------------------------------------------------------------------------
// S = sum over directions, argmin (lowest d wins ties), sgm_ref.c:117-124
// ---------------------------------------------------------------------------
// A lane owns 16 consecutive disparities of one pixel (one 16-byte load per
// direction), so a pixel spans LPP = D/16 lanes and a wave covers 64/LPP
// pixels.  Sums stay packed two per 32-bit word (S <= 8*255 fits 16 bits):
// even bytes in one word, odd bytes in another.  U batches of pixels are
// loaded before any is reduced, to keep enough bytes in flight per wave.
template <int LPP>
__device__ __forceinline__ uint32_t group_min(uint32_t v)
{
    if constexpr (LPP >= 2)
        v = umin(v, dpp<kQuadSwap1>(v));
    if constexpr (LPP >= 4)
        v = umin(v, dpp<kQuadSwap2>(v));
    if constexpr (LPP >= 8)
        v = umin(v, dpp<kRowHalfMirror>(v));
    if constexpr (LPP >= 16)
        v = umin(v, dpp<kRowMirror>(v));
    return v;
}

template <int LPP, int NDIRS>
__global__ __launch_bounds__(256) void k_wta(const uint8_t *__restrict__ L, size_t dir_stride,
                                             int npix, uint8_t *__restrict__ disp)
{
    constexpr int D = 16 * LPP, PPW = 64 / LPP, U = 2;
    const int lane = threadIdx.x & 63, sub = lane % LPP, slot = lane / LPP;
    const int wave = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    const int nwaves = (int)gridDim.x * 4;
    for (int b0 = wave * PPW * U; b0 < npix; b0 += nwaves * PPW * U) {
        uint4 v[U][8];
        int pix[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int p = b0 + u * PPW + slot;
            pix[u] = p;
            const int pc = p < npix ? p : npix - 1;
            const uint8_t *src = L + (size_t)pc * D + sub * 16;
#pragma unroll
            for (int k = 0; k < 8; k++)
                if (k < NDIRS)
                    v[u][k] = *reinterpret_cast<const uint4 *>(src + k * dir_stride);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            uint32_t ev[4] = {0, 0, 0, 0}, od[4] = {0, 0, 0, 0};
#pragma unroll
            for (int k = 0; k < 8; k++)
                if (k < NDIRS) {
                    const uint32_t w[4] = {v[u][k].x, v[u][k].y, v[u][k].z, v[u][k].w};
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        ev[q] += w[q] & 0x00FF00FFu;
                        od[q] += (w[q] >> 8) & 0x00FF00FFu;
                    }
                }
            // key = S << 8 | d: the smallest key has the smallest S, then the smallest d
            const uint32_t d0 = (uint32_t)sub * 16;
            uint32_t key = 0xFFFFFFFFu;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t db = d0 + 4 * q;
                key = umin(key, umin(((ev[q] & 0xFFFFu) << 8) | db, ((od[q] & 0xFFFFu) << 8) | (db + 1)));
                key = umin(key, umin(((ev[q] >> 16) << 8) | (db + 2), ((od[q] >> 16) << 8) | (db + 3)));
            }
            key = group_min<LPP>(key);
            if (sub == 0 && pix[u] < npix)
                disp[pix[u]] = (uint8_t)(key & 0xFFu);
        }
    }
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
enum Phase { PH_CENSUS = 0, PH_COST, PH_AGG, PH_ARGMIN, PH_XFER, PH_N };

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

struct sgmh_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool timing = false;
    std::vector<hipEvent_t> events;
    size_t ev_used = 0;
    std::vector<std::pair<int, std::pair<int, int>>> marks; // phase, ev begin, ev end
    hipStream_t last_stream = nullptr;
    // workspaces
    DevBuf img_l, img_r, disp, disp1;          // host-buffer entry points
    DevBuf cl, cr;                             // census codes
    DevBuf vol;                                // direction volumes
    DevBuf half_l, half_r, sob_l, sob_r;       // multiscale images
    DevBuf cl1, cr1, cost1, cost2, vol1;       // multiscale stage buffers
};

static int ensure(DevBuf &b, size_t bytes)
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

static void free_buf(DevBuf &b)
{
    if (b.p)
        (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

// Phase timing: events are recorded only when timing is on.
struct PhaseScope {
    sgmh_ctx *ctx;
    hipStream_t s;
    int phase, ev0 = -1;
    PhaseScope(sgmh_ctx *c, hipStream_t st, int ph) : ctx(c), s(st), phase(ph)
    {
        if (ctx->timing)
            ev0 = record();
    }
    int record()
    {
        if (ctx->ev_used == ctx->events.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess)
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
        if (ctx->timing && ev0 >= 0) {
            const int ev1 = record();
            if (ev1 >= 0)
                ctx->marks.push_back({phase, {ev0, ev1}});
        }
    }
};

static void begin_call(sgmh_ctx *ctx, hipStream_t s)
{
    ctx->ev_used = 0;
    ctx->marks.clear();
    ctx->last_stream = s;
}

static int check_dims(int W, int H, int D)
{
    REQUIRE(W >= 16 && H >= 16 && W <= 16384 && H <= 16384,
            "sgmh: image %dx%d outside 16..16384", W, H);
    REQUIRE(D == 16 || D == 32 || D == 64 || D == 128 || D == 256,
            "sgmh: D=%d unsupported (16, 32, 64, 128 or 256)", D);
    return 0;
}

static int launch_census(hipStream_t s, const uint8_t *img, uint64_t *out, int W, int H,
                         int ostride)
{
    dim3 grid((W + kCenTX - 1) / kCenTX, (H + kCenTY - 1) / kCenTY);
    hipLaunchKernelGGL(k_census, grid, dim3(256), 0, s, img, out, W, H, ostride);
    CK(hipGetLastError());
    return 0;
}

template <bool FROM_CENSUS, bool SAT, bool DECIM>
static int launch_path_t(hipStream_t s, const PathArgs &a, int blocks)
{
    switch (a.D / 16) {
    case 1: hipLaunchKernelGGL((k_path<1, FROM_CENSUS, SAT, DECIM>), dim3(blocks), dim3(64), 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_path<2, FROM_CENSUS, SAT, DECIM>), dim3(blocks), dim3(64), 0, s, a); break;
    case 4: hipLaunchKernelGGL((k_path<4, FROM_CENSUS, SAT, DECIM>), dim3(blocks), dim3(64), 0, s, a); break;
    case 8: hipLaunchKernelGGL((k_path<8, FROM_CENSUS, SAT, DECIM>), dim3(blocks), dim3(64), 0, s, a); break;
    case 16: hipLaunchKernelGGL((k_path<16, FROM_CENSUS, SAT, DECIM>), dim3(blocks), dim3(64), 0, s, a); break;
    default: set_err("sgmh: bad D %d", a.D); return -1;
    }
    CK(hipGetLastError());
    return 0;
}

// Fill the direction table (longest lines first, so the first blocks the
// dispatcher places are the longest-running ones) and launch.  Two variants
// are instantiated: census-fed plain paths and cost-volume-fed saturating
// multiscale paths (full or decimated store).
static int launch_paths(hipStream_t s, PathArgs a, const int (*dirs)[3], int ndirs,
                        bool from_census, bool sat, uint8_t *vol, size_t dir_stride, int decimate)
{
    // every byte offset the kernels form must fit the 32-bit buffer offsets
    const size_t in_bytes = (size_t)a.W * a.H * (from_census ? 8 : (size_t)a.D);
    REQUIRE(dir_stride < (1ull << 31) && in_bytes < (1ull << 31),
            "sgmh: %dx%d D=%d exceeds the 2 GiB per-volume limit", a.W, a.H, a.D);
    a.out_bytes = (uint32_t)dir_stride;
    int order[8];
    for (int k = 0; k < ndirs; k++)
        order[k] = k;
    auto len = [&](int k) { return dirs[k][1] == 0 ? a.W : (dirs[k][0] == 0 ? a.H : std::min(a.W, a.H)); };
    std::stable_sort(order, order + ndirs, [&](int p, int q) { return len(p) > len(q); });
    int blocks = 0;
    a.ndirs = ndirs;
    for (int i = 0; i < ndirs; i++) {
        const int k = order[i];
        DirTask &t = a.dir[i];
        t.dx = dirs[k][0];
        t.dy = dirs[k][1];
        t.P1 = dirs[k][2];
        t.nlines = t.dy == 0 ? a.H : (t.dx == 0 ? a.W : a.W + a.H - 1);
        t.first_block = blocks;
        t.L = vol + (size_t)k * dir_stride;
        // line groups of this direction, padded to a multiple of the XCD count
        const int groups = (t.nlines + kLinesPerWave - 1) / kLinesPerWave;
        blocks += (groups + kXcds - 1) / kXcds * kXcds;
    }
    if (from_census && !sat && !decimate)
        return launch_path_t<true, false, false>(s, a, blocks);
    if (!from_census && sat && !decimate)
        return launch_path_t<false, true, false>(s, a, blocks);
    if (!from_census && sat && decimate)
        return launch_path_t<false, true, true>(s, a, blocks);
    set_err("sgmh: unsupported path-kernel variant");
    return -1;
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
    const size_t waves = (npix + (size_t)ppw * 2 - 1) / ((size_t)ppw * 2);
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
    const size_t N = (size_t)W * H;
    const int blocks = (int)((N + 15) / 16);
    switch (D / 16) {
    case 1: hipLaunchKernelGGL(k_cost<1>, dim3(blocks), dim3(256), 0, s, cl, cr, W, H, D, invalid, merge, C); break;
    case 2: hipLaunchKernelGGL(k_cost<2>, dim3(blocks), dim3(256), 0, s, cl, cr, W, H, D, invalid, merge, C); break;
    case 4: hipLaunchKernelGGL(k_cost<4>, dim3(blocks), dim3(256), 0, s, cl, cr, W, H, D, invalid, merge, C); break;
    case 8: hipLaunchKernelGGL(k_cost<8>, dim3(blocks), dim3(256), 0, s, cl, cr, W, H, D, invalid, merge, C); break;
    case 16: hipLaunchKernelGGL(k_cost<16>, dim3(blocks), dim3(256), 0, s, cl, cr, W, H, D, invalid, merge, C); break;
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

static int plain_paths(sgmh_ctx *ctx, hipStream_t s, const uint64_t *cl, const uint64_t *cr, int W,
                       int H, int D, const int (*dirs)[3], int ndirs, uint8_t *vol, size_t dir_stride)
{
    (void)ctx;
    PathArgs a{};
    a.cl = cl;
    a.cr = cr;
    a.W = W;
    a.H = H;
    a.D = D;
    a.P2 = kPlainP2;
    a.invalid = kPlainInvalid;
    return launch_paths(s, a, dirs, ndirs, true, false, vol, dir_stride, 0);
}

static int plain_enqueue(sgmh_ctx *ctx, hipStream_t s, const uint8_t *l, const uint8_t *r, int W,
                         int H, int D, int paths, uint8_t *disp)
{
    const size_t N = (size_t)W * H;
    if (ensure(ctx->cl, (N + kLead) * 8) || ensure(ctx->cr, (N + kLead) * 8) ||
        ensure(ctx->vol, N * D * paths))
        return -1;
    uint64_t *cl = (uint64_t *)ctx->cl.p, *cr = (uint64_t *)ctx->cr.p; // padded bases
    uint8_t *vol = (uint8_t *)ctx->vol.p;
    {
        PhaseScope ps(ctx, s, PH_CENSUS);
        if (launch_census(s, l, cl + kLead, W, H, W) || launch_census(s, r, cr + kLead, W, H, W))
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
        if (plain_paths(ctx, s, cl, cr, W, H, D, dirs, paths, vol, N * D))
            return -1;
    }
    {
        PhaseScope ps(ctx, s, PH_ARGMIN);
        if (launch_wta(s, vol, N * D, paths, N, D, disp))
            return -1;
    }
    return 0;
}

static int ms_enqueue(sgmh_ctx *ctx, hipStream_t s, const uint8_t *l, const uint8_t *r, int W, int H,
                      int D, uint8_t *disp, uint8_t *disp1)
{
    const int hw = W / 2, hh = H / 2;
    const size_t N = (size_t)W * H, Nh = (size_t)hw * hh;
    if (ensure(ctx->half_l, Nh) || ensure(ctx->half_r, Nh) || ensure(ctx->sob_l, N) ||
        ensure(ctx->sob_r, N) || ensure(ctx->cl1, Nh * 8) || ensure(ctx->cr1, Nh * 8) ||
        ensure(ctx->cl, N * 8) || ensure(ctx->cr, N * 8) || ensure(ctx->cost1, Nh * D) ||
        ensure(ctx->cost2, N * D) || ensure(ctx->vol1, Nh * D * 8) || ensure(ctx->vol, Nh * D * 8) ||
        (disp1 == nullptr && ensure(ctx->disp1, Nh)))
        return -1;
    uint8_t *hl = (uint8_t *)ctx->half_l.p, *hr = (uint8_t *)ctx->half_r.p;
    uint8_t *sl = (uint8_t *)ctx->sob_l.p, *sr = (uint8_t *)ctx->sob_r.p;
    uint64_t *cl1 = (uint64_t *)ctx->cl1.p, *cr1 = (uint64_t *)ctx->cr1.p;
    uint64_t *cl2 = (uint64_t *)ctx->cl.p, *cr2 = (uint64_t *)ctx->cr.p;
    uint8_t *C1 = (uint8_t *)ctx->cost1.p, *C2 = (uint8_t *)ctx->cost2.p;
    uint8_t *V1 = (uint8_t *)ctx->vol1.p, *V2 = (uint8_t *)ctx->vol.p;
    uint8_t *d1 = disp1 ? disp1 : (uint8_t *)ctx->disp1.p;
    int dirs[8][3];
    memcpy(dirs, kDirsMs, sizeof dirs);
    {
        PhaseScope ps(ctx, s, PH_CENSUS);
        const dim3 gh((hw + 63) / 64, (hh + 3) / 4), gf((W + 63) / 64, (H + 3) / 4);
        hipLaunchKernelGGL(k_half, gh, dim3(256), 0, s, l, W, H, hl);
        hipLaunchKernelGGL(k_half, gh, dim3(256), 0, s, r, W, H, hr);
        hipLaunchKernelGGL(k_sobel, dim3((hw + 63) / 64, (hh + 3) / 4), dim3(256), 0, s, hl, hw, hh, sl);
        hipLaunchKernelGGL(k_sobel, dim3((hw + 63) / 64, (hh + 3) / 4), dim3(256), 0, s, hr, hw, hh, sr);
        CK(hipGetLastError());
        if (launch_census(s, sl, cl1, hw, hh, hw) || launch_census(s, sr, cr1, hw, hh, hw))
            return -1;
        hipLaunchKernelGGL(k_sobel, gf, dim3(256), 0, s, l, W, H, sl);
        hipLaunchKernelGGL(k_sobel, gf, dim3(256), 0, s, r, W, H, sr);
        CK(hipGetLastError());
        if (launch_census(s, sl, cl2, W, H, W) || launch_census(s, sr, cr2, W, H, W))
            return -1;
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
        a.cost = C1;
        a.W = hw;
        a.H = hh;
        if (launch_paths(s, a, dirs, 8, false, true, V1, Nh * D, 0))
            return -1;
        a.cost = C2;
        a.W = W;
        a.H = H;
        if (launch_paths(s, a, dirs, 8, false, true, V2, Nh * D, 1))
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
    sgmh_ctx *c = new sgmh_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        set_err("sgmh_ctx_create: hipStreamCreateWithFlags failed");
        delete c;
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
                      &c->half_l, &c->half_r, &c->sob_l, &c->sob_r, &c->cl1, &c->cr1,
                      &c->cost1, &c->cost2, &c->vol1};
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
    if (ensure(c->cl, (N + kLead) * 8) || ensure(c->cr, (N + kLead) * 8))
        return -1;
    uint64_t *pl = (uint64_t *)c->cl.p, *pr = (uint64_t *)c->cr.p;
    CK(hipMemcpyAsync(pl + kLead, cl, N * 8, hipMemcpyDeviceToDevice, s));
    CK(hipMemcpyAsync(pr + kLead, cr, N * 8, hipMemcpyDeviceToDevice, s));
    const int dirs[1][3] = {{dx, dy, (int)kPlainP1}};
    return plain_paths(c, s, pl, pr, W, H, D, dirs, 1, L, N * D);
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
