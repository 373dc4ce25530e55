// sgm_device.h -- device-side pieces of the SGM pipeline for gfx950: census,
// the multiscale pre-filters and cost volume, the path-aggregation scanline
// groups and the sum + argmin, as inline device functions.  The single-pair
// kernels (sgm_hip.hip) and the persistent batch kernels (sgm_batch.hip) are
// thin loops over these.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <algorithm>
#include <type_traits>

#include "sgm_kernels.h"

namespace sgmk {

// ---- parameters fixed by the reference ----
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
// census: a TX x TY output tile per 256-thread block, each thread 4 consecutive x
// (TX * TY = 1024; 64x16 by default, 32x32 the batch's other tile shape)
// ---------------------------------------------------------------------------
template <int TX, int TY>
struct CenTile {
    static_assert(TX * TY == 1024 && TX % 4 == 0, "a tile is 256 threads x 4 pixels");
    static constexpr int RowBytes = TX + 8; // x0-4 .. x0+TX+3
    static constexpr int Rows = TY + 6;     // y0-3 .. y0+TY+2
    static constexpr int RowWords = RowBytes / 2; // the tile holds each pixel as a uint16
    static constexpr int Words = Rows * RowWords;
};
constexpr int kCenTX = 64, kCenTY = 16;
constexpr int kCenRowBytes = CenTile<64, 16>::RowBytes;
constexpr int kCenRows = CenTile<64, 16>::Rows;
constexpr int kCenRowWords = CenTile<64, 16>::RowWords;
// LDS words a census tile of either shape needs
constexpr int kCenTileWords = CenTile<64, 16>::Words > CenTile<32, 32>::Words ? CenTile<64, 16>::Words
                                                                              : CenTile<32, 32>::Words;

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

typedef uint16_t cen_u16x2 __attribute__((ext_vector_type(2)));

// (a - b) on the two uint16 halves independently (v_pk_sub_u16)
__device__ __forceinline__ uint32_t pk_sub_u16(uint32_t a, uint32_t b)
{
    cen_u16x2 x, y;
    __builtin_memcpy(&x, &a, 4);
    __builtin_memcpy(&y, &b, 4);
    const cen_u16x2 d = x - y;
    uint32_t r;
    __builtin_memcpy(&r, &d, 4);
    return r;
}

// (a & m) | b in one v_and_or_b32: written out, the compiler re-associates a
// chain of them into ands and three-input ors, which costs a third more
__device__ __forceinline__ uint32_t and_or(uint32_t a, uint32_t m, uint32_t b)
{
    uint32_t r;
    asm("v_and_or_b32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(m), "v"(b));
    return r;
}

// One TX x TY output tile by 256 threads (all of them must call it): the
// (TX+8)x(TY+6) input window is staged in `tile` (CenTile::Words words of
// LDS, 8-byte aligned, one uint16 per pixel) with clamped coordinates, then
// each thread builds the codes of 4 consecutive pixels from 7 rows x 6 dwords
// of LDS.
//
// Two pixels are compared at once: with a neighbour pair n and the centre
// pair c as uint16 halves, (n - c) per half is n - c < 256 when n >= c and
// 0xFF01..0xFFFF when n < c, so bits 8..15 of each half are eight copies of
// the census bit (nb < centre, common/sgm_params.h:45-47).  Bit n of a code
// is taken into bit 8 + n % 8 of word n / 8 by one v_and_or_b32, so word k
// ends up holding code byte k of the pair's first pixel in its byte 1 and of
// the second pixel in its byte 3; byte permutes then assemble the codes.
template <int TX, int TY>
__device__ __forceinline__ void census_from_tile(uint64_t *__restrict__ out, int W, int H, int ostride, int x0,
                                                 int y0, const uint32_t *__restrict__ tile);

template <int TX = 64, int TY = 16>
__device__ __forceinline__ void census_tile(const uint8_t *__restrict__ img, uint64_t *__restrict__ out,
                                            int W, int H, int ostride, int x0, int y0,
                                            uint32_t *__restrict__ tile)
{
    uint16_t *t16 = reinterpret_cast<uint16_t *>(tile);
    // every staging load is issued before the first LDS write, so a thread
    // waits for one memory latency rather than one per load.  A tile whose
    // columns x0-4 .. x0+67 all lie inside a row, of a width that is a multiple
    // of 4 (and an image base aligned to 4 bytes) is staged by aligned 4-byte
    // words; the others by clamped bytes.
    using T = CenTile<TX, TY>;
    constexpr int kCenRows = T::Rows, kCenRowBytes = T::RowBytes;
    if (x0 >= 4 && x0 + TX + 4 <= W && (W & 3) == 0 && (reinterpret_cast<uintptr_t>(img) & 3) == 0) {
        constexpr int kWords = kCenRows * (kCenRowBytes / 4), kIters = (kWords + 255) / 256;
        uint32_t wv[kIters];
#pragma unroll
        for (int k = 0; k < kIters; k++) {
            const int i = (int)threadIdx.x + 256 * k;
            const int r = i / (kCenRowBytes / 4), q = i - r * (kCenRowBytes / 4);
            const int yy = clampi(y0 - 3 + r, 0, H - 1);
            if (k + 1 < kIters || i < kWords)
                wv[k] = *reinterpret_cast<const uint32_t *>(img + (size_t)yy * W + (x0 - 4 + 4 * q));
        }
        __syncthreads(); // the previous tile of a persistent loop is no longer read
#pragma unroll
        for (int k = 0; k < kIters; k++) {
            const int i = (int)threadIdx.x + 256 * k;
            if (k + 1 < kIters || i < kWords) {
                // 4 bytes -> 4 uint16: pixels 4i .. 4i+3 of the tile
                const uint2 v = {__builtin_amdgcn_perm(0u, wv[k], 0x0C010C00u),
                                 __builtin_amdgcn_perm(0u, wv[k], 0x0C030C02u)};
                reinterpret_cast<uint2 *>(tile)[i] = v;
            }
        }
    } else {
        constexpr int kBytes = kCenRows * kCenRowBytes, kIters = (kBytes + 255) / 256;
        uint32_t px[kIters];
#pragma unroll
        for (int k = 0; k < kIters; k++) {
            const int i = (int)threadIdx.x + 256 * k;
            const int r = i / kCenRowBytes, c = i - r * kCenRowBytes;
            const int yy = clampi(y0 - 3 + r, 0, H - 1), xx = clampi(x0 - 4 + c, 0, W - 1);
            if (k + 1 < kIters || i < kBytes)
                px[k] = img[(size_t)yy * W + xx];
        }
        __syncthreads(); // the previous tile of a persistent loop is no longer read
#pragma unroll
        for (int k = 0; k < kIters; k++) {
            const int i = (int)threadIdx.x + 256 * k;
            if (k + 1 < kIters || i < kBytes)
                t16[i] = (uint16_t)px[k];
        }
    }
    __syncthreads();
    census_from_tile<TX, TY>(out, W, H, ostride, x0, y0, tile);
}

// The codes of a TX x TY output tile from its staged window (CenTile layout,
// one uint16 per pixel), 4 consecutive pixels per thread.  Every thread of the
// block calls it after the staging barrier.
template <int TX, int TY>
__device__ __forceinline__ void census_from_tile(uint64_t *__restrict__ out, int W, int H, int ostride, int x0,
                                                 int y0, const uint32_t *__restrict__ tile)
{
    using T = CenTile<TX, TY>;
    constexpr int kCenRowWords = T::RowWords;
    const int tx = threadIdx.x & (TX / 4 - 1), ty = threadIdx.x / (TX / 4);
    const int x = x0 + 4 * tx, y = y0 + ty;
    if (y >= H || x >= W)
        return;
    // pixels x-4 .. x+7 of the 7 tile rows around this thread's row, as
    // uint16 pairs: win[r][k] holds pixels x-4+2k (low half) and x-3+2k
    uint32_t win[7][6];
    const uint2 *t2 = reinterpret_cast<const uint2 *>(tile);
#pragma unroll
    for (int r = 0; r < 7; r++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const uint2 v = t2[(ty + r) * (kCenRowWords / 2) + tx + k];
            win[r][2 * k] = v.x;
            win[r][2 * k + 1] = v.y;
        }
    // the pair of pixels x-4+o, x-3+o of row r
    auto pair = [&](int r, int o) -> uint32_t {
        return (o & 1) ? __builtin_amdgcn_alignbit(win[r][(o + 1) >> 1], win[r][(o - 1) >> 1], 16)
                       : win[r][o >> 1];
    };
    uint64_t code[4];
#pragma unroll
    for (int j = 0; j < 4; j += 2) {
        const uint32_t centre = pair(3, j + 4);
        uint32_t wd[8];
        int n = 0; // bit n: n-th neighbour in (dy outer -3..3, dx inner -4..4) order, centre skipped
#pragma unroll
        for (int r = 0; r < 7; r++)
#pragma unroll
            for (int c = 0; c < 9; c++) {
                if (r == 3 && c == 4)
                    continue;
                const uint32_t f = pk_sub_u16(pair(r, j + c), centre);
                const int k = n >> 3, m = n & 7;
                const uint32_t mask = 0x01000100u << m;
                if (m == 0)
                    wd[k] = f & mask;
                else
                    wd[k] = and_or(f, mask, wd[k]);
                n++;
            }
        // byte 1 of each word -> the first pixel's code, byte 3 -> the second's
        const uint32_t t01 = __builtin_amdgcn_perm(wd[1], wd[0], 0x07030501u);
        const uint32_t t23 = __builtin_amdgcn_perm(wd[3], wd[2], 0x07030501u);
        const uint32_t t45 = __builtin_amdgcn_perm(wd[5], wd[4], 0x07030501u);
        const uint32_t t67 = __builtin_amdgcn_perm(wd[7], wd[6], 0x07030501u);
        code[j] = ((uint64_t)__builtin_amdgcn_perm(t67, t45, 0x05040100u) << 32) |
                  __builtin_amdgcn_perm(t23, t01, 0x05040100u);
        code[j + 1] = ((uint64_t)__builtin_amdgcn_perm(t67, t45, 0x07060302u) << 32) |
                      __builtin_amdgcn_perm(t23, t01, 0x07060302u);
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
__device__ __forceinline__ void half_px(const uint8_t *__restrict__ img, int W, uint8_t *__restrict__ out,
                                        int w, int x, int y)
{
    const uint8_t *a = img + (size_t)(2 * y) * W + 2 * x;
    out[(size_t)y * w + x] = (uint8_t)((a[0] + a[1] + a[W] + a[W + 1] + 2) >> 2);
}

// The multiscale census input staged straight into a 64x16 census tile:
// Sobel-X (multiscale/sgm_ms_ref.c:47-54) of the image, or of its 2x2 box
// mean (HALF, sgm_ms_ref.c:145-150), at the census window's clamped
// coordinates, so neither the half-scale image nor the Sobel image is
// written.  W, H: the input image; the codes are of the (W/2)x(H/2) half-scale
// image when HALF.  win: LDS for the scaled image over the window's Sobel
// neighbourhood, (64+10) x (16+8) bytes.  Then the codes as census_tile makes
// them.  Every thread of the block calls it.
constexpr int kSobWinX = kCenTX + 10, kSobWinY = kCenTY + 8;
template <bool HALF>
__device__ __forceinline__ void sobel_census_tile(const uint8_t *__restrict__ img, int W, int H,
                                                  uint64_t *__restrict__ out, int x0, int y0,
                                                  uint32_t *__restrict__ tile, uint8_t *__restrict__ win)
{
    const int w = HALF ? W / 2 : W, h = HALF ? H / 2 : H;
    constexpr int kWin = kSobWinX * kSobWinY, kIters = (kWin + 255) / 256;
    // the scaled image over [x0-5, x0+69] x [y0-4, y0+19], clamped: every load
    // is issued before the first LDS write
    uint32_t v[kIters];
#pragma unroll
    for (int k = 0; k < kIters; k++) {
        const int i = (int)threadIdx.x + 256 * k;
        const int r = i / kSobWinX, c = i - r * kSobWinX;
        const int X = clampi(x0 - 5 + c, 0, w - 1), Y = clampi(y0 - 4 + r, 0, h - 1);
        if (k + 1 < kIters || i < kWin) {
            if constexpr (HALF) {
                const uint8_t *a = img + (size_t)(2 * Y) * W + 2 * X;
                v[k] = ((uint32_t)a[0] + a[1] + a[W] + a[W + 1] + 2) >> 2;
            } else {
                v[k] = img[(size_t)Y * W + X];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < kIters; k++) {
        const int i = (int)threadIdx.x + 256 * k;
        if (k + 1 < kIters || i < kWin)
            win[i] = (uint8_t)v[k];
    }
    __syncthreads();
    // Sobel-X at each census window pixel's clamped (cx, cy), from the scaled
    // pixels at clamped cx +- 1, cy +- 1 (all inside win)
    uint16_t *t16 = reinterpret_cast<uint16_t *>(tile);
    constexpr int kTile = kCenRows * kCenRowBytes;
    for (int i = (int)threadIdx.x; i < kTile; i += 256) {
        const int r = i / kCenRowBytes, c = i - r * kCenRowBytes;
        const int cx = clampi(x0 - 4 + c, 0, w - 1), cy = clampi(y0 - 3 + r, 0, h - 1);
        const int xl = clampi(cx - 1, 0, w - 1) - (x0 - 5), xr = clampi(cx + 1, 0, w - 1) - (x0 - 5);
        const int yu = (clampi(cy - 1, 0, h - 1) - (y0 - 4)) * kSobWinX;
        const int ym = (cy - (y0 - 4)) * kSobWinX;
        const int yd = (clampi(cy + 1, 0, h - 1) - (y0 - 4)) * kSobWinX;
        const int gx = ((int)win[yu + xr] - win[yu + xl]) + 2 * ((int)win[ym + xr] - win[ym + xl]) +
                       ((int)win[yd + xr] - win[yd + xl]);
        t16[i] = (uint16_t)clampi(128 + (gx >> 3), 0, 255);
    }
    __syncthreads();
    census_from_tile<kCenTX, kCenTY>(out, w, h, w, x0, y0, tile);
}

// Sobel-X, out = clamp(128 + (gx >> 3)) with arithmetic shift,
// multiscale/sgm_ms_ref.c:47-54
__device__ __forceinline__ void sobel_px(const uint8_t *__restrict__ img, int W, int H,
                                         uint8_t *__restrict__ out, int x, int y)
{
    const int xl = clampi(x - 1, 0, W - 1), xr = clampi(x + 1, 0, W - 1);
    const uint8_t *up = img + (size_t)clampi(y - 1, 0, H - 1) * W;
    const uint8_t *md = img + (size_t)y * W;
    const uint8_t *dn = img + (size_t)clampi(y + 1, 0, H - 1) * W;
    const int gx = (up[xr] - up[xl]) + 2 * (md[xr] - md[xl]) + (dn[xr] - dn[xl]);
    out[(size_t)y * W + x] = (uint8_t)clampi(128 + (gx >> 3), 0, 255);
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

// Materialised cost volume (multiscale): pixel (x, y), one pixel per 16-lane
// row, lane l owning disparities l*DPL .. l*DPL+DPL-1.
//   merge == nullptr: C = popcount(cl ^ cr[x-d]) or `invalid`      (sgm_ms_ref.c:66-72)
//   merge != nullptr: C = (C + merge[(y/2)*hw + x/2][d/2] + 1) >> 1 (sgm_ms_ref.c:184-188)
// The lane's right codes cr[x-d] are contiguous and descending in d: they
// come 16 bytes at a time through a buffer resource (on rows below the first,
// codes left of x = 0 are the previous row's and are replaced by `invalid`).
// On row 0 a window starting left of x = 0 is read code by code with clamped
// indices: a 16-byte load straddling offset 0 would be out of range as a
// whole.  The merge terms of a lane are DPL/2 consecutive bytes, one load.
template <int DPL>
struct CostIn {
    uint64_t a;       // left code
    uint64_t b[DPL];  // b[j] = cr[row + x - l*DPL - j]
    uint64_t mw;      // merge bytes (merge != nullptr)

    __device__ __forceinline__ void load(const uint64_t *__restrict__ cl, const uint64_t *__restrict__ cr,
                                         rsrc_t rr, int W, int D, const uint8_t *__restrict__ merge, int x,
                                         int y, int l)
    {
        const uint32_t p = (uint32_t)(y * W + x);
        a = cl[p];
        const int lo = x - l * DPL - (DPL - 1); // lowest code of the window, relative to the row start
        const uint32_t base = (uint32_t)(y * W + lo) * 8u;
        if (DPL >= 2 && (y > 0 || lo >= 0)) {
#pragma unroll
            for (int i = 0; i < DPL; i += 2) {
                const auto q = __builtin_amdgcn_raw_buffer_load_b128(rr, base + 8u * i, 0, 0);
                b[DPL - 1 - i] = ((uint64_t)q[1] << 32) | q[0];
                b[DPL - 2 - i] = ((uint64_t)q[3] << 32) | q[2];
            }
        } else {
            // row 0, window starting left of x = 0: clamped in-range loads
            // (codes left of x = 0 are replaced by `invalid`)
            const uint64_t *crow = cr + (size_t)y * W;
#pragma unroll
            for (int j = 0; j < DPL; j++) {
                const int xr = x - (l * DPL + j);
                b[j] = crow[xr < 0 ? 0 : xr];
            }
        }
        mw = 0;
        if (merge) {
            const uint8_t *m = merge + ((size_t)(y / 2) * (W / 2) + x / 2) * D;
            if constexpr (DPL == 1)
                mw = m[l >> 1];
            else if constexpr (DPL == 2)
                mw = m[l];
            else if constexpr (DPL == 4)
                mw = *reinterpret_cast<const uint16_t *>(m + 2 * l);
            else if constexpr (DPL == 8)
                mw = *reinterpret_cast<const uint32_t *>(m + 4 * l);
            else
                mw = *reinterpret_cast<const uint64_t *>(m + 8 * l);
        }
    }

    __device__ __forceinline__ void finish(int W, int D, uint32_t invalid, bool merged, uint8_t *__restrict__ C,
                                           int x, int y, int l) const
    {
        uint32_t c[DPL];
#pragma unroll
        for (int j = 0; j < DPL; j++) {
            const uint32_t h = (uint32_t)__popcll(a ^ b[j]);
            c[j] = x - (l * DPL + j) < 0 ? invalid : h;
            if (merged)
                c[j] = (c[j] + (uint32_t)((mw >> (8 * (j >> 1))) & 0xFFu) + 1) >> 1;
        }
        store_bytes<DPL>(C + (size_t)(y * W + x) * D + l * DPL, c);
    }
};

template <int DPL>
__device__ __forceinline__ void cost_px(const uint64_t *__restrict__ cl, const uint64_t *__restrict__ cr,
                                        int W, int H, int D, uint32_t invalid, const uint8_t *__restrict__ merge,
                                        uint8_t *__restrict__ C, int x, int y, int l)
{
    const rsrc_t rr = make_rsrc(cr, (uint32_t)(W * H) * 8u);
    CostIn<DPL> in;
    in.load(cl, cr, rr, W, D, merge, x, y, l);
    in.finish(W, D, invalid, merge != nullptr, C, x, y, l);
}

// Store DPL packed bytes at byte offset `off` (kDrop: no store).  NT: the
// stores are non-temporal (cache policy NT).  The path kernels use it from 8
// disparities per lane up: those volumes (4 x N x D bytes, about 1 GB at 1080p
// D=128) are written once, read once by the argmin and outgrow the last-level
// cache, so they bypass it and leave it to the census codes every direction
// re-reads (plain 1080p D=128 0.715 -> 0.657 ms).  Narrower lanes store 4
// bytes or fewer per lane and step, which NT writes reach DRAM as partial
// lines (21 % more write traffic at D=64, for no gain in time).  The windowed
// path stores its full-layout tiles NT at D=128 too, so that its small dense
// narrow volumes stay cached for the argmin.
constexpr int kStoreNT = 2; // buffer-store cache policy: non-temporal
template <int DPL, bool NT = true>
__device__ __forceinline__ void st_bytes(rsrc_t r, uint32_t off, const uint32_t (&v)[DPL])
{
    constexpr int pol = NT ? kStoreNT : 0;
    if constexpr (DPL == 1) {
        __builtin_amdgcn_raw_buffer_store_b8((uint8_t)v[0], r, off, 0, pol);
    } else if constexpr (DPL == 2) {
        __builtin_amdgcn_raw_buffer_store_b16((uint16_t)(v[0] | (v[1] << 8)), r, off, 0, pol);
    } else {
        uint32_t w[DPL / 4];
#pragma unroll
        for (int k = 0; k < DPL / 4; k++)
            w[k] = pack4(v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]);
        if constexpr (DPL == 4) {
            __builtin_amdgcn_raw_buffer_store_b32(w[0], r, off, 0, pol);
        } else if constexpr (DPL == 8) {
            typedef uint32_t v2u __attribute__((ext_vector_type(2)));
            const v2u q = {w[0], w[1]};
            __builtin_amdgcn_raw_buffer_store_b64(q, r, off, 0, pol);
        } else {
#pragma unroll
            for (int k = 0; k < DPL / 4; k += 4) {
                typedef uint32_t v4u __attribute__((ext_vector_type(4)));
                const v4u q = {w[k], w[k + 1], w[k + 2], w[k + 3]};
                __builtin_amdgcn_raw_buffer_store_b128(q, r, off == kDrop ? kDrop : off + 4 * k, 0, pol);
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
    int nlines;      // scanlines in this direction (dec: the surviving ones, dec_lines)
    int first_block; // first block of this direction in the grid
    int dec;         // 1: the decimated multiscale stage, which aggregates only dec_line's lines
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
__host__ __device__ __forceinline__ LineGeom line_geom(int i, int dx, int dy, int W, int H)
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

// The decimated multiscale stage keeps L only at even (x, y)
// (multiscale/sgm_ms_ref.c:194-199), and a scanline's L depends on the earlier
// pixels of that line alone.  So only the lines through even-even pixels are
// aggregated: the even rows, the even columns, and the diagonals whose x + y
// is even (a diagonal step changes x + y by 0 or 2).  The other half of every
// direction's lines store nothing and feed nothing.  W and H are even there.
// dec_lines: how many lines survive; dec_line: the line_geom index of
// surviving line m.  A diagonal's first W lines start on the edge row at x = i
// or W-1-i, the next H-1 on the edge column at row j = i-W+1 or H-1-j; with
// a = (dx < 0) + (dy < 0) mod 2, x + y is even exactly for i = a mod 2 on the
// edge row and j = a mod 2 on the edge column.
__host__ __device__ __forceinline__ int dec_lines(int dx, int dy, int W, int H)
{
    if (dy == 0)
        return H / 2;
    if (dx == 0)
        return W / 2;
    const int a = ((dx < 0) + (dy < 0)) & 1;
    return W / 2 + H / 2 - 1 + a;
}
__host__ __device__ __forceinline__ int dec_line(int m, int dx, int dy, int W)
{
    if (dy == 0 || dx == 0)
        return 2 * m;
    const int a = ((dx < 0) + (dy < 0)) & 1;
    return m < W / 2 ? 2 * m + a : W - 1 + 2 * (m - W / 2) + 2 - a;
}

// 64-bit row shift by one lane; the lane without a source (lane 0 for shr,
// 15 for shl) reads 0 and takes `in` instead, which is 0 in every other lane:
// the OR folds into the DPP move (v_or_b32_dpp, bound_ctrl:1), one op per half.
// (LPL = 64: the same across the whole wave, lanes 0 / 63 without a source)
template <int LPL = kRowLanes>
__device__ __forceinline__ uint64_t dpp64_shr1(uint64_t in, uint64_t v)
{
    constexpr int ctl = LineLanes<LPL>::shr;
    const uint32_t lo = dpp<ctl>((uint32_t)v) | (uint32_t)in;
    const uint32_t hi = dpp<ctl>((uint32_t)(v >> 32)) | (uint32_t)(in >> 32);
    return ((uint64_t)hi << 32) | lo;
}
template <int LPL = kRowLanes>
__device__ __forceinline__ uint64_t dpp64_shl1(uint64_t in, uint64_t v)
{
    constexpr int ctl = LineLanes<LPL>::shl;
    const uint32_t lo = dpp<ctl>((uint32_t)v) | (uint32_t)in;
    const uint32_t hi = dpp<ctl>((uint32_t)(v >> 32)) | (uint32_t)(in >> 32);
    return ((uint64_t)hi << 32) | lo;
}

// An offset bias that keeps a lane's loads out of range (they return 0) for a
// whole line: buffers stay below 2^31 bytes and a line advances the offset by
// less than 2^31, so biased offsets never wrap back into range.
constexpr uint32_t kFar = 0x80000000u;

// Horizontal direction DX (+1: L->R, -1: R->L), cost from census codes.
// Lane l keeps the window w[j] = cr[y][x - l*DPL - j] in registers; one step
// along the row shifts it by one disparity: the code leaving lane l's slice
// enters lane l+1's (L->R) or lane l-1's (R->L) through one row_shr:1 /
// row_shl:1 DPP move, and the end lane takes the one new code of the step.
// The window is a register ring indexed by the step modulo DPL (PF is a
// multiple of DPL), so the shift itself moves nothing.
// The path starts from Lp = 0, mn = 0: with P1 > 0 that step yields L = C
// exactly, the reference's first-pixel rule (sgm_ref.c:90-91), branch-free.
// BCAST: lane 0 of each row loads the left code and the XORs take it through
// a row_newbcast DPP operand.  Same instruction count, but it changes how the
// row steps issue beside the column waves that share their SIMDs: measured
// 3-4 % faster path kernels at D = 64 where there is at least one row wave
// per CU (1280x720, 1920x1080), 6 % slower where rows are few (640x480) and
// 1 % slower at D = 128, so the launch picks it by frame height and D.
// LPL: lanes per scanline (16, or 64 for the wave-line layout, where lane l
// owns disparities l*DPL .. with DPL = D/64 and the shift crosses the wave).
template <int DPL, bool SAT, int DX, bool BCAST, class A, int LPL = kRowLanes>
__device__ __forceinline__ void path_horizontal(const A &a, const DirTask &t, LineGeom ln,
                                                int l, int maxsteps)
{
    static_assert(!BCAST || LPL == kRowLanes, "the left-code broadcast is per DPP row");
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
    // prefetched cl[x_s] and the window code entering at step s; only the
    // lane that takes the entering code loads it, the others read 0
    const int enter = DX > 0 ? 0 : -(D - 1);
    const uint32_t nbias = l == (DX > 0 ? 0 : LPL - 1) ? 0u : kFar;
    const uint32_t lbias = BCAST && l != 0 ? kFar : 0u; // BCAST: lane 0 alone loads the left code
    uint64_t pl[PF], pn[PF];
#pragma unroll
    for (int k = 0; k < PF; k++) {
        pl[k] = ld64(rl, (uint32_t)(row + x + k * DX) * 8u + lbias);
        pn[k] = ld64(rr, (uint32_t)(row + x + k * DX + enter) * 8u + nbias);
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
    // The costs of step k of a block, at x = xk: shift the window to that
    // step, then the Hamming costs (CHECK: with the invalid select).
    auto costs = [&](int k, int xk, auto check_tag, uint32_t (&C)[DPL]) {
        constexpr bool CHECK = decltype(check_tag)::value;
        if constexpr (DX > 0) {
            const int slot = ((-k) % DPL + DPL) % DPL;        // logical 0 at step k
            phys[slot] = dpp64_shr1<LPL>(pn[k], phys[slot]); // was logical DPL-1
        } else {
            const int slot = ((k - 1) % DPL + DPL) % DPL;     // logical DPL-1 at step k
            phys[slot] = dpp64_shl1<LPL>(pn[k], phys[slot]); // was logical 0
        }
        const uint64_t cL = BCAST ? row_bcast0(pl[k]) : pl[k];
        const int rel = xk - lbase;
#pragma unroll
        for (int j = 0; j < DPL; j++) {
            const int slot = DX > 0 ? (((j - k) % DPL) + DPL) % DPL : (j + k) % DPL;
            const uint32_t h = (uint32_t)__popcll(cL ^ phys[slot]);
            C[j] = CHECK ? (rel >= j ? h : inval) : h;
        }
    };
    // Forming step k+1's costs in step k's scheduling region lets the
    // scheduler fill the minimum's DPP wait states with them (no s_nop left in
    // the D = 64 row step), and a row direction alone gets 5-8 % faster.  Beside
    // the column waves it was not faster: measured on one box against this
    // form, 1080p D = 64 frames 0.352 -> 0.361 ms one at a time and 0.327 ->
    // 0.323 two in flight; D = 128 0.645 -> 0.641 and 0.552 -> 0.557.  The
    // freed issue slots go to the row, which then takes more of the SIMD it
    // shares with a column wave.  Each step forms its costs at its start.
    auto block = [&](int s0, auto check_tag) {
        constexpr bool CHECK = decltype(check_tag)::value;
#pragma unroll
        for (int k = 0; k < PF; k++) {
            const int s = s0 + k;
            uint32_t C[DPL];
            costs(k, x, check_tag, C);
            // ring slot k is dead now: refill it for step s + PF in place
            __builtin_amdgcn_sched_barrier(0);
            const int xp = x + PF * DX;
            pl[k] = ld64(rl, (uint32_t)(row + xp) * 8u + lbias);
            pn[k] = ld64(rr, (uint32_t)(row + xp + enter) * 8u + nbias);
            sgm_step<DPL, SAT, LPL>(L, C, mn, P1, P2);
            st_bytes<DPL, (DPL >= 8)>(ro, (!CHECK || s < steps) ? off : kDrop, L);
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
// VERT: 0 any direction (cost select per step), 1 a column (x fixed: the
// invalid disparities are fixed too, folded in by a max), 2 columns with no
// invalid disparity at all (x >= D-1 on every line of the wave).
// LPL: lanes per scanline, as for path_horizontal (64: census-fed only).
template <int DPL, bool FROM_CENSUS, bool SAT, bool DECIM, int VERT, class A, int LPL = kRowLanes>
__device__ __forceinline__ void path_generic(const A &a, const DirTask &t, LineGeom ln, int l,
                                             int maxsteps)
{
    static_assert(LPL == kRowLanes || FROM_CENSUS, "the wave-line layout reads census codes");
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
    // PK: the multiscale step on uint16 pairs at D = 64 and 128 (sgm_step_sat_pk)
    constexpr bool PK = !FROM_CENSUS && SAT && (DPL == 4 || DPL == 8);
    constexpr int NW = DPL >= 2 ? DPL / 2 : 1; // pair words (used when PK)
    uint32_t L[DPL], mn = 0, Lw[NW];
#pragma unroll
    for (int w = 0; w < NW; w++)
        Lw[w] = 0;
#pragma unroll
    for (int j = 0; j < DPL; j++)
        L[j] = 0;

    // census-fed costs of ring slot k, at x = xk (a generic lambda: formed only
    // where it is called, in the census-fed instantiations)
    auto costs = [&](int k, int xk, auto &C) {
        const uint64_t cL = rq[k][DPL];
        const int rel = xk - lbase;
#pragma unroll
        for (int j = 0; j < DPL; j++) {
            const uint32_t h = (uint32_t)__popcll(cL ^ rq[k][j]);
            if constexpr (VERT == 2)
                C[j] = h;
            else if constexpr (VERT == 1)
                C[j] = h > inv[j] ? h : inv[j];
            else
                C[j] = rel >= j ? h : inval;
        }
    };
    // One block of PF steps; CHECK: a step of the block may lie past the end
    // of its line (always possible on diagonals, whose lines differ in length).
    auto block = [&](int s0, auto check_tag) {
        constexpr bool CHECK = decltype(check_tag)::value;
#pragma unroll
        for (int k = 0; k < PF; k++) {
            const int s = s0 + k;
            uint32_t C[DPL];
            if constexpr (FROM_CENSUS) {
                costs(k, x, C);
            } else if constexpr (!PK) {
#pragma unroll
                for (int j = 0; j < DPL; j++)
                    C[j] = (rw[k].w[j / 4] >> (8 * (j % 4))) & 0xFFu;
            }
            uint32_t Cw[NW];
            if constexpr (PK) { // the cost bytes as uint16 pairs
#pragma unroll
                for (int q = 0; q < DPL / 4; q++) {
                    Cw[2 * q] = __builtin_amdgcn_perm(0u, rw[k].w[q], 0x0C010C00u);
                    Cw[2 * q + 1] = __builtin_amdgcn_perm(0u, rw[k].w[q], 0x0C030C02u);
                }
            }
            // ring slot k is dead now: refill it for step s + PF in place
            __builtin_amdgcn_sched_barrier(0);
            fetch(pix + PF * pstep, k);
            if constexpr (PK)
                sgm_step_sat_pk<NW>(Lw, Cw, mn, P1, P2);
            else
                sgm_step<DPL, SAT, LPL>(L, C, mn, P1, P2);
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
            if constexpr (PK && DPL == 4) { // bytes 0 and 2 of each pair word
                __builtin_amdgcn_raw_buffer_store_b32(__builtin_amdgcn_perm(Lw[1], Lw[0], 0x06040200u), ro,
                                                      off | dead, 0, 0);
            } else if constexpr (PK) { // D = 128: 8 bytes, non-temporal like st_bytes at 8 per lane
                typedef uint32_t v2u __attribute__((ext_vector_type(2)));
                const v2u q = {__builtin_amdgcn_perm(Lw[1], Lw[0], 0x06040200u),
                               __builtin_amdgcn_perm(Lw[3], Lw[2], 0x06040200u)};
                __builtin_amdgcn_raw_buffer_store_b64(q, ro, off | dead, 0, kStoreNT);
            } else
                st_bytes<DPL, (DPL >= 8)>(ro, off | dead, L);
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

// Scanlines grp*4 .. grp*4+3 of direction t, one per 16-lane row of the wave
// (LPL = 16), or scanline grp alone across the wave (LPL = 64).
// A is any struct with PathArgs' frame fields (cl, cr, cost, W, H, D, P2,
// invalid, out_bytes).
template <int DPL, bool FROM_CENSUS, bool SAT, bool DECIM, class A, bool BCAST = false, int LPL = kRowLanes>
__device__ __forceinline__ void path_group(const A &a, const DirTask &t, int grp)
{
    constexpr int LPW = 64 / LPL; // scanlines per wave
    const int lane = threadIdx.x & 63, g = lane / LPL, l = lane % LPL;
    const int base = grp * LPW;
    // line_geom's index of this direction's line m
    auto index = [&](int m) { return t.dec ? dec_line(m, t.dx, t.dy, a.W) : m; };
    int maxsteps = 0;
#pragma unroll
    for (int q = 0; q < LPW; q++)
        if (base + q < t.nlines)
            maxsteps = max(maxsteps, line_geom(index(base + q), t.dx, t.dy, a.W, a.H).steps);
    const int line = base + g;
    LineGeom ln = {0, 0, 0};
    if (line < t.nlines)
        ln = line_geom(index(line), t.dx, t.dy, a.W, a.H);
    if constexpr (FROM_CENSUS && !DECIM) {
        if (t.dy == 0) {
            if (t.dx > 0)
                path_horizontal<DPL, SAT, 1, BCAST, A, LPL>(a, t, ln, l, maxsteps);
            else
                path_horizontal<DPL, SAT, -1, BCAST, A, LPL>(a, t, ln, l, maxsteps);
            return;
        }
    }
    if (FROM_CENSUS && t.dx == 0) {
        // line i of a column direction is column x = i
        if (base >= a.D - 1)
            path_generic<DPL, FROM_CENSUS, SAT, DECIM, 2, A, LPL>(a, t, ln, l, maxsteps);
        else
            path_generic<DPL, FROM_CENSUS, SAT, DECIM, 1, A, LPL>(a, t, ln, l, maxsteps);
    } else {
        path_generic<DPL, FROM_CENSUS, SAT, DECIM, 0, A, LPL>(a, t, ln, l, maxsteps);
    }
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

constexpr int kWtaBatch = 4; // pixel batches per wta_batch call: 4 x 4 directions x 16 B in flight per lane

// One batch of U*PPW pixels starting at b0 (every lane of the wave calls it
// with its lane index).
template <int LPP, int NDIRS, int U = kWtaBatch>
__device__ __forceinline__ void wta_batch(const uint8_t *__restrict__ L, size_t dir_stride, int npix,
                                          uint8_t *__restrict__ disp, int b0, int lane)
{
    constexpr int D = 16 * LPP, PPW = 64 / LPP;
    const int sub = lane % LPP, slot = lane / LPP;
    uint4 v[U][NDIRS];
    int pix[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int p = b0 + u * PPW + slot;
        pix[u] = p;
        const int pc = p < npix ? p : npix - 1;
        const uint8_t *src = L + (size_t)pc * D + sub * 16;
        // each volume byte is read exactly once: stream it past the caches
        typedef uint32_t v4u __attribute__((ext_vector_type(4)));
#pragma unroll
        for (int k = 0; k < NDIRS; k++) {
            const v4u q = __builtin_nontemporal_load(reinterpret_cast<const v4u *>(src + k * dir_stride));
            v[u][k] = make_uint4(q[0], q[1], q[2], q[3]);
        }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
        uint32_t ev[4] = {0, 0, 0, 0}, od[4] = {0, 0, 0, 0};
#pragma unroll
        for (int k = 0; k < NDIRS; k++) {
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


// ---------------------------------------------------------------------------
// windowed SGM (build-defined; no reference counterpart)
// ---------------------------------------------------------------------------
// Each 16x16 tile t has an active range of 16-disparity blocks
// [blk[t] & 15, blk[t] >> 4]; the recurrence sees only the predecessor's
// active disparities (the rest are absent), mn is the minimum over the
// predecessor's active range, each candidate of m(d) saturates at 255 as in
// the reference step, and the argmin runs over the pixel's own range.  With
// every block active it is the plain path exactly.
//
// Absent disparities are held as 255.  A real L never exceeds 254 (m <= mn +
// P2, so L <= C + P2), so 255 is never a real value and never the minimum
// while any real value exists, and as a candidate Lp(d) = 255 caps m(d) at 255
// -- the reference's saturation -- with no extra instruction: the plain
// sgm_step is exact for windows as it stands.
//
// A line is cut into 16-step segments, one per tile it crosses; the 4 lines of
// a wave share every tile (groups of 4 rows or 4 columns never straddle one),
// so a segment's block range is wave-uniform.  Three segment modes:
//   NARROW  two blocks [b0, b0+1]: lane l holds d = 16*b0 + 2l + j, j = 0, 1
//           (the plain layout at DPL = 2 over a 32-disparity window)
//   FULL    every block: the plain layout at DPL = D/16
//   MASKED  other ranges (one block included): the plain layout, absent
//           disparities pinned to 255
// The state moves between layouts through a per-wave LDS row (one byte per
// line and disparity) only where a segment's mode or window differs from the
// previous one.
//
// Per-step work is what bounds these kernels (a wave issues one instruction
// per cycle slot, scalar or vector), so every test that only boundary
// segments need -- steps before or past the line, disparities with x - d < 0,
// MASKED penalties -- runs in a CHECK variant of the segment body, and the
// steady-state body has none; addresses are a per-segment base plus constant
// offsets.  Lane groups past the last line recompute the last line and store
// the same bytes to the same places.
constexpr uint32_t kWinAbsent = 255u;
constexpr int kWinMaxTiles = 1024;      // tiles along a line (W, H <= 16384)
constexpr uint32_t kWinPenalty = 1000u; // cost of an absent disparity inside a step: L > 255

// LDS one wave of the windowed kernels uses: the layout exchange row of its
// four lines and its line's tile ranges
template <int NB>
constexpr int win_wave_lds() { return kLinesPerWave * 16 * NB + kWinMaxTiles; }

struct WinArgs {
    const uint64_t *cl, *cr; // census codes, pixel i at index kLead + i
    const uint8_t *blk;      // per-tile block ranges
    int W, H, D, tiles_x;
    uint32_t out_bytes;      // bytes of each direction's full volume (N*D)
    int ndirs;
    DirTask dir[4];          // dir[k].L: full volume [N][D] (FULL / MASKED tiles)
    uint8_t *Ln[4];          // narrow volumes [N][32]: d - 16*b0 of NARROW tiles
};

struct WinSeg {
    int mode; // 0 NARROW, 1 FULL, 2 MASKED
    int b0;   // NARROW: first block
    uint32_t mask;
};

// A tile's range is NARROW iff it holds exactly two blocks (the argmin kernel
// applies the same rule to pick the narrow or the full volume).
template <int NB>
__device__ __forceinline__ WinSeg win_seg(uint32_t rng)
{
    const int b1 = min((int)(rng >> 4), NB - 1), b0 = min((int)(rng & 15u), b1); // clamped
    WinSeg w;
    w.mask = ((2u << b1) - 1u) & ~((1u << b0) - 1u);
    w.b0 = b0;
    w.mode = b1 - b0 == 1 ? 0 : (w.mask == (1u << NB) - 1u ? 1 : 2);
    return w;
}

// State of the 4 lines of a wave and its move between segment layouts
// through the wave's LDS row.  The row is written and read with different
// access widths, so every access goes through memcpy (byte semantics: no
// type-based aliasing).  Entering a NARROW window, Lp+P1 of the predecessor's
// disparities just below / above it become the first step's edge neighbours
// (its range may reach past the new window).
template <int NB>
struct WinState {
    static constexpr int D = 16 * NB;
    uint32_t SN[2], SF[NB], mn;
    uint32_t edge_lo, edge_hi;
    __device__ __forceinline__ void reset()
    {
        SN[0] = SN[1] = 0;
#pragma unroll
        for (int j = 0; j < NB; j++)
            SF[j] = 0;
        mn = 0;
        edge_lo = edge_hi = 0xFFFFFFFFu;
    }
    static __device__ __forceinline__ void put(uint8_t *dst, const uint32_t *v, int n)
    {
        uint8_t b[NB];
        for (int j = 0; j < n; j++)
            b[j] = (uint8_t)v[j];
        __builtin_memcpy(dst, b, n);
    }
    static __device__ __forceinline__ void get(const uint8_t *src, uint32_t *v, int n)
    {
        uint8_t b[NB];
        __builtin_memcpy(b, src, n);
        for (int j = 0; j < n; j++)
            v[j] = b[j];
    }
    // The edges are formed as values and assigned on every path: a store to
    // one member on one branch and to another on the other lets the optimiser
    // merge them into one store through a pointer select, which puts the
    // whole state in scratch memory (and a scratch load waits for every
    // prefetch in flight).
    __device__ __forceinline__ void exchange(uint8_t *row, int l, uint32_t P1, const WinSeg &o, const WinSeg &c)
    {
        if (o.mode == 0) {
            uint32_t ff[NB];
#pragma unroll
            for (int j = 0; j < NB; j++)
                ff[j] = kWinAbsent;
            put(row + NB * l, ff, NB);
            put(row + 16 * o.b0 + 2 * l, SN, 2);
        } else {
            put(row + NB * l, SF, NB);
        }
        uint32_t lo = 0xFFFFFFFFu, hi = 0xFFFFFFFFu;
        if (c.mode == 0) {
            get(row + 16 * c.b0 + 2 * l, SN, 2);
            const int dl = 16 * c.b0 - 1, dh = 16 * c.b0 + 32;
            lo = dl >= 0 ? (uint32_t)row[dl] + P1 : 0xFFFFFFFFu;
            hi = dh < D ? (uint32_t)row[dh] + P1 : 0xFFFFFFFFu;
        } else {
            get(row + NB * l, SF, NB);
        }
        edge_lo = lo;
        edge_hi = hi;
    }
};

template <int NB>
__device__ __forceinline__ bool win_changes(const WinSeg &o, const WinSeg &c)
{
    return (c.mode == 0) != (o.mode == 0) || (c.mode == 0 && c.b0 != o.b0);
}

// Rows: lines grp*4 .. grp*4+3 of a horizontal direction, DIR = +1 (left to
// right) or -1.  Each lane keeps its window of right codes in registers, as the
// plain row kernel does: one step along the row shifts it by one disparity
// through a row_shr:1 / row_shl:1 DPP move, and the end lane takes the one new
// code (8 bytes, the only right-code load of the step).  The window is a
// register ring indexed by the step modulo the mode's disparities per lane (2
// NARROW, NB otherwise; both divide 16).  Each segment starts from a window
// loaded whole for its mode and window (NB/2 16-byte loads issued PF steps
// ahead), so every step of every mode issues the same loads in the same
// order and the prefetch ring's counters are exact.
template <int NB, int DIR>
__device__ __forceinline__ void path_window_rows(const WinArgs &a, const DirTask &t, uint8_t *narrow, int grp,
                                                 uint8_t *wlds)
{
    constexpr int PF = 8;
    constexpr int D = 16 * NB;
    constexpr int NC = NB / 2;
    typedef uint32_t v4u __attribute__((ext_vector_type(4)));
    uint8_t *xs = wlds;                           // layout exchange, kLinesPerWave * D bytes of this wave's LDS
    __builtin_amdgcn_s_setprio(2); // rows are the frame's longest lines: its critical path
    const int lane = threadIdx.x & 63, g = lane >> 4, l = lane & 15;
    const int W = a.W, H = a.H;
    const int nt = (W + 15) >> 4;
    const int ln = min(grp * kLinesPerWave + g, t.nlines - 1);
    const int tile_y = (grp * kLinesPerWave) >> 4;
    const uint32_t nb8 = (uint32_t)(W * H + kLead) * 8u;
    const rsrc_t rl = make_rsrc(a.cl, nb8), rr = make_rsrc(a.cr, nb8);
    const rsrc_t ro = make_rsrc(t.L, a.out_bytes);
    const rsrc_t rn = make_rsrc(narrow, (uint32_t)(W * H) * 32u);
    const uint32_t P1 = (uint32_t)t.P1, P2 = kPlainP2, inval = kPlainInvalid;
    // steps q in [0, 16*nt) on a tile-aligned axis: x = q (DIR > 0) or 16*nt-1-q
    const int q0 = DIR > 0 ? 0 : 16 * nt - W, qend = DIR > 0 ? W : 16 * nt;
    const int rowc = ln * W + kLead, rowp = ln * W; // code / pixel index of (0, ln)
    // The line's tile ranges in traversal order, staged in LDS once: a segment
    // reads its range a few segments ahead from there, so no wait for a global
    // load (which would drain the code prefetch ring) falls inside the line.
    uint8_t *segs = wlds + kLinesPerWave * D; // kWinMaxTiles bytes of this wave's LDS
    for (int k = lane; k < nt; k += 64)
        segs[k] = a.blk[tile_y * a.tiles_x + (DIR > 0 ? k : nt - 1 - k)];
    // this wave alone reads what it staged, and one wave's LDS accesses run in
    // order: keeping the compiler from moving the reads above suffices
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    auto seg_raw = [&](int k) -> uint32_t { return segs[k < nt ? k : nt - 1]; }; // the same in every lane
    auto half_code = [](const v4u &v, int half) -> uint64_t {
        return half ? ((uint64_t)v[3] << 32) | v[2] : ((uint64_t)v[1] << 32) | v[0];
    };
    // Step i of a segment (up to 15 + PF for the prefetches into the next one)
    // lies at x = x0 + DIR*i; offsets are a per-segment base plus a constant,
    // the base at i = 0 (DIR > 0) or at i = 15 + PF for codes, 15 for stores (DIR < 0).
    auto cstep = [](int i) { return DIR > 0 ? i : 15 + PF - i; };
    auto sstep = [](int i) { return DIR > 0 ? i : 15 - i; };
    // disparity entering at the end lane, and the code offset of a lane's window
    auto d_enter = [](const WinSeg &sg) {
        return DIR > 0 ? (sg.mode == 0 ? 16 * sg.b0 : 0) : (sg.mode == 0 ? 16 * sg.b0 + 31 : D - 1);
    };
    auto d_lo = [&](const WinSeg &sg) { return sg.mode == 0 ? 16 * sg.b0 + 2 * l + 1 : NB * l + NB - 1; };
    const uint32_t nbias = l == (DIR > 0 ? 0 : 15) ? 0u : kFar;

    uint64_t rcl[PF], ren[PF];
    v4u nw[NC];
    uint64_t win[NB];
    // per-segment bases (bytes)
    uint32_t cb, eb_cur, eb_nxt, wb_nxt, sb_n, sb_f;
    auto bases = [&](int k, const WinSeg &c, const WinSeg &n) {
        const int x0 = DIR > 0 ? 16 * k : 16 * nt - 1 - 16 * k;
        const int cbase = rowc + x0 - (DIR > 0 ? 0 : 15 + PF);
        cb = (uint32_t)cbase * 8u;
        eb_cur = (uint32_t)(cbase - d_enter(c)) * 8u + nbias;
        eb_nxt = (uint32_t)(cbase - d_enter(n)) * 8u + nbias;
        wb_nxt = (uint32_t)(rowc + x0 + 16 * DIR - d_lo(n)) * 8u;
        const int sbase = rowp + x0 - (DIR > 0 ? 0 : 15);
        sb_n = (uint32_t)sbase * 32u + (uint32_t)(2 * l);
        sb_f = (uint32_t)sbase * (uint32_t)D + (uint32_t)(NB * l);
    };
    auto reload = [&](const WinSeg &n) {
        const int st = n.mode == 0 ? 0 : 16;
#pragma unroll
        for (int i = 0; i < NC; i++)
            nw[i] = __builtin_amdgcn_raw_buffer_load_b128(rr, wb_nxt + (uint32_t)(st * i), 0, 0);
    };
    // the loads of step i + PF, issued by step i
    auto prefetch = [&](int i, const WinSeg &n) {
        const int k = i % PF, cs = cstep(i + PF) * 8;
        rcl[k] = ld64(rl, cb + (uint32_t)cs);
        ren[k] = ld64(rr, (i + PF < 16 ? eb_cur : eb_nxt) + (uint32_t)cs);
        if (i + PF == 16)
            reload(n);
    };

    WinState<NB> S;
    S.reset();
    WinSeg cur = win_seg<NB>(__builtin_amdgcn_readfirstlane(seg_raw(0))),
           nxt = win_seg<NB>(__builtin_amdgcn_readfirstlane(seg_raw(1))), prev = cur;
    uint32_t raw_nn = seg_raw(2);
    // prologue: the loads of segment 0's first PF steps, as the last PF steps
    // of a segment issue them (a dropped store stands for each step's store)
    bases(-1, cur, cur);
#pragma unroll
    for (int i = 16 - PF; i < 16; i++) {
        prefetch(i, cur);
        __builtin_amdgcn_raw_buffer_store_b16((uint16_t)0, rn, kDrop, 0, 0);
        __builtin_amdgcn_sched_barrier(0);
    }

    // one segment: MODE 0 NARROW, 1 FULL / MASKED; CHECK: the boundary tests
    auto segment = [&](int k, auto mode_tag, auto check_tag) {
        constexpr int MODE = decltype(mode_tag)::value;
        constexpr bool CHECK = decltype(check_tag)::value;
        constexpr int DPL = MODE == 0 ? 2 : NB;
        const int x0 = DIR > 0 ? 16 * k : 16 * nt - 1 - 16 * k;
        const int qs = 16 * k;
        const bool masked = cur.mode == 2;
        const int dbase = MODE == 0 ? 16 * cur.b0 + 2 * l : NB * l;
        uint32_t pen[DPL];
#pragma unroll
        for (int j = 0; j < DPL; j++)
            pen[j] = MODE == 1 && masked && !((cur.mask >> ((NB * l + j) >> 4)) & 1u) ? kWinPenalty : 0u;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int k8 = i % PF;
            if (i == 0) {
#pragma unroll
                for (int j = 0; j < DPL; j++)
                    win[j] = MODE == 0 ? half_code(nw[0], 1 - j) : half_code(nw[(NB - 1 - j) >> 1], (NB - 1 - j) & 1);
            } else if constexpr (DIR > 0) {
                const int sl = ((-i) % DPL + DPL) % DPL;
                win[sl] = dpp64_shr1(ren[k8], win[sl]);
            } else {
                const int sl = ((i - 1) % DPL + DPL) % DPL;
                win[sl] = dpp64_shl1(ren[k8], win[sl]);
            }
            uint32_t C[DPL];
            const uint64_t cL = rcl[k8];
            const int x = x0 + DIR * i;
#pragma unroll
            for (int j = 0; j < DPL; j++) {
                const uint64_t cr = win[DIR > 0 ? ((j - i) % DPL + DPL) % DPL : (j + i) % DPL];
                uint32_t c = (uint32_t)__popcll(cL ^ cr);
                if constexpr (CHECK) {
                    c = x - dbase >= j ? c : inval;
                    c = umax(c, pen[j]);
                }
                C[j] = c;
            }
            __builtin_amdgcn_sched_barrier(0);
            prefetch(i, nxt);
            const int q = qs + i;
            const bool real = !CHECK || (q >= q0 && q < qend);
            if constexpr (MODE == 0) {
                sgm_step<2, false>(S.SN, C, S.mn, P1, P2, i == 0 ? S.edge_lo : 0xFFFFFFFFu,
                                   i == 0 ? S.edge_hi : 0xFFFFFFFFu);
                const uint32_t off = real ? sb_n + (uint32_t)(32 * sstep(i)) : kDrop;
                __builtin_amdgcn_raw_buffer_store_b16((uint16_t)(S.SN[0] | (S.SN[1] << 8)), rn, off, 0, 0);
                if constexpr (CHECK) { // before the line's first pixel (backward lines): keep the start state
                    const bool pre = q < q0;
                    S.SN[0] = pre ? 0u : S.SN[0];
                    S.SN[1] = pre ? 0u : S.SN[1];
                    S.mn = pre ? 0u : S.mn;
                }
            } else {
                sgm_step<NB, false>(S.SF, C, S.mn, P1, P2);
                if constexpr (CHECK) {
                    if (masked) {
#pragma unroll
                        for (int j = 0; j < NB; j++)
                            S.SF[j] = umin(S.SF[j], kWinAbsent);
                    }
                }
                const uint32_t off = real ? sb_f + (uint32_t)(D * sstep(i)) : kDrop;
                st_bytes<NB, (NB >= 8)>(ro, off, S.SF);
                if constexpr (CHECK) {
                    const bool pre = q < q0;
#pragma unroll
                    for (int j = 0; j < NB; j++)
                        S.SF[j] = pre ? 0u : S.SF[j];
                    S.mn = pre ? 0u : S.mn;
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        S.edge_lo = 0xFFFFFFFFu;
        S.edge_hi = 0xFFFFFFFFu;
    };

    // boundary segments: steps outside the line, x within D of the row start
    auto variant = [&](int k) {
        const int xmin = DIR > 0 ? 16 * k : 16 * nt - 16 - 16 * k;
        const bool check = 16 * k < q0 || 16 * k + 16 > qend || xmin < D || cur.mode == 2;
        return (cur.mode == 0 ? 0 : 2) + (check ? 1 : 0);
    };
    // One segment of a run of the variant V (mode and boundary tests), then
    // the move to the next segment.
    auto one = [&](int k, auto v_tag) {
        constexpr int V = decltype(v_tag)::value;
        if (k > 0 && win_changes<NB>(prev, cur))
            S.exchange(xs + g * D, l, P1, prev, cur);
        bases(k, cur, nxt);
        segment(k, std::integral_constant<int, (V >> 1)>{}, std::integral_constant<bool, (V & 1) != 0>{});
        prev = cur;
        cur = nxt;
        nxt = win_seg<NB>(__builtin_amdgcn_readfirstlane(raw_nn));
        raw_nn = seg_raw(k + 3);
    };
    // Segments run in maximal runs of one variant, each run its own loop.  A
    // variant's unrolled body leaves the prefetch ring in registers of its own
    // choosing; where two variants' bodies meet, the copies between them wait
    // for every load in flight.  With one loop per run that happens only
    // where the variant changes, not at every segment.
    int k = 0;
#pragma nounroll
    while (k < nt) {
        const int v = variant(k);
        if (v == 0) {
#pragma nounroll
            do
                one(k++, std::integral_constant<int, 0>{});
            while (k < nt && variant(k) == 0);
        } else if (v == 1) {
#pragma nounroll
            do
                one(k++, std::integral_constant<int, 1>{});
            while (k < nt && variant(k) == 1);
        } else if (v == 2) {
#pragma nounroll
            do
                one(k++, std::integral_constant<int, 2>{});
            while (k < nt && variant(k) == 2);
        } else {
#pragma nounroll
            do
                one(k++, std::integral_constant<int, 3>{});
            while (k < nt && variant(k) == 3);
        }
    }
}

// Columns: lines grp*4 .. grp*4+3 of a vertical direction (t.dy = +1 or -1).
// Each step loads its codes: a NARROW step one 16-byte load per lane, a FULL /
// MASKED step NB/2.  Inside a segment the count is the segment's own; the
// last PF steps of a segment prefetch the next segment's first steps with NB/2
// loads whatever its mode (NARROW reads its codes from the first and repeats
// the address), so every path into a segment has the same loads in flight and
// the waits stay exact.  On a column x is fixed, so which disparities have
// x - d < 0 is fixed too: folded, with the MASKED penalties, into one
// per-segment floor under the cost.
template <int NB>
__device__ __forceinline__ void path_window_cols(const WinArgs &a, const DirTask &t, uint8_t *narrow, int grp,
                                                 uint8_t *wlds)
{
    constexpr int PF = 4;
    constexpr int D = 16 * NB;
    constexpr int NC = NB / 2;
    typedef uint32_t v4u __attribute__((ext_vector_type(4)));
    uint8_t *xs = wlds; // layout exchange, kLinesPerWave * D bytes of this wave's LDS
    const int lane = threadIdx.x & 63, g = lane >> 4, l = lane & 15;
    const int W = a.W, H = a.H;
    const int dir = t.dy;
    const int nt = (H + 15) >> 4;
    const int ln = min(grp * kLinesPerWave + g, t.nlines - 1); // column x
    const int tile_x = (grp * kLinesPerWave) >> 4;
    const uint32_t nb8 = (uint32_t)(W * H + kLead) * 8u;
    const rsrc_t rl = make_rsrc(a.cl, nb8), rr = make_rsrc(a.cr, nb8);
    const rsrc_t ro = make_rsrc(t.L, a.out_bytes);
    const rsrc_t rn = make_rsrc(narrow, (uint32_t)(W * H) * 32u);
    const uint32_t P1 = (uint32_t)t.P1, P2 = kPlainP2, inval = kPlainInvalid;
    const int q0 = dir > 0 ? 0 : 16 * nt - H, qend = dir > 0 ? H : 16 * nt;
    // the line's tile ranges in traversal order, staged in LDS once (as for rows)
    uint8_t *segs = wlds + kLinesPerWave * D; // kWinMaxTiles bytes of this wave's LDS
    for (int k = lane; k < nt; k += 64)
        segs[k] = a.blk[(dir > 0 ? k : nt - 1 - k) * a.tiles_x + tile_x];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront"); // as in path_window_rows
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    auto seg_raw = [&](int k) -> uint32_t { return segs[k < nt ? k : nt - 1]; }; // the same in every lane
    auto half_code = [](const v4u &v, int half) -> uint64_t {
        return half ? ((uint64_t)v[3] << 32) | v[2] : ((uint64_t)v[1] << 32) | v[0];
    };
    auto d_lo = [&](const WinSeg &sg) { return sg.mode == 0 ? 16 * sg.b0 + 2 * l + 1 : NB * l + NB - 1; };
    // cursors: the code byte offset of the step being prefetched, the store offsets of the current step
    const int y0 = dir > 0 ? 0 : 16 * nt - 1;
    const uint32_t cstep = (uint32_t)(dir * W) * 8u, sstep_n = (uint32_t)(dir * W) * 32u,
                   sstep_f = (uint32_t)(dir * W) * (uint32_t)D;
    uint32_t cf = (uint32_t)(y0 * W + ln + kLead) * 8u;
    uint32_t so_n = (uint32_t)(y0 * W + ln) * 32u + (uint32_t)(2 * l);
    uint32_t so_f = (uint32_t)(y0 * W + ln) * (uint32_t)D + (uint32_t)(NB * l);

    uint64_t rcl[PF];
    v4u rcr[PF][NC];
    auto fetch = [&](int k, const WinSeg &sg, auto n_tag) {
        constexpr int N = decltype(n_tag)::value;
        rcl[k] = ld64(rl, cf);
        const uint32_t lo = cf - (uint32_t)d_lo(sg) * 8u;
        const int st = sg.mode == 0 ? 0 : 16;
#pragma unroll
        for (int i = 0; i < N; i++)
            rcr[k][i] = __builtin_amdgcn_raw_buffer_load_b128(rr, lo + (uint32_t)(st * i), 0, 0);
        cf += cstep;
    };
    using kOne = std::integral_constant<int, 1>;
    using kAll = std::integral_constant<int, NC>;

    WinState<NB> S;
    S.reset();
    WinSeg cur = win_seg<NB>(__builtin_amdgcn_readfirstlane(seg_raw(0))),
           nxt = win_seg<NB>(__builtin_amdgcn_readfirstlane(seg_raw(1))), prev = cur;
    uint32_t raw_nn = seg_raw(2);
#pragma unroll
    for (int k = 0; k < PF; k++) {
        fetch(k, cur, kAll{});
        __builtin_amdgcn_raw_buffer_store_b16((uint16_t)0, rn, kDrop, 0, 0);
        __builtin_amdgcn_sched_barrier(0);
    }

    auto segment = [&](int k, auto mode_tag, auto check_tag) {
        constexpr int MODE = decltype(mode_tag)::value;
        constexpr bool CHECK = decltype(check_tag)::value;
        constexpr int DPL = MODE == 0 ? 2 : NB;
        const int qs = 16 * k;
        const bool masked = cur.mode == 2;
        const int dbase = MODE == 0 ? 16 * cur.b0 + 2 * l : NB * l;
        uint32_t floor_[DPL];
#pragma unroll
        for (int j = 0; j < DPL; j++) {
            const uint32_t pen =
                MODE == 1 && masked && !((cur.mask >> ((NB * l + j) >> 4)) & 1u) ? kWinPenalty : 0u;
            floor_[j] = umax(ln - dbase >= j ? 0u : inval, pen);
        }
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int rk = i % PF;
            uint32_t C[DPL];
            const uint64_t cL = rcl[rk];
#pragma unroll
            for (int j = 0; j < DPL; j++) {
                const uint64_t cr = MODE == 0 ? half_code(rcr[rk][0], 1 - j)
                                              : half_code(rcr[rk][(NB - 1 - j) >> 1], (NB - 1 - j) & 1);
                uint32_t c = (uint32_t)__popcll(cL ^ cr);
                if constexpr (CHECK)
                    c = umax(c, floor_[j]);
                C[j] = c;
            }
            __builtin_amdgcn_sched_barrier(0);
            if (i + PF < 16) {
                if constexpr (MODE == 0)
                    fetch(rk, cur, kOne{});
                else
                    fetch(rk, cur, kAll{});
            } else {
                fetch(rk, nxt, kAll{});
            }
            const int q = qs + i;
            const bool real = !CHECK || (q >= q0 && q < qend);
            if constexpr (MODE == 0) {
                sgm_step<2, false>(S.SN, C, S.mn, P1, P2, i == 0 ? S.edge_lo : 0xFFFFFFFFu,
                                   i == 0 ? S.edge_hi : 0xFFFFFFFFu);
                __builtin_amdgcn_raw_buffer_store_b16((uint16_t)(S.SN[0] | (S.SN[1] << 8)), rn, real ? so_n : kDrop, 0, 0);
                if constexpr (CHECK) {
                    const bool pre = q < q0;
                    S.SN[0] = pre ? 0u : S.SN[0];
                    S.SN[1] = pre ? 0u : S.SN[1];
                    S.mn = pre ? 0u : S.mn;
                }
            } else {
                sgm_step<NB, false>(S.SF, C, S.mn, P1, P2);
                if constexpr (CHECK) {
                    if (masked) {
#pragma unroll
                        for (int j = 0; j < NB; j++)
                            S.SF[j] = umin(S.SF[j], kWinAbsent);
                    }
                }
                st_bytes<NB, (NB >= 8)>(ro, real ? so_f : kDrop, S.SF);
                if constexpr (CHECK) {
                    const bool pre = q < q0;
#pragma unroll
                    for (int j = 0; j < NB; j++)
                        S.SF[j] = pre ? 0u : S.SF[j];
                    S.mn = pre ? 0u : S.mn;
                }
            }
            so_n += sstep_n;
            so_f += sstep_f;
            __builtin_amdgcn_sched_barrier(0);
        }
        S.edge_lo = 0xFFFFFFFFu;
        S.edge_hi = 0xFFFFFFFFu;
    };

    // x >= D - 1 on every line of the wave: no disparity is ever invalid
    const bool all_valid = grp * kLinesPerWave >= D - 1;
    auto variant = [&](int k) {
        const bool check = 16 * k < q0 || 16 * k + 16 > qend || !all_valid || cur.mode == 2;
        return (cur.mode == 0 ? 0 : 2) + (check ? 1 : 0);
    };
    auto one = [&](int k, auto v_tag) {
        constexpr int V = decltype(v_tag)::value;
        if (k > 0 && win_changes<NB>(prev, cur))
            S.exchange(xs + g * D, l, P1, prev, cur);
        segment(k, std::integral_constant<int, (V >> 1)>{}, std::integral_constant<bool, (V & 1) != 0>{});
        prev = cur;
        cur = nxt;
        nxt = win_seg<NB>(__builtin_amdgcn_readfirstlane(raw_nn));
        raw_nn = seg_raw(k + 3);
    };
    // maximal runs of one variant, one loop each (see path_window_rows)
    int k = 0;
#pragma nounroll
    while (k < nt) {
        const int v = variant(k);
        if (v == 0) {
#pragma nounroll
            do
                one(k++, std::integral_constant<int, 0>{});
            while (k < nt && variant(k) == 0);
        } else if (v == 1) {
#pragma nounroll
            do
                one(k++, std::integral_constant<int, 1>{});
            while (k < nt && variant(k) == 1);
        } else if (v == 2) {
#pragma nounroll
            do
                one(k++, std::integral_constant<int, 2>{});
            while (k < nt && variant(k) == 2);
        } else {
#pragma nounroll
            do
                one(k++, std::integral_constant<int, 3>{});
            while (k < nt && variant(k) == 3);
        }
    }
}

// Sum over the directions and argmin over the pixel's active range, one tile
// (16x16 pixels, uniform range) at a time per wave.  One lane per active
// 16-disparity block: a range of exactly two blocks puts 2 lanes on a pixel
// (32 pixels per pass) reading the dense narrow volumes, a wider one NB lanes
// (64/NB pixels per pass) reading the full ones; lanes of inactive blocks load
// nothing and bid the largest key.
template <int NB>
__device__ __forceinline__ uint32_t wta_block_key(const uint8_t *__restrict__ src, size_t dir_stride, int ndirs,
                                                  uint32_t d0)
{
    uint32_t ev[4] = {0, 0, 0, 0}, od[4] = {0, 0, 0, 0};
    typedef uint32_t v4u __attribute__((ext_vector_type(4)));
    for (int k = 0; k < ndirs; k++) {
        // read once: stream past the caches
        const v4u q = __builtin_nontemporal_load(reinterpret_cast<const v4u *>(src + k * dir_stride));
        const uint32_t w[4] = {q[0], q[1], q[2], q[3]};
#pragma unroll
        for (int i = 0; i < 4; i++) {
            ev[i] += w[i] & 0x00FF00FFu;
            od[i] += (w[i] >> 8) & 0x00FF00FFu;
        }
    }
    uint32_t key = 0xFFFFFFFFu;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint32_t db = d0 + 4 * i;
        key = umin(key, umin(((ev[i] & 0xFFFFu) << 8) | db, ((od[i] & 0xFFFFu) << 8) | (db + 1)));
        key = umin(key, umin(((ev[i] >> 16) << 8) | (db + 2), ((od[i] >> 16) << 8) | (db + 3)));
    }
    return key;
}

// `part` (0..3): this wave's quarter of the tile's passes, so a tile is spread
// over the 4 waves of a workgroup and no wave runs a whole full-D tile alone.
template <int NB>
__device__ __forceinline__ void wta_window_tile(const uint8_t *__restrict__ L, size_t dir_stride,
                                                const uint8_t *__restrict__ Ln, size_t nstride, int ndirs,
                                                const uint8_t *__restrict__ blk, int tiles_x, int W, int H,
                                                uint8_t *__restrict__ disp, int t, int part)
{
    constexpr int D = 16 * NB;
    const int lane = threadIdx.x & 63;
    const uint32_t rng = __builtin_amdgcn_readfirstlane((uint32_t)blk[t]);
    const int b1 = min((int)(rng >> 4), NB - 1), b0 = min((int)(rng & 15u), b1);
    const int ty = t / tiles_x, tx = t - ty * tiles_x;
    if (b1 - b0 == 1) {
        const int b = b0 + (lane & 1);
        const bool act = b <= b1;
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const int it = 2 * part + i;
            const int r = it * 32 + (lane >> 1), y = ty * 16 + (r >> 4), x = tx * 16 + (r & 15);
            const bool valid = y < H && x < W;
            const size_t p = valid ? (size_t)y * W + x : 0;
            uint32_t key = 0xFFFFFFFFu;
            if (act && valid)
                key = wta_block_key<NB>(Ln + p * 32 + 16 * (lane & 1), nstride, ndirs, 16u * b);
            key = umin(key, dpp<kQuadSwap1>(key));
            if (!(lane & 1) && valid)
                disp[p] = (uint8_t)(key & 0xFFu);
        }
    } else {
        constexpr int PPW = 64 / NB, PASSES = 256 / PPW / 4; // passes per part
        const int sub = lane % NB;
        const bool act = sub >= b0 && sub <= b1;
#pragma unroll 2
        for (int i = 0; i < PASSES; i++) {
            const int it = PASSES * part + i;
            const int r = it * PPW + lane / NB, y = ty * 16 + (r >> 4), x = tx * 16 + (r & 15);
            const bool valid = y < H && x < W;
            const size_t p = valid ? (size_t)y * W + x : 0;
            uint32_t key = 0xFFFFFFFFu;
            if (act && valid)
                key = wta_block_key<NB>(L + p * D + 16 * sub, dir_stride, ndirs, 16u * sub);
            key = group_min<NB>(key);
            if (sub == 0 && valid)
                disp[p] = (uint8_t)(key & 0xFFu);
        }
    }
}

} // namespace sgmk
