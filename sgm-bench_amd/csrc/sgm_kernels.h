// sgm_kernels.h -- device-side building blocks shared by the SGM kernels
// (gfx950, wave64).
//
// Work layout of the aggregation kernels: one scanline per 16-lane DPP row,
// four scanlines per wave64.  Lane l of a row owns the DPL = D/16
// consecutive disparities d = l*DPL + j (j = 0..DPL-1) in registers, so
//   * the d-1 / d+1 neighbours inside a lane are registers,
//   * the two that cross a lane boundary are one row_shr:1 / row_shl:1 DPP move,
//   * min over all D is an in-lane min3 chain plus four DPP min steps
//     (quad_perm x2, row_half_mirror, row_mirror) that leave the row minimum in
//     every lane of the row: no LDS, no barrier, no readlane.
// The plain path kernel also has a wave-line layout for frames too small to
// fill the GPU with four lines per wave: one scanline per wave64, DPL = D/64
// (one disparity per lane at D = 64), neighbours by wave_shr:1 / wave_shl:1
// and min over D by a 64-lane DPP reduction read from lane 63.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sgmk {

constexpr int kRowLanes = 16;    // lanes per scanline (one DPP row)
constexpr int kLinesPerWave = 4; // scanlines per wave64

__device__ __forceinline__ uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }
__device__ __forceinline__ uint32_t umax(uint32_t a, uint32_t b) { return a > b ? a : b; }

// DPP move with every lane valid (bound_ctrl irrelevant for these patterns).
template <int CTRL>
__device__ __forceinline__ uint32_t dpp(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, true);
}

constexpr int kRowBcast0 = 0x150; // row_newbcast:0, lane 0 of each row to the whole row

// The value lane 0 of each 16-lane row holds, in every lane of the row.
__device__ __forceinline__ uint64_t row_bcast0(uint64_t v)
{
    return ((uint64_t)dpp<kRowBcast0>((uint32_t)(v >> 32)) << 32) | dpp<kRowBcast0>((uint32_t)v);
}

// DPP move where lanes without a source keep `old`.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_or(uint32_t old, uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, CTRL, 0xF, 0xF, false);
}

constexpr int kQuadSwap1 = 0xB1;      // quad_perm [1,0,3,2]
constexpr int kQuadSwap2 = 0x4E;      // quad_perm [2,3,0,1]
constexpr int kRowHalfMirror = 0x141; // lane i <- 7-i within each half row
constexpr int kRowMirror = 0x140;     // lane i <- 15-i within the row
constexpr int kRowShr1 = 0x111;       // lane i <- lane i-1 (lane 0 has none)
constexpr int kRowShl1 = 0x101;       // lane i <- lane i+1 (lane 15 has none)

// Minimum over the 16 lanes of this lane's DPP row, returned in every lane.
__device__ __forceinline__ uint32_t row_min16(uint32_t v)
{
    v = umin(v, dpp<kQuadSwap1>(v));
    v = umin(v, dpp<kQuadSwap2>(v));
    v = umin(v, dpp<kRowHalfMirror>(v));
    v = umin(v, dpp<kRowMirror>(v));
    return v;
}

// The wave-line layout: one scanline per wave64, lane l owning disparities
// l*DPL .. l*DPL+DPL-1 (DPL = D/64).  Its neighbours across lanes move by the
// whole-wave DPP shifts, and min over D is a reduction over all 64 lanes.
constexpr int kWaveShr1 = 0x138;   // lane i <- lane i-1 across the wave (lane 0 has none)
constexpr int kWaveShl1 = 0x130;   // lane i <- lane i+1 across the wave (lane 63 has none)
constexpr int kRowBcast15 = 0x142; // lane 15 of each row to the whole next row
constexpr int kRowBcast31 = 0x143; // lane 31 to rows 2 and 3

// Minimum over all 64 lanes, wave-uniform: the row minimum in every lane,
// then rows 1 and 3 take in rows 0 and 2 (row_bcast:15, rows 0 and 2 masked
// off) and rows 2 and 3 take in row 1 (row_bcast:31, rows 0 and 1 masked
// off), which leaves the minimum of all four rows in row 3; lane 63 is read.
// Each broadcast step is one v_min_u32_dpp whose masked-off rows keep their
// value (the destination is also the second operand).  Written out, the
// compiler forms it as a move, a masked DPP move and a min.  The s_nop gives
// the DPP read the two wait states a VALU result needs before DPP reads it
// (the compiler's hazard checks do not look inside inline asm).
__device__ __forceinline__ uint32_t wave_min64(uint32_t v)
{
    v = row_min16(v);
    asm("s_nop 1\n\tv_min_u32_dpp %0, %0, %0 row_bcast:15 row_mask:0xa bank_mask:0xf" : "+v"(v));
    asm("s_nop 1\n\tv_min_u32_dpp %0, %0, %0 row_bcast:31 row_mask:0xc bank_mask:0xf" : "+v"(v));
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// Lanes per scanline LPL: 16 (four lines per wave, DPP row moves) or 64 (one
// line per wave, whole-wave moves).
template <int LPL>
struct LineLanes {
    static_assert(LPL == 16 || LPL == 64, "a scanline spans one DPP row or the whole wave");
    static constexpr int shr = LPL == 16 ? kRowShr1 : kWaveShr1;
    static constexpr int shl = LPL == 16 ? kRowShl1 : kWaveShl1;
    static __device__ __forceinline__ uint32_t min(uint32_t v) { return LPL == 16 ? row_min16(v) : wave_min64(v); }
};

// One recurrence step for this lane's DPL disparities (common/sgm_params.h:50-55,
// sgm_ref.c:61-72 and sgm_ms_ref.c:74-84):
//   m(d) = min(Lp(d), Lp(d-1)+P1, Lp(d+1)+P1, mn+P2);  L(d) = C(d) + m(d) - mn
// where mn = min_k Lp(k).  The reference saturates each bracketed add at 255;
// since Lp(d) <= 255 is always one of the candidates, min() gives the same
// result without the clamps.  Out-of-range neighbours (d = -1, d = D) are
// absent, which the DPP `old` value 0xFFFFFFFF expresses.  SAT clamps the
// final sum at 255 (multiscale rule; the plain rule cannot exceed 254).
//
// edge_lo / edge_hi: Lp+P1 of the disparities just below / above the line's
// range, seen by its first / last lane (0xFFFFFFFF: none).  LPL: lanes per
// scanline (16 or 64, LineLanes).
template <int DPL, bool SAT, int LPL = kRowLanes>
__device__ __forceinline__ void sgm_step(uint32_t (&L)[DPL], const uint32_t (&C)[DPL],
                                         uint32_t &mn, uint32_t P1, uint32_t P2,
                                         uint32_t edge_lo = 0xFFFFFFFFu, uint32_t edge_hi = 0xFFFFFFFFu)
{
    using LL = LineLanes<LPL>;
    uint32_t a[DPL];
#pragma unroll
    for (int j = 0; j < DPL; j++)
        a[j] = L[j] + P1;
    const uint32_t below = dpp_or<LL::shr>(edge_lo, a[DPL - 1]); // Lp(l*DPL-1)+P1
    const uint32_t above = dpp_or<LL::shl>(edge_hi, a[0]);       // Lp(l*DPL+DPL)+P1
    const uint32_t jump = mn + P2;
    // -mn through an opaque VALU op: the compiler would fold `C + m + (0 - mn)`
    // back into (C - mn) + m, two adds per disparity instead of one v_add3.
    // With one line per wave mn is wave-uniform: -mn is a scalar op
    // (s_mul_i32, which leaves SCC alone) and v_add3 reads it from an SGPR.
    uint32_t neg;
    if constexpr (LPL == 64)
        asm("s_mul_i32 %0, %1, -1" : "=s"(neg) : "s"(__builtin_amdgcn_readfirstlane(mn)));
    else
        asm("v_sub_u32 %0, 0, %1" : "=v"(neg) : "v"(mn));
    uint32_t lo = 0xFFFFFFFFu;
#pragma unroll
    for (int j = 0; j < DPL; j++) {
        const uint32_t nb_lo = j == 0 ? below : a[j - 1];
        const uint32_t nb_hi = j == DPL - 1 ? above : a[j + 1];
        uint32_t m = umin(umin(L[j], jump), umin(nb_lo, nb_hi));
        uint32_t v = C[j] + m + neg;
        if (SAT)
            v = umin(v, 255u);
        L[j] = v;
        lo = umin(lo, v);
    }
    mn = LL::min(lo);
}

// The saturating step on uint16 pairs (D = 32 * NW): lane l's disparities
// l*2NW .. l*2NW+2NW-1 are Lw[w] = (L(l*2NW+2w), L(l*2NW+2w+1)), low half
// first, and Cw likewise.  Same rule as sgm_step<2*NW, true>: every
// intermediate stays below 2^16 (L <= 255, P1 <= 255, mn + P2 <= 455,
// C + m <= 255 + 455), and m(d) >= mn, so C + m - mn does not wrap.  The
// neighbours d-1 / d+1 inside a lane are one v_alignbit_b32 of two adjacent
// words; across lanes they come by the same row_shr / row_shl DPP moves
// (lane 0 / 15: 0xFFFF, none).
typedef uint16_t u16x2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t pk_u16(u16x2_t v)
{
    uint32_t r;
    __builtin_memcpy(&r, &v, 4);
    return r;
}

__device__ __forceinline__ u16x2_t as_u16x2(uint32_t v)
{
    u16x2_t r;
    __builtin_memcpy(&r, &v, 4);
    return r;
}

__device__ __forceinline__ uint32_t pk_min_u16(uint32_t a, uint32_t b)
{
    const u16x2_t x = as_u16x2(a), y = as_u16x2(b);
    return pk_u16(x < y ? x : y);
}

template <int NW>
__device__ __forceinline__ void sgm_step_sat_pk(uint32_t (&Lw)[NW], const uint32_t (&Cw)[NW], uint32_t &mn,
                                                uint32_t P1, uint32_t P2)
{
    const uint32_t p1 = P1 | (P1 << 16);
    uint32_t a[NW];
#pragma unroll
    for (int w = 0; w < NW; w++)
        a[w] = pk_u16(as_u16x2(Lw[w]) + as_u16x2(p1));
    const uint32_t below = dpp_or<kRowShr1>(0xFFFFFFFFu, a[NW - 1]); // high half: Lp(first d - 1)+P1
    const uint32_t above = dpp_or<kRowShl1>(0xFFFFFFFFu, a[0]);      // low half: Lp(last d + 1)+P1
    // nb[w]: the d-1 neighbours of word w, which are the d+1 neighbours of word w-1
    uint32_t nb[NW + 1];
    nb[0] = __builtin_amdgcn_alignbit(a[0], below, 16);
#pragma unroll
    for (int w = 1; w < NW; w++)
        nb[w] = __builtin_amdgcn_alignbit(a[w], a[w - 1], 16);
    nb[NW] = __builtin_amdgcn_alignbit(above, a[NW - 1], 16);
    const uint32_t jump = mn + P2, jump2 = jump | (jump << 16), mn2 = mn | (mn << 16);
    const uint32_t sat = 255u | (255u << 16);
    uint32_t lo = 0xFFFFFFFFu;
#pragma unroll
    for (int w = 0; w < NW; w++) {
        const uint32_t m = pk_min_u16(pk_min_u16(Lw[w], jump2), pk_min_u16(nb[w], nb[w + 1]));
        Lw[w] = pk_min_u16(pk_u16(as_u16x2(Cw[w]) + as_u16x2(m) - as_u16x2(mn2)), sat);
        lo = pk_min_u16(lo, Lw[w]);
    }
    mn = row_min16(umin(lo & 0xFFFFu, lo >> 16));
}

// Four byte values (each < 256) into one word: two v_lshl_or_b32 and one
// v_perm_b32 taking the low half of each (the perm also keeps the compiler
// from re-associating the ors into four ops).
__device__ __forceinline__ uint32_t pack4(uint32_t a, uint32_t b, uint32_t c, uint32_t d)
{
    return __builtin_amdgcn_perm((d << 8) | c, (b << 8) | a, 0x05040100u);
}

template <int DPL>
__device__ __forceinline__ uint32_t lane_min(const uint32_t (&v)[DPL])
{
    uint32_t m = v[0];
#pragma unroll
    for (int j = 1; j < DPL; j++)
        m = umin(m, v[j]);
    return m;
}

// Pack DPL byte values (each < 256) and store them contiguously.
template <int DPL>
__device__ __forceinline__ void store_bytes(uint8_t *dst, const uint32_t (&v)[DPL])
{
    if constexpr (DPL == 1) {
        *dst = (uint8_t)v[0];
    } else if constexpr (DPL == 2) {
        *reinterpret_cast<uint16_t *>(dst) = (uint16_t)(v[0] | (v[1] << 8));
    } else {
        static_assert(DPL % 4 == 0, "DPL must be 1, 2 or a multiple of 4");
        uint32_t w[DPL / 4];
#pragma unroll
        for (int k = 0; k < DPL / 4; k++)
            w[k] = pack4(v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]);
        if constexpr (DPL == 4) {
            *reinterpret_cast<uint32_t *>(dst) = w[0];
        } else if constexpr (DPL == 8) {
            *reinterpret_cast<uint2 *>(dst) = make_uint2(w[0], w[1]);
        } else {
#pragma unroll
            for (int k = 0; k < DPL / 4; k += 4)
                *reinterpret_cast<uint4 *>(dst + 4 * k) =
                    make_uint4(w[k], w[k + 1], w[k + 2], w[k + 3]);
        }
    }
}

// Load DPL contiguous bytes and widen them.
template <int DPL>
__device__ __forceinline__ void load_bytes(const uint8_t *src, uint32_t (&v)[DPL])
{
    if constexpr (DPL == 1) {
        v[0] = *src;
    } else if constexpr (DPL == 2) {
        const uint32_t h = *reinterpret_cast<const uint16_t *>(src);
        v[0] = h & 0xFFu;
        v[1] = h >> 8;
    } else {
        static_assert(DPL % 4 == 0, "DPL must be 1, 2 or a multiple of 4");
        uint32_t w[DPL / 4];
        if constexpr (DPL == 4) {
            w[0] = *reinterpret_cast<const uint32_t *>(src);
        } else if constexpr (DPL == 8) {
            const uint2 q = *reinterpret_cast<const uint2 *>(src);
            w[0] = q.x;
            w[1] = q.y;
        } else {
#pragma unroll
            for (int k = 0; k < DPL / 4; k += 4) {
                const uint4 q = *reinterpret_cast<const uint4 *>(src + 4 * k);
                w[k] = q.x;
                w[k + 1] = q.y;
                w[k + 2] = q.z;
                w[k + 3] = q.w;
            }
        }
#pragma unroll
        for (int j = 0; j < DPL; j++)
            v[j] = (w[j / 4] >> (8 * (j % 4))) & 0xFFu;
    }
}

} // namespace sgmk
