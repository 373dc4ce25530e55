// sgm_cnn.hip -- the window-proposal network of the learned path: inference
// and training on gfx950, bf16 MFMA with fp32 accumulation.
//
// Network (no reference counterpart; oracle/cnn_oracle.py is its float
// restatement): features of a 4x-downsampled pair (32 bins of 3x3 absolute
// differences of 4x4 block sums, the right image at 4 sub-block phases, the
// best phase per bin), 6 conv layers 3x3 32->32 (ReLU on 1-5, fp32 logits
// from 6 plus the fixed residual -32 * features: a low matching cost is a high
// logit before the network's correction), a softmax over the 32 disparity
// bins (4 disparities each, D = 128) per low-resolution pixel, and per
// 16x16-pixel tile the mean of its 4x4 pixels' distributions: the 32-disparity
// window (two 16-blocks) with the most mass wins and that mass is the
// confidence.  Training minimises the per-pixel cross-entropy against the
// histogram of the plain full-D SGM disparities of the pixel's 4x4
// full-resolution block.
//
// Convolutions are implicit GEMMs on v_mfma_f32_16x16x16bf16_1k: a wave owns
// 16 consecutive pixels of a row x 32 output channels (two 16-channel
// tiles); K = 9 taps x 32 input channels = 18 steps of 16.  The layer's packed
// bf16 weights (one operand of all 18 steps) stay in registers for the
// wave's whole pixel loop; activations are NHWC bf16, 64 bytes per pixel, and
// a pixel operand is one 8-byte LDS read per lane from the staged patch.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <vector>

#include "sgm_device.h"
#include "sgm_internal.h"

namespace {

constexpr int kC = 32;             // channels = disparity bins
constexpr int kLayers = 6;
constexpr int kWPer = kC * kC * 9; // weights per layer
constexpr int kPPer = kWPer + kC;  // parameters per layer (weights + bias)
constexpr int kNParam = kLayers * kPPer;
constexpr int kDisp = 128;
constexpr float kSkip = 32.f; // logits = last conv - kSkip * features

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t to_bf16(float f)
{
    uint32_t u = __float_as_uint(f);
    u = u + 0x7FFFu + ((u >> 16) & 1u); // round to nearest, ties to even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// ---- features ----
// f_k(y, x) = min over the 4 phases ph of sum over the clamped 3x3 window of
// |Ld - Rd_ph(x' - k)| (4080 where x' - k < 0), divided by 9*4080, bf16.  Ld is
// the 4x4 block sum of the left image; Rd_ph that of the right image at phase
// ph: block x covers full-resolution columns 4x - ph .. 4x - ph + 3, clamped at
// 0, so bin k matches full-resolution disparities 4k + ph.
//
// A workgroup makes 8 rows x 32 columns of pixels, all 32 bins.  It first
// forms the block sums it needs (rows y0-1 .. y0+8, clamped) into LDS straight
// from the images: Ld from one aligned word per image row, all four Rd_ph from
// two (a byte-aligned window per phase, v_alignbyte) -- v_dot4 with ones sums
// the bytes.  Thread (k, column group) then takes columns one at a time: per
// row and phase the three horizontal differences of the window (v_sad), then
// the vertical sum of three rows, so each output costs a few operations per
// phase instead of nine.  The 32 threads of a pixel write its 64 bytes
// contiguously.
constexpr int kFeatTX = 32, kFeatTY = 8;
constexpr int kPhases = 4;
__device__ __forceinline__ uint32_t sum4(uint32_t v, uint32_t acc) { return __builtin_amdgcn_udot4(v, 0x01010101u, acc, false); }
// |a - b| + acc in one v_sad_u32 (the compiler forms sub, max and add)
__device__ __forceinline__ uint32_t sad(uint32_t a, uint32_t b, uint32_t acc)
{
    uint32_t r;
    asm("v_sad_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(acc));
    return r;
}

__global__ __launch_bounds__(256) void k_features(const uint8_t *__restrict__ Lim, const uint8_t *__restrict__ Rim,
                                                  int W, int H, uint16_t *__restrict__ X)
{
    const int w = W / 4, h = H / 4;
    constexpr int LC = kFeatTX + 2, RC = kFeatTX + 33, NR = kFeatTY + 2;
    __shared__ int sl[NR][LC];              // Ld, columns x0-1 .. x0+32
    __shared__ int sr[kPhases][NR][RC];     // Rd_ph, columns x0-32 .. x0+32
    const int x0 = blockIdx.x * kFeatTX, y0 = blockIdx.y * kFeatTY;
    const int tid = threadIdx.x;
    // Staging: every load unconditional (indices clamped, stores predicated), so
    // all of a thread's loads are in flight at once.
    constexpr int LIT = (NR * LC + 255) / 256, RIT = (NR * RC + 255) / 256;
    uint32_t lw[LIT][4], rhi[RIT][4], rlo[RIT][4];
#pragma unroll
    for (int it = 0; it < LIT; it++) {
        const int i = min(tid + 256 * it, NR * LC - 1);
        const int r = i / LC, c = i - r * LC;
        const int Y = min(max(y0 - 1 + r, 0), h - 1), Xc = min(max(x0 - 1 + c, 0), w - 1);
#pragma unroll
        for (int dy = 0; dy < 4; dy++)
            lw[it][dy] = *reinterpret_cast<const uint32_t *>(Lim + (size_t)(4 * Y + dy) * W + 4 * Xc);
    }
#pragma unroll
    for (int it = 0; it < RIT; it++) {
        const int i = min(tid + 256 * it, NR * RC - 1);
        const int r = i / RC, c = i - r * RC;
        const int Y = min(max(y0 - 1 + r, 0), h - 1), Xc = min(max(x0 - 32 + c, 0), w - 1);
#pragma unroll
        for (int dy = 0; dy < 4; dy++) {
            const uint8_t *row = Rim + (size_t)(4 * Y + dy) * W;
            rhi[it][dy] = *reinterpret_cast<const uint32_t *>(row + 4 * Xc);
            rlo[it][dy] = *reinterpret_cast<const uint32_t *>(row + 4 * max(Xc - 1, 0));
        }
    }
#pragma unroll
    for (int it = 0; it < LIT; it++) {
        const int i = tid + 256 * it;
        uint32_t acc = 0;
#pragma unroll
        for (int dy = 0; dy < 4; dy++)
            acc = sum4(lw[it][dy], acc);
        if (i < NR * LC)
            sl[i / LC][i % LC] = (int)acc;
    }
#pragma unroll
    for (int it = 0; it < RIT; it++) {
        const int i = tid + 256 * it;
        const int r = min(i, NR * RC - 1) / RC, c = min(i, NR * RC - 1) % RC;
        const bool edge = x0 - 32 + c <= 0; // columns 4Xc-4 .. 4Xc-1 left of the image repeat column 0
        uint32_t acc[kPhases] = {0, 0, 0, 0};
#pragma unroll
        for (int dy = 0; dy < 4; dy++) {
            const uint32_t hi = rhi[it][dy];
            const uint32_t lo = edge ? (hi & 0xFFu) * 0x01010101u : rlo[it][dy];
            acc[0] = sum4(hi, acc[0]);
#pragma unroll
            for (int ph = 1; ph < kPhases; ph++)
                acc[ph] = sum4(__builtin_amdgcn_alignbyte(hi, lo, 4 - ph), acc[ph]);
        }
        if (i < NR * RC) {
#pragma unroll
            for (int ph = 0; ph < kPhases; ph++)
                sr[ph][r][c] = (int)acc[ph];
        }
    }
    __syncthreads();
    const int k = tid & 31, cg = tid >> 5;
    // interior workgroups: no window column is clamped and x' - k >= 0 for every bin
    const bool interior = x0 >= 32 && x0 + kFeatTX < w;
    for (int xl = cg; xl < kFeatTX; xl += 8) {
        const int x = x0 + xl;
        if (x >= w)
            break;
        int best[kFeatTY];
#pragma unroll
        for (int y = 0; y < kFeatTY; y++)
            best[y] = 0x7FFFFFFF;
        if (interior) {
#pragma unroll
            for (int ph = 0; ph < kPhases; ph++) {
                int h3[NR];
#pragma unroll
                for (int r = 0; r < NR; r++) {
                    uint32_t acc = 0;
#pragma unroll
                    for (int j = 0; j < 3; j++)
                        acc = sad((uint32_t)sl[r][xl + j], (uint32_t)sr[ph][r][xl + j - k + 31], acc);
                    h3[r] = (int)acc;
                }
#pragma unroll
                for (int y = 0; y < kFeatTY; y++)
                    best[y] = min(best[y], h3[y] + h3[y + 1] + h3[y + 2]);
            }
        } else {
            // edge workgroups: clamped window columns, and bins whose x' - k < 0
            // cost 4080 -- fixed per (column, bin), so that column's term is
            // made 0 (R := L) and the 4080s start the row sum
            int li[3], ri[3];
            bool valid[3];
            uint32_t base = 0;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const int cx = min(max(x - 1 + j, 0), w - 1);
                li[j] = cx - x0 + 1;
                ri[j] = cx - k - x0 + 32; // in the staged range for every k (cx >= x0 - 1)
                valid[j] = cx - k >= 0;
                base += valid[j] ? 0u : 4080u;
            }
#pragma unroll
            for (int ph = 0; ph < kPhases; ph++) {
                int h3[NR];
#pragma unroll
                for (int r = 0; r < NR; r++) {
                    uint32_t acc = base;
#pragma unroll
                    for (int j = 0; j < 3; j++) {
                        const uint32_t L = (uint32_t)sl[r][li[j]], R = (uint32_t)sr[ph][r][ri[j]];
                        acc = sad(L, valid[j] ? R : L, acc);
                    }
                    h3[r] = (int)acc;
                }
#pragma unroll
                for (int y = 0; y < kFeatTY; y++)
                    best[y] = min(best[y], h3[y] + h3[y + 1] + h3[y + 2]);
            }
        }
#pragma unroll
        for (int y = 0; y < kFeatTY; y++)
            if (y0 + y < h)
                X[((size_t)(y0 + y) * w + x) * kC + k] = to_bf16((float)best[y] / (float)(9 * 4080));
    }
}

// ---- convolution (implicit GEMM on MFMA) ----
// mode 0: Y = relu(conv + b) as bf16      (hidden layers)
// mode 1: Y = conv + b - kSkip * features as fp32 (logits; `mask` = features)
// mode 2: Y = conv * (mask > 0) as bf16    (backward data through a ReLU)
// Wp: packed bf16 B operands, [co][k] with k = tap*32 + ci, 32 x 288.
// A workgroup of 4 waves makes a block of 4 rows x 16 pixels x 32 channels.
// Its input patch -- 6 rows x 18 pixels x 32 channels, out-of-image pixels as
// zeros -- is staged in LDS by 16-byte loads (two per thread), double
// buffered: the next block's patch is in flight while the MFMAs of this one
// run.  Wave w reads its 18 A operands (row w + ky, pixel lr + kx, 4 channels)
// from the patch; the layer's B operands stay in registers.
constexpr int kConvWaves = 4;
constexpr int kPatchRows = kConvWaves + 2, kPatchCols = 18;
constexpr int kPatchPx = 72;                          // bytes per staged pixel (64 + 8 against bank conflicts)
constexpr int kPatchBytes = kPatchRows * kPatchCols * kPatchPx;
constexpr int kPatchChunks = kPatchRows * kPatchCols * 4; // 16-byte chunks (4 per pixel)

template <int MODE>
__global__ __launch_bounds__(256) void k_conv(const uint16_t *__restrict__ X, int w, int h,
                                              const uint16_t *__restrict__ Wp, const float *__restrict__ bias,
                                              const uint16_t *__restrict__ mask, void *__restrict__ Y)
{
    __shared__ __align__(16) uint8_t patch[2][kPatchBytes];
    __shared__ uint2 wlds[kC * 288 / 4]; // 32 x 288 bf16
    for (int i = threadIdx.x; i < kC * 288 / 4; i += 256)
        wlds[i] = reinterpret_cast<const uint2 *>(Wp)[i];
    const int l = threadIdx.x & 63, lr = l & 15, lq = l >> 4, wv = threadIdx.x >> 6;
    const int bx = (w + 15) / 16, by = (h + kConvWaves - 1) / kConvWaves, nblk = bx * by;
    const sgmk::rsrc_t rx = sgmk::make_rsrc(X, (uint32_t)((size_t)w * h * kC * 2));
    typedef uint32_t v4u __attribute__((ext_vector_type(4)));
    // this thread's (up to) two 16-byte chunks of a block's patch
    v4u stage[2];
    auto fetch = [&](int bk) {
        const int y0 = (bk / bx) * kConvWaves - 1, x0 = (bk % bx) * 16 - 1;
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int c = threadIdx.x + 256 * u;
            const int px = c >> 2, q = c & 3, r = px / kPatchCols, col = px - r * kPatchCols;
            const int yy = y0 + r, xx = x0 + col;
            const bool in = bk < nblk && c < kPatchChunks && yy >= 0 && yy < h && xx >= 0 && xx < w;
            const uint32_t off = (uint32_t)(((yy * w + xx) * kC + 8 * q) * 2);
            stage[u] = __builtin_amdgcn_raw_buffer_load_b128(rx, in ? off : sgmk::kDrop, 0, 0);
        }
    };
    auto commit = [&](int buf) {
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int c = threadIdx.x + 256 * u;
            if (c < kPatchChunks)
                *reinterpret_cast<v4u *>(&patch[buf][(c >> 2) * kPatchPx + 16 * (c & 3)]) = stage[u];
        }
    };
    float bch[2][4] = {};
    if constexpr (MODE != 2) {
#pragma unroll
        for (int nt = 0; nt < 2; nt++)
#pragma unroll
            for (int i = 0; i < 4; i++)
                bch[nt][i] = bias[16 * nt + 4 * lq + i];
    }
    int bk = blockIdx.x;
    fetch(bk);
    commit(0);
    __syncthreads(); // weights and the first patch staged
    s16x4 B[18][2];
#pragma unroll
    for (int kb = 0; kb < 18; kb++)
#pragma unroll
        for (int nt = 0; nt < 2; nt++)
            B[kb][nt] = __builtin_bit_cast(s16x4, wlds[((16 * nt + lr) * 288 + 16 * kb + 4 * lq) / 4]);
    int buf = 0;
    for (; bk < nblk; bk += gridDim.x) {
        fetch(bk + gridDim.x);
        const uint8_t *pt = patch[buf];
        f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < 18; kb++) {
            const int tap = kb >> 1, ky = tap / 3, kx = tap % 3;
            const int ci = (kb & 1) * 16 + 4 * lq;
            const uint2 v = *reinterpret_cast<const uint2 *>(pt + ((wv + ky) * kPatchCols + lr + kx) * kPatchPx + 2 * ci);
            const s16x4 a = __builtin_bit_cast(s16x4, v);
            // D^T = W X^T: the weights are the A operand, the pixels the B
            // operand, so each lane ends with 4 consecutive channels of one pixel
            acc0 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(B[kb][0], a, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(B[kb][1], a, acc1, 0, 0, 0);
        }
        // lane holds D[channel 16*nt + 4*lq + i][pixel lr]
        const int y = (bk / bx) * kConvWaves + wv, x = (bk % bx) * 16 + lr;
        if (y < h && x < w) {
            const size_t o = ((size_t)y * w + x) * kC;
#pragma unroll
            for (int nt = 0; nt < 2; nt++) {
                const f32x4 acc = nt ? acc1 : acc0;
                const int c0 = 16 * nt + 4 * lq;
                if constexpr (MODE == 0) {
                    uint32_t pk[2];
#pragma unroll
                    for (int i = 0; i < 4; i += 2)
                        pk[i / 2] = (uint32_t)to_bf16(fmaxf(acc[i] + bch[nt][i], 0.f)) |
                                    ((uint32_t)to_bf16(fmaxf(acc[i + 1] + bch[nt][i + 1], 0.f)) << 16);
                    *reinterpret_cast<uint2 *>(reinterpret_cast<uint16_t *>(Y) + o + c0) = make_uint2(pk[0], pk[1]);
                } else if constexpr (MODE == 1) {
                    // logits: the fixed residual from the matching costs (passed as `mask`)
                    const uint2 fm = *reinterpret_cast<const uint2 *>(mask + o + c0);
                    const uint16_t f[4] = {(uint16_t)fm.x, (uint16_t)(fm.x >> 16), (uint16_t)fm.y,
                                           (uint16_t)(fm.y >> 16)};
                    float4 r;
                    r.x = acc[0] + bch[nt][0] - kSkip * from_bf16(f[0]);
                    r.y = acc[1] + bch[nt][1] - kSkip * from_bf16(f[1]);
                    r.z = acc[2] + bch[nt][2] - kSkip * from_bf16(f[2]);
                    r.w = acc[3] + bch[nt][3] - kSkip * from_bf16(f[3]);
                    *reinterpret_cast<float4 *>(reinterpret_cast<float *>(Y) + o + c0) = r;
                } else {
                    const uint2 mm = *reinterpret_cast<const uint2 *>(mask + o + c0);
                    const uint16_t m[4] = {(uint16_t)mm.x, (uint16_t)(mm.x >> 16), (uint16_t)mm.y,
                                           (uint16_t)(mm.y >> 16)};
                    uint32_t pk[2];
#pragma unroll
                    for (int i = 0; i < 4; i += 2)
                        pk[i / 2] = (uint32_t)to_bf16(from_bf16(m[i]) > 0.f ? acc[i] : 0.f) |
                                    ((uint32_t)to_bf16(from_bf16(m[i + 1]) > 0.f ? acc[i + 1] : 0.f) << 16);
                    *reinterpret_cast<uint2 *>(reinterpret_cast<uint16_t *>(Y) + o + c0) = make_uint2(pk[0], pk[1]);
                }
            }
        }
        buf ^= 1;
        commit(buf); // the next block's patch, into the buffer not read above
        __syncthreads();
    }
}

// ---- weight gradient: dW[co][ci][tap] += sum_p dY[p][co] * X[p + off(tap)][ci] ----
// One wave per (row band, kernel row ky): K = pixels in steps of 16 through
// LDS tiles stored channel-major, so each MFMA operand is one 8-byte LDS read.
constexpr int kBandRows = 2;
__global__ __launch_bounds__(64) void k_conv_wgrad(const uint16_t *__restrict__ X, const uint16_t *__restrict__ dY,
                                                   int w, int h, float *__restrict__ dW)
{
    __shared__ uint16_t sdy[kC][16];
    __shared__ uint16_t sx[kC][20]; // 16 pixels + halo, padded
    const int l = threadIdx.x, lr = l & 15, lq = l >> 4;
    const int ky = blockIdx.y;
    const int y0 = blockIdx.x * kBandRows;
    f32x4 acc[3][2][2];
#pragma unroll
    for (int kx = 0; kx < 3; kx++)
#pragma unroll
        for (int mt = 0; mt < 2; mt++)
#pragma unroll
            for (int nt = 0; nt < 2; nt++)
                acc[kx][mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int y = y0; y < min(y0 + kBandRows, h); y++) {
        const int ys = y + ky - 1;
        for (int x0 = 0; x0 < w; x0 += 16) {
            // stage dY[y][x0..x0+15][:] and X[ys][x0-1..x0+16][:] channel-major
            for (int i = l; i < 16 * kC; i += 64) {
                const int px = i / kC, c = i % kC, x = x0 + px;
                sdy[c][px] = x < w ? dY[((size_t)y * w + x) * kC + c] : (uint16_t)0;
            }
            for (int i = l; i < 18 * kC; i += 64) {
                const int px = i / kC, c = i % kC, x = x0 - 1 + px;
                sx[c][px] = (x >= 0 && x < w && ys >= 0 && ys < h) ? X[((size_t)ys * w + x) * kC + c]
                                                                  : (uint16_t)0;
            }
            __syncthreads();
#pragma unroll
            for (int mt = 0; mt < 2; mt++) {
                s16x4 a;
#pragma unroll
                for (int i = 0; i < 4; i++)
                    a[i] = (short)sdy[16 * mt + lr][4 * lq + i];
#pragma unroll
                for (int kx = 0; kx < 3; kx++)
#pragma unroll
                    for (int nt = 0; nt < 2; nt++) {
                        s16x4 b;
#pragma unroll
                        for (int i = 0; i < 4; i++)
                            b[i] = (short)sx[16 * nt + lr][4 * lq + i + kx];
                        acc[kx][mt][nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[kx][mt][nt], 0, 0, 0);
                    }
            }
            __syncthreads();
        }
    }
    // D[co = 16mt + 4lq + i][ci = 16nt + lr]
#pragma unroll
    for (int kx = 0; kx < 3; kx++)
#pragma unroll
        for (int mt = 0; mt < 2; mt++)
#pragma unroll
            for (int nt = 0; nt < 2; nt++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int co = 16 * mt + 4 * lq + i, ci = 16 * nt + lr;
                    atomicAdd(&dW[(co * kC + ci) * 9 + 3 * ky + kx], acc[kx][mt][nt][i]);
                }
}

// db[c] += sum over a band of pixels of dY[p][c]
__global__ __launch_bounds__(256) void k_bias_grad(const uint16_t *__restrict__ dY, int npix, float *__restrict__ db)
{
    const int c = threadIdx.x & 31, sub = threadIdx.x >> 5;
    float s = 0.f;
    for (int p = blockIdx.x * 8 + sub; p < npix; p += gridDim.x * 8)
        s += from_bf16(dY[(size_t)p * kC + c]);
    atomicAdd(&db[c], s);
}

// ---- head ----
// hist[p][k]: how many of low-resolution pixel p's 16 target disparities (its
// 4x4 full-resolution block) fall in bin k
__global__ __launch_bounds__(256) void k_pixel_hist(const uint8_t *__restrict__ disp, int W, int H,
                                                    uint8_t *__restrict__ hist)
{
    const int w = W / 4, h = H / 4;
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= w * h)
        return;
    const int y = p / w, x = p - y * w;
    uint32_t cnt[kC / 4] = {}; // four 8-bit counters per word
    for (int r = 0; r < 4; r++) {
        const uint32_t v = *reinterpret_cast<const uint32_t *>(disp + (size_t)(4 * y + r) * W + 4 * x);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const uint32_t bin = min((v >> (8 * i) & 0xFFu) >> 2, (uint32_t)kC - 1);
#pragma unroll
            for (int q = 0; q < kC / 4; q++)
                cnt[q] += (bin >> 2) == (uint32_t)q ? 1u << (8 * (bin & 3)) : 0u;
        }
    }
    uint4 *dst = reinterpret_cast<uint4 *>(hist + (size_t)p * kC);
    dst[0] = make_uint4(cnt[0], cnt[1], cnt[2], cnt[3]);
    dst[1] = make_uint4(cnt[4], cnt[5], cnt[6], cnt[7]);
}

// One 32-lane half-wave per 16x16 tile (lane = bin): the tile distribution is
// the mean of its low-resolution pixels' softmax.  Inference writes blk /
// conf; training (hist != nullptr) also writes each pixel's dlogits = (p - q) /
// npix and adds the pixels' cross-entropy / npix to *loss.
__global__ __launch_bounds__(64) void k_tile_head(const float *__restrict__ logits, int w, int h,
                                                  const uint8_t *__restrict__ hist, float thresh,
                                                  uint8_t *__restrict__ blk, float *__restrict__ conf,
                                                  uint16_t *__restrict__ dlogits, float *__restrict__ loss)
{
    const int k = threadIdx.x & 31;
    const int tw = (w + 3) / 4, th = (h + 3) / 4, T = tw * th;
    const int t = blockIdx.x * 2 + (threadIdx.x >> 5);
    if (t >= T)
        return;
    const int ty = t / tw, tx = t - ty * tw;
    const float inv_np = 1.f / (float)(w * h);
    float P = 0.f, ce = 0.f;
    int n = 0;
    for (int y = 4 * ty; y < min(4 * ty + 4, h); y++)
        for (int x = 4 * tx; x < min(4 * tx + 4, w); x++) {
            const size_t o = ((size_t)y * w + x) * kC + k;
            const float z = logits[o];
            float mx = z;
            for (int s = 16; s; s >>= 1)
                mx = fmaxf(mx, __shfl_xor(mx, s, 32));
            const float e = expf(z - mx);
            float se = e;
            for (int s = 16; s; s >>= 1)
                se += __shfl_xor(se, s, 32);
            const float p = e / se;
            P += p;
            n++;
            if (hist) {
                const float q = (float)hist[o] * (1.f / 16.f);
                ce -= q * logf(fmaxf(p, 1e-30f));
                dlogits[o] = to_bf16((p - q) * inv_np);
            }
        }
    P /= (float)n;
    // probability mass of the window of bins [4 b0, 4 b0 + 8)
    float best = -1.f;
    int bb = 0;
    for (int b0 = 0; b0 < kDisp / 16 - 1; b0++) {
        float m = (k >= 4 * b0 && k < 4 * b0 + 8) ? P : 0.f;
        for (int s = 16; s; s >>= 1)
            m += __shfl_xor(m, s, 32);
        if (m > best) {
            best = m;
            bb = b0;
        }
    }
    if (k == 0) {
        conf[t] = best;
        blk[t] = best >= thresh ? (uint8_t)(bb | ((bb + 1) << 4)) : (uint8_t)((kDisp / 16 - 1) << 4);
    }
    if (hist) {
        for (int s = 16; s; s >>= 1)
            ce += __shfl_xor(ce, s, 32);
        if (k == 0)
            atomicAdd(loss, ce * inv_np);
    }
}

// ---- fused inference: all six layers and the head in one launch ----
// A workgroup of 16 waves owns 16x16 low-resolution pixels (4x4 tiles).
// Their receptive field through six 3x3 layers is a 28x28 region, staged once
// in LDS; the layers ping-pong between two LDS copies of that region, so the
// hidden activations never leave the CU.  Layer i computes the region pixels
// that layer i+1 reads, as a flat run [29(i+1), 784-29(i+1)) in row-major
// region order: a 3x3 tap of flat pixel f is f + 28 dy + dx, so 16 consecutive
// flat pixels are one MFMA operand for every tap (pixels off the valid window
// wrap across region rows; they are never read by a valid pixel of the next
// layer).  Pixels outside the image are stored as zeros: the zero padding of
// every layer.  The last layer computes the 16 inner rows, adds the bias and
// the residual, and takes the per-pixel softmax in registers (a pixel's 32
// logits sit in the 4 lanes n, n+16, n+32, n+48).  The head then averages the
// tile's pixels and picks the window, as k_tile_head does.
//
// MFMA v_mfma_f32_16x16x32_bf16: A = weights (row = output channel, 8
// consecutive input channels of one tap per lane), B = pixels (column = pixel,
// the same 8 input channels), so the K order inside the instruction is the
// same for both operands and each lane ends with output channels 4q..4q+3 of
// pixel n.  A region pixel takes 80 bytes (64 of channels and 16 of padding):
// 16-byte reads of 16 consecutive pixels then fall in 16 different bank
// groups, and every tap of a lane is its base address plus a constant.
constexpr int kInfOut = 16;
constexpr int kInfR = kInfOut + 2 * kLayers; // 28
constexpr int kInfPx = kInfR * kInfR;        // 784
constexpr int kInfStride = 80;               // bytes per region pixel
constexpr int kInfThreads = 1024;
constexpr int kWBytes = kC * 288 * 2;        // one layer's packed weights
constexpr int kWRow = 288 * 2 + 16;          // LDS bytes per output channel: 16-byte reads of 16 channels hit 16 bank groups
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t pk_bf16(float a, float b)
{
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{a, b}, bf16x2)); // v_cvt_pk_bf16_f32, RNE
}

__global__ __launch_bounds__(kInfThreads) void k_infer(const uint16_t *__restrict__ X, int w, int h,
                                                       const uint16_t *__restrict__ Wf,
                                                       const float *__restrict__ theta, float thresh,
                                                       uint8_t *__restrict__ blk, float *__restrict__ conf)
{
    __shared__ __align__(16) uint8_t act[2][kInfPx * kInfStride];
    __shared__ __align__(16) uint8_t wts[kC * kWRow];
    __shared__ float bsh[2][kC];
    typedef uint32_t v4u __attribute__((ext_vector_type(4)));
    const int tid = threadIdx.x, wv = tid >> 6, l = tid & 63, n = l & 15, q = l >> 4;
    constexpr int kWaves = kInfThreads / 64;
    const int nbx = (w + kInfOut - 1) / kInfOut;
    const int bx = (int)blockIdx.x % nbx, by = (int)blockIdx.x / nbx;
    const int Y0 = by * kInfOut - kLayers, X0 = bx * kInfOut - kLayers;
    const sgmk::rsrc_t rx = sgmk::make_rsrc(X, (uint32_t)((size_t)w * h * kC * 2));
    const sgmk::rsrc_t rw = sgmk::make_rsrc(Wf, (uint32_t)(kLayers * kWPer * 2));
    auto in_image = [&](int f) {
        const int ry = f / kInfR, Y = Y0 + ry, Xc = X0 + (f - ry * kInfR);
        return Y >= 0 && Y < h && Xc >= 0 && Xc < w;
    };
    auto image_px = [&](int f) {
        const int ry = f / kInfR;
        return (Y0 + ry) * w + X0 + (f - ry * kInfR);
    };

    constexpr int kWChunks = kWBytes / 16;                                 // 1152
    constexpr int kWPerThread = (kWChunks + kInfThreads - 1) / kInfThreads; // 2
    v4u pw[kWPerThread];
    float pb = 0.f;
    auto fetch_layer = [&](int li) {
#pragma unroll
        for (int u = 0; u < kWPerThread; u++) {
            const int c = tid + kInfThreads * u;
            pw[u] = __builtin_amdgcn_raw_buffer_load_b128(
                rw, c < kWChunks ? (uint32_t)(li * kWPer * 2 + 16 * c) : sgmk::kDrop, 0, 0);
        }
        if (tid < kC)
            pb = theta[li * kPPer + kWPer + tid];
    };
    auto commit_layer = [&](int li) {
#pragma unroll
        for (int u = 0; u < kWPerThread; u++) {
            const int c = tid + kInfThreads * u;
            if (c < kWChunks)
                *reinterpret_cast<v4u *>(&wts[(c / 36) * kWRow + 16 * (c % 36)]) = pw[u];
        }
        if (tid < kC)
            bsh[li & 1][tid] = pb;
    };
    // features of the region (zeros outside the image), layer 0's weights and bias
    fetch_layer(0);
    for (int c = tid; c < kInfPx * 4; c += kInfThreads) {
        const int f = c >> 2, ch = c & 3;
        const uint32_t off = in_image(f) ? (uint32_t)((image_px(f) * kC + 8 * ch) * 2) : sgmk::kDrop;
        *reinterpret_cast<v4u *>(&act[0][f * kInfStride + 16 * ch]) =
            __builtin_amdgcn_raw_buffer_load_b128(rx, off, 0, 0);
    }
    commit_layer(0);
    __syncthreads();

    for (int li = 0; li < kLayers; li++) {
        const uint8_t *src = act[li & 1];
        uint8_t *dst = act[(li + 1) & 1];
        bf16x8 A[9][2];
#pragma unroll
        for (int t = 0; t < 9; t++)
#pragma unroll
            for (int nt = 0; nt < 2; nt++)
                A[t][nt] = *reinterpret_cast<const bf16x8 *>(wts + (16 * nt + n) * kWRow + (32 * t + 8 * q) * 2);
        // this lane's 8 biases, read from LDS where they are used (registers are the limit)
        auto bias4 = [&](int nt) { return *reinterpret_cast<const float4 *>(&bsh[li & 1][16 * nt + 4 * q]); };
        if (li + 1 < kLayers)
            fetch_layer(li + 1); // in flight during this layer's MFMAs
        __syncthreads();         // every wave holds this layer's weights: the buffer is free

        // 16 pixels from flat pixel f0 (reads clamped at fmax); taps are constants from one base
        auto conv16 = [&](int f0, int fmax, f32x4 &acc0, f32x4 &acc1) {
            const int fr = min(f0 + n, fmax);
            const uint8_t *base = src + (fr - kInfR - 1) * kInfStride + 16 * q;
            acc0 = f32x4{0.f, 0.f, 0.f, 0.f};
            acc1 = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int t = 0; t < 9; t++) {
                const bf16x8 B = *reinterpret_cast<const bf16x8 *>(base + ((t / 3) * kInfR + (t % 3)) * kInfStride);
                acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[t][0], B, acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[t][1], B, acc1, 0, 0, 0);
            }
        };
        if (li + 1 < kLayers) {
            const int s = 29 * (li + 1), e = kInfPx - 29 * (li + 1), ngr = (e - s + 15) / 16;
            for (int gi = wv; gi < ngr; gi += kWaves) {
                f32x4 acc[2];
                conv16(s + 16 * gi, e - 1, acc[0], acc[1]);
                const int f = s + 16 * gi + n;
                if (f < e) {
                    const bool in = in_image(f);
                    uint8_t *o = dst + f * kInfStride + 8 * q;
#pragma unroll
                    for (int nt = 0; nt < 2; nt++) {
                        const float4 bb = bias4(nt);
                        const uint32_t lo = pk_bf16(fmaxf(acc[nt][0] + bb.x, 0.f), fmaxf(acc[nt][1] + bb.y, 0.f));
                        const uint32_t hi = pk_bf16(fmaxf(acc[nt][2] + bb.z, 0.f), fmaxf(acc[nt][3] + bb.w, 0.f));
                        *reinterpret_cast<uint2 *>(o + 32 * nt) = in ? make_uint2(lo, hi) : make_uint2(0u, 0u);
                    }
                }
            }
        } else {
            // last layer: inner rows; logits, softmax, probabilities into dst as fp32 [16][16][32]
            float *prob = reinterpret_cast<float *>(dst);
            for (int r = wv; r < kInfOut; r += kWaves) {
                f32x4 acc[2];
                const int f0 = (kLayers + r) * kInfR + kLayers;
                conv16(f0, kInfPx - 1, acc[0], acc[1]);
                const int f = f0 + n;
                const uint32_t base = in_image(f) ? (uint32_t)(image_px(f) * kC * 2) : sgmk::kDrop;
                float z[2][4];
#pragma unroll
                for (int nt = 0; nt < 2; nt++) {
                    const float4 bb = bias4(nt);
                    const float b4[4] = {bb.x, bb.y, bb.z, bb.w};
                    const uint32_t o = base == sgmk::kDrop ? base : base + (uint32_t)((16 * nt + 4 * q) * 2);
                    const uint2 fm = __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(rx, o, 0, 0));
                    const uint16_t fv[4] = {(uint16_t)fm.x, (uint16_t)(fm.x >> 16), (uint16_t)fm.y,
                                            (uint16_t)(fm.y >> 16)};
#pragma unroll
                    for (int i = 0; i < 4; i++)
                        z[nt][i] = acc[nt][i] + b4[i] - kSkip * from_bf16(fv[i]);
                }
                float mx = z[0][0];
#pragma unroll
                for (int nt = 0; nt < 2; nt++)
#pragma unroll
                    for (int i = 0; i < 4; i++)
                        mx = fmaxf(mx, z[nt][i]);
                mx = fmaxf(mx, __shfl_xor(mx, 16));
                mx = fmaxf(mx, __shfl_xor(mx, 32));
                float se = 0.f;
#pragma unroll
                for (int nt = 0; nt < 2; nt++)
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        z[nt][i] = expf(z[nt][i] - mx);
                        se += z[nt][i];
                    }
                se += __shfl_xor(se, 16);
                se += __shfl_xor(se, 32);
#pragma unroll
                for (int nt = 0; nt < 2; nt++) {
                    float4 pv;
                    pv.x = z[nt][0] / se;
                    pv.y = z[nt][1] / se;
                    pv.z = z[nt][2] / se;
                    pv.w = z[nt][3] / se;
                    *reinterpret_cast<float4 *>(prob + (r * kInfOut + n) * kC + 16 * nt + 4 * q) = pv;
                }
            }
        }
        if (li + 1 < kLayers)
            commit_layer(li + 1);
        __syncthreads();
    }

    // head: half-wave (tile, lane = bin) as in k_tile_head; 16 tiles
    if (tid >= 16 * 32)
        return;
    const float *prob = reinterpret_cast<const float *>(act[kLayers & 1]);
    const int tw = (w + 3) / 4, th = (h + 3) / 4;
    const int tt = tid >> 5, k = tid & 31;
    const int TY = by * 4 + (tt >> 2), TX = bx * 4 + (tt & 3);
    float P = 0.f;
    int cnt = 0;
    for (int yy = 0; yy < 4; yy++)
        for (int xx = 0; xx < 4; xx++) {
            const int y = 4 * TY + yy, x = 4 * TX + xx;
            if (y < h && x < w) {
                P += prob[((4 * (tt >> 2) + yy) * kInfOut + 4 * (tt & 3) + xx) * kC + k];
                cnt++;
            }
        }
    P /= (float)max(cnt, 1);
    float best = -1.f;
    int bb = 0;
    for (int b0 = 0; b0 < kDisp / 16 - 1; b0++) {
        float m = (k >= 4 * b0 && k < 4 * b0 + 8) ? P : 0.f;
        for (int s = 16; s; s >>= 1)
            m += __shfl_xor(m, s, 32);
        if (m > best) {
            best = m;
            bb = b0;
        }
    }
    if (k == 0 && TY < th && TX < tw) {
        const int t = TY * tw + TX;
        conf[t] = best;
        blk[t] = best >= thresh ? (uint8_t)(bb | ((bb + 1) << 4)) : (uint8_t)((kDisp / 16 - 1) << 4);
    }
}

// ---- parameters ----
// forward pack: Wf[co][tap*32 + ci] = W[co][ci][tap]; backward-data pack
// (conv of dY with flipped taps, channels swapped): Wb[ci][tap*32 + co] = W[co][ci][8 - tap]
__global__ __launch_bounds__(256) void k_pack(const float *__restrict__ theta, uint16_t *__restrict__ Wf,
                                              uint16_t *__restrict__ Wb)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= kLayers * kWPer)
        return;
    const int layer = i / kWPer, r = i - layer * kWPer;
    const int co = r / (kC * 9), ci = (r / 9) % kC, tap = r % 9;
    const uint16_t v = to_bf16(theta[layer * kPPer + r]);
    Wf[layer * kWPer + co * 288 + tap * kC + ci] = v;
    Wb[layer * kWPer + ci * 288 + (8 - tap) * kC + co] = v;
}

// Adam on fp32 master parameters (grad layout = parameter layout)
__global__ __launch_bounds__(256) void k_adam(float *__restrict__ theta, const float *__restrict__ g,
                                              float *__restrict__ m, float *__restrict__ v, int n, float lr,
                                              float b1, float b2, float eps, float c1, float c2)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n)
        return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi, vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    theta[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ __launch_bounds__(256) void k_zero(float *p, int n)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n)
        p[i] = 0.f;
}

} // namespace

struct sgmh_cnn {
    sgmh_ctx *ctx = nullptr;
    float *theta = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
    uint16_t *Wf = nullptr, *Wb = nullptr;
    int step = 0;
    bool packed = false;
    // activations for the current frame size
    int W = 0, H = 0;
    uint16_t *act[kLayers] = {};   // act[i] = input of layer i (bf16 NHWC)
    float *logits = nullptr;
    uint16_t *g0 = nullptr, *g1 = nullptr; // gradient ping-pong (bf16)
    uint8_t *hist = nullptr; // per low-resolution pixel target histograms
    float *loss = nullptr;
    uint8_t *blk = nullptr;
    float *conf = nullptr;
    // the learned frame's census runs on `side` beside the proposal
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
};

namespace {

int cnn_resize(sgmh_cnn *c, int W, int H)
{
    if (c->W == W && c->H == H)
        return 0;
    CK(hipDeviceSynchronize());
    void **bufs[] = {(void **)&c->logits, (void **)&c->g0, (void **)&c->g1,
                     (void **)&c->hist, (void **)&c->loss, (void **)&c->blk, (void **)&c->conf};
    for (void **b : bufs)
        if (*b) {
            CK(hipFree(*b));
            *b = nullptr;
        }
    for (auto &a : c->act)
        if (a) {
            CK(hipFree(a));
            a = nullptr;
        }
    const size_t w = W / 4, h = H / 4, np = w * h, T = ((w + 3) / 4) * ((h + 3) / 4);
    for (auto &a : c->act)
        CK(hipMalloc(&a, np * kC * 2));
    CK(hipMalloc(&c->logits, np * kC * 4));
    CK(hipMalloc(&c->g0, np * kC * 2));
    CK(hipMalloc(&c->g1, np * kC * 2));
    CK(hipMalloc(&c->hist, np * kC));
    CK(hipMalloc(&c->loss, 64));
    CK(hipMalloc(&c->blk, T));
    CK(hipMalloc(&c->conf, T * 4));
    c->W = W;
    c->H = H;
    return 0;
}

int cnn_check(sgmh_cnn *c, int W, int H)
{
    REQUIRE(c, "sgmh_cnn: NULL network");
    REQUIRE(W >= 16 && H >= 16 && W % 4 == 0 && H % 4 == 0 && W <= 16384 && H <= 16384,
            "sgmh_cnn: frame %dx%d (W and H must be multiples of 4, >= 16)", W, H);
    CK(hipSetDevice(c->ctx->device));
    return cnn_resize(c, W, H);
}

int cnn_pack(sgmh_cnn *c, hipStream_t s)
{
    hipLaunchKernelGGL(k_pack, dim3((kLayers * kWPer + 255) / 256), dim3(256), 0, s, c->theta, c->Wf, c->Wb);
    CK(hipGetLastError());
    c->packed = true;
    return 0;
}

int conv_blocks(int w, int h) { return std::max(1, std::min(((w + 15) / 16) * ((h + kConvWaves - 1) / kConvWaves), 1024)); }

// features (layer 0 input) into c->act[0]; packs the weights if needed
int cnn_features(sgmh_cnn *c, const uint8_t *l, const uint8_t *r, int W, int H, hipStream_t s)
{
    const int w = W / 4, h = H / 4;
    REQUIRE((((uintptr_t)l | (uintptr_t)r) & 3) == 0, "sgmh_cnn: images must be 4-byte aligned");
    if (!c->packed && cnn_pack(c, s))
        return -1;
    hipLaunchKernelGGL(k_features, dim3((w + kFeatTX - 1) / kFeatTX, (h + kFeatTY - 1) / kFeatTY), dim3(256), 0, s,
                       l, r, W, H, c->act[0]);
    CK(hipGetLastError());
    return 0;
}

// features + 6 layers, every activation kept for the backward pass; logits in c->logits
int cnn_forward(sgmh_cnn *c, const uint8_t *l, const uint8_t *r, int W, int H, hipStream_t s)
{
    const int w = W / 4, h = H / 4;
    if (cnn_features(c, l, r, W, H, s))
        return -1;
    for (int i = 0; i < kLayers; i++) {
        const bool last = i == kLayers - 1;
        if (last)
            hipLaunchKernelGGL(k_conv<1>, dim3(conv_blocks(w, h)), dim3(256), 0, s, c->act[i], w, h,
                               c->Wf + i * kWPer, c->theta + i * kPPer + kWPer, (const uint16_t *)c->act[0],
                               (void *)c->logits);
        else
            hipLaunchKernelGGL(k_conv<0>, dim3(conv_blocks(w, h)), dim3(256), 0, s, c->act[i], w, h,
                               c->Wf + i * kWPer, c->theta + i * kPPer + kWPer, (const uint16_t *)nullptr,
                               (void *)c->act[i + 1]);
    }
    CK(hipGetLastError());
    return 0;
}

// backward from dlogits (in c->g0) into c->grad (zeroed here)
int cnn_backward(sgmh_cnn *c, int W, int H, hipStream_t s)
{
    const int w = W / 4, h = H / 4, np = w * h;
    hipLaunchKernelGGL(k_zero, dim3((kNParam + 255) / 256), dim3(256), 0, s, c->grad, kNParam);
    uint16_t *dy = c->g0, *dx = c->g1;
    for (int i = kLayers - 1; i >= 0; i--) {
        float *gw = c->grad + i * kPPer, *gb = gw + kWPer;
        hipLaunchKernelGGL(k_conv_wgrad, dim3((h + kBandRows - 1) / kBandRows, 3), dim3(64), 0, s, c->act[i], dy,
                           w, h, gw);
        hipLaunchKernelGGL(k_bias_grad, dim3(std::min((np + 7) / 8, 1024)), dim3(256), 0, s, dy, np, gb);
        if (i > 0) {
            // dX = conv(dY, flipped / transposed W) * (act[i] > 0)
            hipLaunchKernelGGL(k_conv<2>, dim3(conv_blocks(w, h)), dim3(256), 0, s, dy, w, h, c->Wb + i * kWPer,
                               (const float *)nullptr, c->act[i], (void *)dx);
            std::swap(dy, dx);
        }
    }
    CK(hipGetLastError());
    return 0;
}

} // namespace

extern "C" {

size_t sgmh_cnn_num_params(void) { return (size_t)kNParam; }

int sgmh_cnn_create(sgmh_ctx *ctx, sgmh_cnn **out)
{
    REQUIRE(ctx && out, "sgmh_cnn_create: NULL argument");
    *out = nullptr;
    CK(hipSetDevice(ctx->device));
    sgmh_cnn *c = new sgmh_cnn();
    c->ctx = ctx;
    const size_t nb = (size_t)kNParam * 4;
    if (hipMalloc(&c->theta, nb) != hipSuccess || hipMalloc(&c->grad, nb) != hipSuccess ||
        hipMalloc(&c->m, nb) != hipSuccess || hipMalloc(&c->v, nb) != hipSuccess ||
        hipMalloc(&c->Wf, (size_t)kLayers * kWPer * 2) != hipSuccess ||
        hipMalloc(&c->Wb, (size_t)kLayers * kWPer * 2) != hipSuccess) {
        set_err("sgmh_cnn_create: out of device memory");
        sgmh_cnn_destroy(c);
        return -1;
    }
    // zero parameters and optimiser state: callers set weights before use
    std::vector<float> z(kNParam, 0.f);
    if (hipMemcpy(c->theta, z.data(), nb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->m, z.data(), nb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->v, z.data(), nb, hipMemcpyHostToDevice) != hipSuccess) {
        set_err("sgmh_cnn_create: initialisation copy failed");
        sgmh_cnn_destroy(c);
        return -1;
    }
    *out = c;
    return 0;
}

void sgmh_cnn_destroy(sgmh_cnn *c)
{
    if (!c)
        return;
    (void)hipSetDevice(c->ctx->device);
    (void)hipDeviceSynchronize();
    void *bufs[] = {c->theta, c->grad, c->m, c->v, c->Wf, c->Wb, c->logits, c->g0, c->g1,
                    c->hist, c->loss, c->blk, c->conf};
    for (void *b : bufs)
        if (b)
            (void)hipFree(b);
    for (auto a : c->act)
        if (a)
            (void)hipFree(a);
    if (c->ev_fork)
        (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join)
        (void)hipEventDestroy(c->ev_join);
    if (c->side)
        (void)hipStreamDestroy(c->side);
    delete c;
}

int sgmh_cnn_set_params(sgmh_cnn *c, const float *theta, size_t n)
{
    REQUIRE(c && theta && n == (size_t)kNParam, "sgmh_cnn_set_params: need %d parameters", kNParam);
    CK(hipSetDevice(c->ctx->device));
    CK(hipMemcpy(c->theta, theta, n * 4, hipMemcpyHostToDevice));
    std::vector<float> z(kNParam, 0.f);
    CK(hipMemcpy(c->m, z.data(), n * 4, hipMemcpyHostToDevice));
    CK(hipMemcpy(c->v, z.data(), n * 4, hipMemcpyHostToDevice));
    c->step = 0;
    c->packed = false;
    return 0;
}

int sgmh_cnn_get_params(sgmh_cnn *c, float *theta, size_t n)
{
    REQUIRE(c && theta && n == (size_t)kNParam, "sgmh_cnn_get_params: need %d parameters", kNParam);
    CK(hipSetDevice(c->ctx->device));
    CK(hipDeviceSynchronize());
    CK(hipMemcpy(theta, c->theta, n * 4, hipMemcpyDeviceToHost));
    return 0;
}

int sgmh_cnn_propose_dev(sgmh_cnn *c, const uint8_t *l, const uint8_t *r, int W, int H, float thresh,
                         uint8_t *d_blk, float *d_conf, void *stream)
{
    REQUIRE(l && r && d_blk, "sgmh_cnn_propose_dev: NULL argument");
    if (cnn_check(c, W, H))
        return -1;
    hipStream_t s = (hipStream_t)stream;
    if (cnn_features(c, l, r, W, H, s))
        return -1;
    const int w = W / 4, h = H / 4;
    hipLaunchKernelGGL(k_infer, dim3(((w + kInfOut - 1) / kInfOut) * ((h + kInfOut - 1) / kInfOut)),
                       dim3(kInfThreads), 0, s, (const uint16_t *)c->act[0], w, h, (const uint16_t *)c->Wf,
                       (const float *)c->theta, thresh, d_blk, d_conf ? d_conf : c->conf);
    CK(hipGetLastError());
    return 0;
}

// The learned frame: census of both images on a side stream, forked from
// `stream`, while the proposal network runs on `stream`; then the windowed
// paths and argmin over the proposal.  Same result as sgmh_cnn_propose_dev
// followed by sgmh_window_dev at D = 128.
int sgmh_learned_dev(sgmh_cnn *c, const uint8_t *l, const uint8_t *r, int W, int H, float thresh,
                     uint8_t *d_blk, uint8_t *d_disp, void *stream)
{
    REQUIRE(l && r && d_disp, "sgmh_learned_dev: NULL argument");
    if (cnn_check(c, W, H))
        return -1;
    sgmh_ctx *ctx = c->ctx;
    hipStream_t s = (hipStream_t)stream;
    if (!c->side) {
        CK(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
        CK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
        CK(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    }
    if (window_prepare(ctx, W, H, kDisp))
        return -1;
    begin_call(ctx, s);
    CK(hipEventRecord(c->ev_fork, s));
    CK(hipStreamWaitEvent(c->side, c->ev_fork, 0));
    if (window_census(ctx, c->side, l, r, W, H))
        return -1;
    CK(hipEventRecord(c->ev_join, c->side));
    uint8_t *blk = d_blk ? d_blk : c->blk;
    if (cnn_features(c, l, r, W, H, s))
        return -1;
    const int w = W / 4, h = H / 4;
    hipLaunchKernelGGL(k_infer, dim3(((w + kInfOut - 1) / kInfOut) * ((h + kInfOut - 1) / kInfOut)),
                       dim3(kInfThreads), 0, s, (const uint16_t *)c->act[0], w, h, (const uint16_t *)c->Wf,
                       (const float *)c->theta, thresh, blk, c->conf);
    CK(hipGetLastError());
    CK(hipStreamWaitEvent(s, c->ev_join, 0));
    return window_rest(ctx, s, W, H, kDisp, blk, d_disp);
}

// One optimisation step on one frame: forward, per-pixel cross-entropy against the
// target disparities, backward, Adam.  *loss (host) receives the loss before
// the update (this synchronises the stream).  lr <= 0: gradients only, kept in
// the network's gradient buffer (sgmh_cnn_get_grads), no update.
int sgmh_cnn_train_step_dev(sgmh_cnn *c, const uint8_t *l, const uint8_t *r, const uint8_t *target, int W,
                            int H, float lr, float *loss, void *stream)
{
    REQUIRE(l && r && target, "sgmh_cnn_train_step_dev: NULL argument");
    if (cnn_check(c, W, H))
        return -1;
    hipStream_t s = (hipStream_t)stream;
    if (cnn_forward(c, l, r, W, H, s))
        return -1;
    const int w = W / 4, h = H / 4, T = ((w + 3) / 4) * ((h + 3) / 4);
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(256), 0, s, c->loss, 1);
    hipLaunchKernelGGL(k_pixel_hist, dim3((w * h + 255) / 256), dim3(256), 0, s, target, W, H, c->hist);
    hipLaunchKernelGGL(k_tile_head, dim3((T + 1) / 2), dim3(64), 0, s, c->logits, w, h, c->hist, 2.f, c->blk,
                       c->conf, c->g0, c->loss);
    if (cnn_backward(c, W, H, s))
        return -1;
    if (lr > 0.f) {
        c->step++;
        const float b1 = 0.9f, b2 = 0.999f;
        hipLaunchKernelGGL(k_adam, dim3((kNParam + 255) / 256), dim3(256), 0, s, c->theta, c->grad, c->m, c->v,
                           kNParam, lr, b1, b2, 1e-8f, 1.f - powf(b1, (float)c->step),
                           1.f - powf(b2, (float)c->step));
        if (cnn_pack(c, s))
            return -1;
    }
    CK(hipGetLastError());
    if (loss) {
        CK(hipMemcpyAsync(loss, c->loss, 4, hipMemcpyDeviceToHost, s));
        CK(hipStreamSynchronize(s));
    }
    return 0;
}

int sgmh_cnn_get_grads(sgmh_cnn *c, float *g, size_t n)
{
    REQUIRE(c && g && n == (size_t)kNParam, "sgmh_cnn_get_grads: need %d values", kNParam);
    CK(hipSetDevice(c->ctx->device));
    CK(hipDeviceSynchronize());
    CK(hipMemcpy(g, c->grad, n * 4, hipMemcpyDeviceToHost));
    return 0;
}

// Features (layer-0 input, bf16 bits) and logits of the last forward pass.
int sgmh_cnn_get_activations(sgmh_cnn *c, uint16_t *features, float *logits)
{
    REQUIRE(c && c->W > 0, "sgmh_cnn_get_activations: no forward pass yet");
    CK(hipSetDevice(c->ctx->device));
    CK(hipDeviceSynchronize());
    const size_t np = (size_t)(c->W / 4) * (c->H / 4);
    if (features)
        CK(hipMemcpy(features, c->act[0], np * kC * 2, hipMemcpyDeviceToHost));
    if (logits)
        CK(hipMemcpy(logits, c->logits, np * kC * 4, hipMemcpyDeviceToHost));
    return 0;
}

} // extern "C"
