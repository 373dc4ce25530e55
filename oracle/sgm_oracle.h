/* sgm_oracle.h -- CPU restatement of the reference's SGM oracles.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (the HIP library, its C
 * host programs, the Python package) links, loads or calls this code. Only the
 * test-suite, the smoke check and the CPU-baseline leg of the benchmark use it,
 * and only as the checker or as the timed CPU baseline.
 *
 * Every function states the reference file:line it restates. All arithmetic is
 * integer; D and the path count are runtime parameters here (compile-time macros
 * in the reference, common/sgm_params.h:13-18).
 *
 * Parity pins (what checks each function against something it cannot simply
 * retype; the tests under tests/ hold them):
 *   orc_census9x7      reference goldens end-to-end; brute-force bit-order
 *                      invariants (bit 0 = (-4,-3), bit 61 = (+4,+3)).
 *   orc_cost           closed form popcount; invalid code; reference goldens.
 *   orc_step           exhaustive small-D brute force against the min-of-all-
 *                      transitions definition; reference goldens.
 *   orc_path_L         single-direction argmin vs an independent DP; goldens.
 *   orc_sgm_plain      the 5 on-disk reference goldens + hashes produced by the
 *                      compiled reference (oracle/_ref) on new scenes.
 *   orc_sgm_ms         the 2 on-disk multiscale goldens + compiled reference.
 *   orc_sgm_ms (D!=64) the reference hard-codes D=64 for the multiscale oracle
 *                      (multiscale/sgm_ms_ref.c:38); every D of 16..256 is pinned
 *                      by an independent numpy restatement of the whole
 *                      composition (tests/test_oracle.py, ms_bruteforce).
 *   orc_argmin ties    tie-rich scenes against the reference's SGM_COUNT_TIES
 *                      build: its hash and its tied-pixel count.
 *   orc_sgm_window     build-defined (no reference counterpart): pinned by the
 *                      identity "all blocks active == orc_sgm_plain" and by a
 *                      brute-force transition-matrix restatement in tests.
 */
#ifndef SGM_ORACLE_H
#define SGM_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

uint64_t orc_fnv1a64(const void *data, size_t n);

/* 9x7 census, bit n set iff neighbour < centre, dy outer, dx inner, centre
 * skipped, clamped coordinates.  common/sgm_ref.c:23-41, sgm_params.h:45-47 */
void orc_census9x7(const uint8_t *im, int W, int H, uint64_t *out);

/* C[(y*W+x)*D+d] = popcount(cl ^ cr[x-d]) or `invalid` when x-d < 0.
 * common/sgm_ref.c:44-56 (invalid 62), multiscale/sgm_ms_ref.c:66-72 (63) */
void orc_cost(const uint64_t *cl, const uint64_t *cr, int W, int H, int D,
              int invalid, uint8_t *C);

/* One recurrence step.  common/sgm_ref.c:61-72 (sat=0), sgm_ms_ref.c:74-84
 * (sat=1: the final add saturates at 255). */
void orc_step(const uint8_t *prev, const uint8_t *c, uint8_t *cur, int D,
              int P1, int P2, int sat);

/* L volume of ONE direction (dx,dy): L[(y*W+x)*D+d], uint8.  The march and the
 * "first pixel on a path: L = C" rule are sgm_ref.c:75-94 / sgm_ms_ref.c:93-133. */
void orc_path_L(const uint8_t *C, int W, int H, int D, int dx, int dy,
                int P1, int P2, int sat, uint8_t *Lout);

/* S += L of one direction (uint16).  sgm_ref.c:75-94 */
void orc_aggregate_dir(const uint8_t *C, int W, int H, int D, int dx, int dy,
                       int P1, int P2, int sat, uint16_t *S);

/* argmin over d, strict '<' so the lowest d wins ties.  sgm_ref.c:120-124 */
void orc_argmin(const uint16_t *S, size_t npix, int D, uint8_t *disp);

/* The direction tables.  sgm_ref.c:96-99 and sgm_ms_ref.c:85-87 */
extern const int orc_dirs_plain[8][2];
extern const int orc_dirs_ms[8][3];

/* Whole plain pipeline (Configuration A semantics, P1=20 P2=192, invalid 62).
 * sgm_ref.c:101-144.  paths = 4 or 8.  Returns 0 on success. */
int orc_sgm_plain(const uint8_t *left, const uint8_t *right, int W, int H,
                  int D, int paths, uint8_t *disp);
/* Same output, OpenMP over rows (horizontal) and columns within a row (others):
 * the output-neutral decomposition of sgm_ms_ref.c:88-92. nthreads<=0: all. */
int orc_sgm_plain_omp(const uint8_t *left, const uint8_t *right, int W, int H,
                      int D, int paths, uint8_t *disp, int nthreads);

/* Multiscale (Configurations B/C): Sobel-X, half, census, cost (invalid 63),
 * merge, 8 weighted paths (P1 40/20/10, P2 200, saturating), decimated argmin.
 * multiscale/sgm_ms_ref.c:47-202.  disp is (W/2)x(H/2).  disp1 (optional, may
 * be NULL) receives the stage-1 half-resolution argmin the reference computes
 * and discards (sgm_ms_ref.c:175-180). */
void orc_sobelx(const uint8_t *im, int W, int H, uint8_t *out);
void orc_half(const uint8_t *im, int W, int H, uint8_t *out);
int orc_sgm_ms(const uint8_t *left, const uint8_t *right, int W, int H, int D,
               uint8_t *disp, uint8_t *disp1);
int orc_sgm_ms_omp(const uint8_t *left, const uint8_t *right, int W, int H, int D,
                   uint8_t *disp, uint8_t *disp1, int nthreads);

/* Windowed SGM (build-defined; no reference counterpart).  Tiles are 16x16
 * pixels, tiles_x = ceil(W/16); tile t's active disparities are the 16-blocks
 * [blk[t] & 15, blk[t] >> 4] (inclusive), i.e. d in [16*b0, 16*(b1+1)).
 * Plain-path rules (P1=20, P2=192, invalid 62) with these changes: the
 * predecessor's L at a disparity outside ITS active range is absent (+inf),
 * mn is the min over its active range, and each candidate of m(d) saturates
 * at 255 as in the reference step (sgm_ref.c:61-72).  S sums the 4 (or 8)
 * directions over the pixel's active range; disp = argmin there, lowest d on
 * ties.  With every tile's range = all blocks this is orc_sgm_plain exactly.
 * paths = 4 or 8.  Returns 0 on success. */
int orc_sgm_window(const uint8_t *left, const uint8_t *right, int W, int H, int D,
                   int paths, const uint8_t *blk, uint8_t *disp);

int orc_omp_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
