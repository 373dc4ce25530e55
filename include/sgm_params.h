/* sgm_params.h -- workload parameters of the plain SGM path.
 *
 * Interface restatement of the reference's common/sgm_params.h (macro names and
 * values must match so that an implementation built against either header
 * agrees with the harness it links into).  Width and height are runtime.
 */
#ifndef SGM_PARAMS_H
#define SGM_PARAMS_H

#ifndef SGM_D
#define SGM_D 64          /* disparity range; a multiple of 16 (sgm_params.h:13-15) */
#endif
#ifndef SGM_PATHS
#define SGM_PATHS 4       /* 4: L->R, R->L, T->B, B->T; 8: plus diagonals */
#endif
#define SGM_CENSUS_W 9
#define SGM_CENSUS_H 7

#define SGM_P1 20         /* penalty for a disparity change of 1 */
#define SGM_P2 192        /* penalty for a larger change */

#define SGM_CENSUS_BITS (SGM_CENSUS_W * SGM_CENSUS_H - 1) /* 62: centre skipped */
#define SGM_COST_MAX SGM_CENSUS_BITS
#define SGM_COST_INVALID SGM_COST_MAX                    /* cost where x-d < 0 */

#if SGM_CENSUS_BITS > 64
#error "census descriptor does not fit 64 bits"
#endif
#if (SGM_COST_MAX + SGM_P2) > 255
#error "uint8 path costs need SGM_COST_MAX + SGM_P2 <= 255"
#endif
#if (SGM_D % 16) != 0
#error "SGM_D must be a multiple of 16"
#endif
#if SGM_PATHS != 4 && SGM_PATHS != 8
#error "SGM_PATHS must be 4 or 8"
#endif

/* Semantics shared by every implementation (reference sgm_params.h:44-59):
 * census bit n set iff neighbour < centre, n running over dy (outer) then dx,
 * centre skipped, clamped coordinates; cost = popcount of the XOR, or
 * SGM_COST_INVALID where x-d < 0; uint8 path cost with renormalisation,
 * saturating adds, out-of-range d+-1 counting as 255, first pixel L = C;
 * uint16 sum over the paths; disparity = argmin with the lowest d winning. */

#endif
