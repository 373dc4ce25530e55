/* sgm_scene.h -- deterministic synthetic stereo scenes and small host utilities
 * (PGM I/O, FNV-1a 64, monotonic clock), in plain C (libsgm_host.so).
 *
 * Two generators:
 *
 *  sgm_scene_reference()  byte-identical restatement of the reference's scene
 *      generator, common/gen_synthetic.c:21-112 (3-octave bilinear value noise,
 *      12 stratified slabs with the last pinned to D-1, background D/8, left
 *      built from right).  The reference goldens data/golden/synthetic.pgm,
 *      data/shared/ and data/multiscale/ were computed on its output.
 *
 *  sgm_scene_periodic()   a seeded tie-rich scene (horizontally periodic
 *      texture, flat patches) for the argmin's tie rule.
 *
 *  sgm_scene_planes()     the seeded piecewise-planar generator the benchmark
 *      configurations are quoted on: left = 3x3 box-filtered uniform noise;
 *      right = left warped by a field of rectangular (optionally slanted)
 *      planes with integer disparities in [0, dmax], +-noise uniform intensity
 *      noise, and a per-mille of occluded pixels of fresh random texture.  It
 *      has no reference counterpart; it is pinned by seed -> hash fixtures.
 *
 * Convention for both: left(x) = right(x - d).  Buffers are W*H bytes, row-major,
 * no stride.
 */
#ifndef SGM_SCENE_H
#define SGM_SCENE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Restatement of common/gen_synthetic.c:46-112 for a given D (the reference's
 * compile-time SGM_D) and seed.  gt may be NULL.  Returns 0 on success. */
int sgm_scene_reference(int W, int H, int D, uint32_t seed,
                        uint8_t *left, uint8_t *right, uint8_t *gt);

typedef struct {
    int W, H;
    int dmax;          /* disparities drawn in [0, dmax]; the last plane is pinned to dmax */
    uint32_t seed;
    int planes_lo;     /* plane count drawn uniformly in [planes_lo, planes_hi] */
    int planes_hi;
    int occl_permille; /* per-mille of right-image pixels replaced by random texture */
    int noise;         /* right = left + U[-noise, noise], clamped */
    int slanted;       /* 0: fronto-parallel planes; 1: slanted planes */
} sgm_scene_spec;

/* gt (may be NULL) receives the disparity field in right-image coordinates. */
int sgm_scene_planes(const sgm_scene_spec *spec, uint8_t *left, uint8_t *right, uint8_t *gt);

/* Tie-rich scene: a texture repeating every `period` columns (2..64), flat
 * rectangles, rectangular planes of integer disparity in [0, dmax] and no
 * noise, so that many pixels' path sums have tied minima.  gt (may be NULL)
 * receives the disparity field in right-image coordinates.  Returns 0 on
 * success. */
int sgm_scene_periodic(int W, int H, int dmax, int period, uint32_t seed,
                       uint8_t *left, uint8_t *right, uint8_t *gt);

/* P5 PGM I/O (maxval 255).  sgm_pgm_read returns a malloc'd buffer or NULL and
 * prints the reason to stderr (the reference's common/util.c:19-42 contract). */
uint8_t *sgm_pgm_read(const char *path, int *W, int *H);
int sgm_pgm_write(const char *path, const uint8_t *img, int W, int H);
void sgm_free(void *p);

uint64_t sgm_fnv1a64(const void *data, size_t n);
double sgm_now_ms(void);

#ifdef __cplusplus
}
#endif
#endif
