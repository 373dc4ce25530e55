/* sgm_util.c -- the helper symbols common/sgm.h declares (reference
 * common/util.c:19-65), forwarded to libsgm_host. */
#include "sgm.h"
#include "sgm_scene.h"

uint8_t *pgm_read(const char *path, int *W, int *H) { return sgm_pgm_read(path, W, H); }
int pgm_write(const char *path, const uint8_t *img, int W, int H) { return sgm_pgm_write(path, img, W, H); }
uint64_t fnv1a64(const void *data, size_t n) { return sgm_fnv1a64(data, n); }
double now_ms(void) { return sgm_now_ms(); }
