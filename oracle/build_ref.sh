#!/usr/bin/env bash
# build_ref.sh -- compile the reference's own CPU oracles and scene generator
# from their sources under /root/reference into oracle/_ref/ (git-ignored).
# Only used to pin the restatement in oracle/sgm_oracle.c and to write the hash
# fixtures under tests/golden/.  Nothing is copied; the sources are compiled
# where they lie, with the reference's oracle flags (its Makefile:17-18).
set -euo pipefail
REF=${SGM_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -f "$REF/common/sgm_ref.c" ]; then
  echo "build_ref: reference sources not found at $REF; skipping" >&2
  exit 0
fi
mkdir -p "$OUT"
CFL="-O2 -Wall -Wextra -std=gnu11 -I$REF/common -w"
for D in 16 32 64 128 256; do
  gcc $CFL -DSGM_D=$D -DCFLAGS_STR='"ref"' -o "$OUT/sgm_ref_d$D" \
      "$REF/common/sgm_ref.c" "$REF/common/harness.c" "$REF/common/util.c" -lm
  gcc $CFL -DSGM_D=$D -DSGM_PATHS=8 -DCFLAGS_STR='"ref"' -o "$OUT/sgm_ref_d${D}_p8" \
      "$REF/common/sgm_ref.c" "$REF/common/harness.c" "$REF/common/util.c" -lm
  gcc -O2 -w -I"$REF/common" -DSGM_D=$D -o "$OUT/gen_synthetic_d$D" "$REF/common/gen_synthetic.c"
done
# the reference's own tie counter (common/sgm_ref.c:14-16,125-138)
for D in 64 128; do
  gcc $CFL -DSGM_D=$D -DSGM_COUNT_TIES -DCFLAGS_STR='"ref"' -o "$OUT/sgm_ref_d${D}_ties" \
      "$REF/common/sgm_ref.c" "$REF/common/harness.c" "$REF/common/util.c" -lm
done
gcc $CFL -o "$OUT/sgm_ms_ref" "$REF/multiscale/sgm_ms_ref.c" "$REF/common/util.c" -lm
gcc $CFL -fopenmp -o "$OUT/sgm_ms_ref_omp" "$REF/multiscale/sgm_ms_ref.c" "$REF/common/util.c" -lm
echo "build_ref: built $(ls "$OUT" | wc -l) binaries in $OUT"
