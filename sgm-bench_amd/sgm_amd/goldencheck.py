"""Golden discrimination gate: can this golden fail a truncated implementation?

Restates the verdict rule of the reference's scripts/check_golden_discriminates.py
(:59-119): the highest disparity present must be D-1, there must be at least
`min_distinct` distinct values, and at least 0.5% of pixels must lie in the top
decile (d >= int(0.9*D)).  A golden failing any of these can be reproduced by an
implementation that never searches the top of the range.
"""
from __future__ import annotations

import numpy as np


def discriminates(disp: np.ndarray, D: int, min_distinct: int = 8):
    px = np.asarray(disp, np.uint8).ravel()
    top = int(px.max()) if px.size else -1
    distinct = int(np.unique(px).size)
    decile = int(D * 0.9)
    n_top = int((px >= decile).sum())
    cov = 100.0 * n_top / max(px.size, 1)
    ok = top >= D - 1 and distinct >= min_distinct and cov >= 0.5
    return ok, {"max": top, "distinct": distinct, "top_decile_px": n_top, "top_decile_pct": cov}


def main(argv=None):
    import sys
    from . import pgm_read
    argv = argv if argv is not None else sys.argv[1:]
    if len(argv) < 2:
        print("usage: python -m sgm_amd.goldencheck GOLDEN.pgm D [min_distinct]")
        return 1
    ok, st = discriminates(pgm_read(argv[0]), int(argv[1]), int(argv[2]) if len(argv) > 2 else 8)
    print("golden %s D=%s: %s %s" % (argv[0], argv[1], "PASS" if ok else "NOT USABLE", st))
    return 0 if ok else 1


if __name__ == "__main__":
    raise SystemExit(main())
