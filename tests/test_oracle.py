"""The CPU oracle is pinned here against the reference's goldens, the compiled
reference's hashes, and independent formulations of each step (never by
retyping its own formula)."""
import itertools
import os
import sys

import numpy as np
import pytest

import oracle
import sgm_amd
from conftest import make_scene, rng_image


# ---------------------------------------------------------------- census
def census_bruteforce(img):
    """numpy restatement by shifted-image comparison (a different algorithm
    from the oracle's per-pixel loop)."""
    H, W = img.shape
    pad = np.pad(img, ((3, 3), (4, 4)), mode="edge").astype(np.int32)
    c = img.astype(np.int32)
    code = np.zeros((H, W), np.uint64)
    n = 0
    for dy in range(-3, 4):
        for dx in range(-4, 5):
            if dx == 0 and dy == 0:
                continue
            nb = pad[3 + dy:3 + dy + H, 4 + dx:4 + dx + W]
            code |= (nb < c).astype(np.uint64) << np.uint64(n)
            n += 1
    assert n == 62
    return code


@pytest.mark.parametrize("H,W", [(1, 1), (7, 9), (13, 40), (61, 97)])
def test_census_matches_bruteforce(H, W):
    img = rng_image(H, W, H * 1000 + W)
    np.testing.assert_array_equal(oracle.census(img), census_bruteforce(img))


def test_census_bit_order():
    """bit 0 is the neighbour at (dx,dy)=(-4,-3), bit 61 at (+4,+3); the centre
    is skipped, so (+1,0) is bit 31 (sgm_params.h:45-47)."""
    for (dx, dy, bit) in [(-4, -3, 0), (4, 3, 61), (-1, 0, 30), (1, 0, 31), (0, -3, 4)]:
        img = np.full((21, 21), 100, np.uint8)
        img[10 + dy, 10 + dx] = 50
        code = int(oracle.census(img)[10, 10])
        assert code == 1 << bit, (dx, dy, hex(code))


def test_census_edge_replication():
    img = rng_image(9, 12, 3)
    # clamped coordinates == census of the edge-padded image, cropped
    big = np.pad(img, ((5, 5), (6, 6)), mode="edge")
    np.testing.assert_array_equal(oracle.census(img), oracle.census(big)[5:-5, 6:-6])


# ---------------------------------------------------------------- cost
def test_cost_closed_form():
    cl = np.random.default_rng(1).integers(0, 2**62, size=(5, 70), dtype=np.uint64)
    cr = np.random.default_rng(2).integers(0, 2**62, size=(5, 70), dtype=np.uint64)
    C = oracle.cost(cl, cr, 64, invalid=62)
    for y, x, d in itertools.product(range(5), range(0, 70, 7), range(0, 64, 5)):
        want = 62 if x - d < 0 else bin(int(cl[y, x]) ^ int(cr[y, x - d])).count("1")
        assert C[y, x, d] == want
    C63 = oracle.cost(cl, cr, 64, invalid=63)
    assert C63[0, 0, 5] == 63


# ---------------------------------------------------------------- step
def step_transition_matrix(prev, c, P1, P2, sat):
    """L(d) = C(d) + min_d' sat(Lp(d') + pen(d, d')) - min Lp, pen 0 / P1 / P2:
    the textbook transition form of the SGM recurrence."""
    D = prev.size
    d = np.arange(D)
    dist = np.abs(d[:, None] - d[None, :])
    pen = np.where(dist == 0, 0, np.where(dist == 1, P1, P2))
    cand = np.minimum(prev[None, :].astype(np.int64) + pen, 255)
    m = cand.min(axis=1)
    v = c.astype(np.int64) + m - prev.min()
    if sat:
        v = np.minimum(v, 255)
    return v.astype(np.uint8)


@pytest.mark.parametrize("sat,P1,P2", [(False, 20, 192), (True, 40, 200), (True, 10, 200)])
def test_step_exhaustive_small(sat, P1, P2):
    rng = np.random.default_rng(P1 + P2)
    for D in (1, 2, 3, 16):
        for _ in range(400):
            prev = rng.integers(0, 256, D).astype(np.uint8)
            if rng.random() < 0.3:
                prev[rng.integers(0, D)] = 255
            if rng.random() < 0.3:
                prev[:] = prev.min()
            cmax = 62 if not sat else 63
            c = rng.integers(0, cmax + 1, D).astype(np.uint8)
            if not sat:
                # plain-path operating range: Lp - min Lp <= P2 keeps L <= 254
                prev = np.minimum(prev.astype(np.int64), int(prev.min()) + P2).astype(np.uint8)
            np.testing.assert_array_equal(oracle.step(prev, c, P1, P2, sat),
                                          step_transition_matrix(prev, c, P1, P2, sat))


# ---------------------------------------------------------------- paths
def path_bruteforce(C, dx, dy, P1, P2, sat):
    H, W, D = C.shape
    L = np.zeros_like(C)
    ys = range(H - 1, -1, -1) if dy < 0 else range(H)
    xs = range(W - 1, -1, -1) if dx < 0 else range(W)
    for y in ys:
        for x in xs:
            px, py = x - dx, y - dy
            if 0 <= px < W and 0 <= py < H:
                L[y, x] = step_transition_matrix(L[py, px], C[y, x], P1, P2, sat)
            else:
                L[y, x] = C[y, x]
    return L


@pytest.mark.parametrize("dx,dy", [(1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (-1, 1), (1, -1), (-1, -1)])
def test_path_L_matches_transition_form(dx, dy):
    img_l, img_r = rng_image(9, 23, 5), rng_image(9, 23, 6)
    C = oracle.cost(oracle.census(img_l), oracle.census(img_r), 16)
    np.testing.assert_array_equal(oracle.path_L(C, dx, dy), path_bruteforce(C, dx, dy, 20, 192, False))
    np.testing.assert_array_equal(oracle.path_L(C, dx, dy, 10, 200, True),
                                  path_bruteforce(C, dx, dy, 10, 200, True))


def window_bruteforce(left, right, blk, D, paths=4):
    """Windowed SGM from its definition, by transition matrices over the active
    disparity sets: a predecessor disparity outside the predecessor tile's range
    is absent, each candidate Lp(d') + pen(d, d') (pen 0 / P1 / P2 for |d-d'| =
    0 / 1 / more) saturates at 255, the first pixel of a path takes L = C, and
    the argmin runs over the pixel's own range (lowest d on ties)."""
    H, W = left.shape
    C = oracle.cost(oracle.census(left), oracle.census(right), D).astype(np.int64)
    d = np.arange(D)
    pen = np.where(d[:, None] == d[None, :], 0, np.where(np.abs(d[:, None] - d[None, :]) == 1, 20, 192))

    def active(y, x):
        b = int(blk[y // 16, x // 16])
        return (d >= 16 * (b & 15)) & (d < 16 * ((b >> 4) + 1))

    S = np.zeros((H, W, D), np.int64)
    dirs = [(1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (-1, 1), (1, -1), (-1, -1)][:paths]
    for dx, dy in dirs:
        Ld = np.zeros((H, W, D), np.int64)
        for y in (range(H - 1, -1, -1) if dy < 0 else range(H)):
            for x in (range(W - 1, -1, -1) if dx < 0 else range(W)):
                a = active(y, x)
                px, py = x - dx, y - dy
                if 0 <= px < W and 0 <= py < H:
                    ap = active(py, px)
                    lp = np.where(ap, Ld[py, px], 1 << 20)
                    mn = lp[ap].min()
                    m = np.minimum(lp[None, :] + pen, 255).min(axis=1)
                    Ld[y, x] = np.where(a, C[y, x] + m - mn, 0)
                else:
                    Ld[y, x] = np.where(a, C[y, x], 0)
        S += Ld
    out = np.empty((H, W), np.uint8)
    for y in range(H):
        for x in range(W):
            a = active(y, x)
            out[y, x] = np.flatnonzero(a)[np.argmin(S[y, x][a])]
    return out


@pytest.mark.parametrize("seed,paths", [(0, 4), (1, 4), (2, 8)])
def test_window_oracle_matches_transition_form(seed, paths):
    rng = np.random.default_rng(seed)
    H, W, D = 21, 37, 64
    L, R, _ = sgm_amd.scene_planes(W, H, 40, seed + 3, planes=(2, 4))
    b0 = rng.integers(0, 3, (2, 3))
    b1 = np.minimum(b0 + rng.integers(0, 3, b0.shape), 3)
    blk = (b0 | (b1 << 4)).astype(np.uint8)
    np.testing.assert_array_equal(oracle.sgm_window(L, R, blk, D, paths), window_bruteforce(L, R, blk, D, paths))


def test_window_oracle_full_range_is_plain():
    L, R = rng_image(40, 70, 11), rng_image(40, 70, 12)
    for D, paths in ((32, 4), (64, 8)):
        blk = np.full((3, 5), (D // 16 - 1) << 4, np.uint8)
        np.testing.assert_array_equal(oracle.sgm_window(L, R, blk, D, paths), oracle.sgm_plain(L, R, D, paths))


def test_argmin_lowest_d_wins_ties():
    S = np.full((1, 3, 8), 100, np.uint16)
    S[0, 0, [2, 5]] = 7
    S[0, 1, [7]] = 1
    S[0, 2, :] = 9
    np.testing.assert_array_equal(oracle.argmin(S), [[2, 7, 0]])


def test_plain_scalar_equals_omp():
    L, R = rng_image(50, 90, 1), rng_image(50, 90, 2)
    for paths in (4, 8):
        np.testing.assert_array_equal(oracle.sgm_plain(L, R, 32, paths, threads=1),
                                      oracle.sgm_plain(L, R, 32, paths, threads=4))


def test_plain_single_direction_consistency():
    """With one direction the disparity is the argmin of that direction's L."""
    L, R = rng_image(20, 70, 3), rng_image(20, 70, 4)
    C = oracle.cost(oracle.census(L), oracle.census(R), 64)
    Lv = oracle.path_L(C, 1, 0)
    np.testing.assert_array_equal(oracle.argmin(Lv.astype(np.uint16)), np.argmin(Lv, axis=2))


# ---------------------------------------------------------------- ms pieces
def test_sobel_and_half_closed_forms():
    img = rng_image(11, 17, 9).astype(np.int64)
    p = np.pad(img, 1, mode="edge")
    gx = (p[:-2, 2:] - p[:-2, :-2]) + 2 * (p[1:-1, 2:] - p[1:-1, :-2]) + (p[2:, 2:] - p[2:, :-2])
    want = np.clip(128 + (gx >> 3), 0, 255)   # >> on negative int64 is arithmetic
    np.testing.assert_array_equal(oracle.sobelx(img.astype(np.uint8)), want)
    img = rng_image(10, 16, 8).astype(np.int64)
    want = (img[0::2, 0::2] + img[0::2, 1::2] + img[1::2, 0::2] + img[1::2, 1::2] + 2) >> 2
    np.testing.assert_array_equal(oracle.half(img.astype(np.uint8)), want)


# ---------------------------------------------------------------- ms composition
# multiscale/sgm_ms_ref.c:85-87: dx, dy and the direction's P1
MS_DIRS = [(1, 0, 40), (-1, 0, 40), (0, 1, 20), (0, -1, 20), (1, 1, 10), (-1, -1, 10), (1, -1, 10), (-1, 1, 10)]


def sobel_closed(img):
    p = np.pad(img.astype(np.int64), 1, mode="edge")
    gx = (p[:-2, 2:] - p[:-2, :-2]) + 2 * (p[1:-1, 2:] - p[1:-1, :-2]) + (p[2:, 2:] - p[2:, :-2])
    return np.clip(128 + (gx >> 3), 0, 255).astype(np.uint8)


def half_closed(img):
    a = img.astype(np.int64)
    return ((a[0::2, 0::2] + a[0::2, 1::2] + a[1::2, 0::2] + a[1::2, 1::2] + 2) >> 2).astype(np.uint8)


def cost_closed(cl, cr, D, invalid):
    H, W = cl.shape
    C = np.full((H, W, D), invalid, np.uint8)
    for d in range(min(D, W)):
        x = cl[:, d:] ^ cr[:, :W - d]
        C[:, d:, d] = np.array([bin(int(v)).count("1") for v in x.ravel()], np.uint8).reshape(x.shape)
    return C


def ms_bruteforce(L, R, D):
    """The whole multiscale composition from already-pinned pieces and the
    reference's stage order (multiscale/sgm_ms_ref.c:166-199): half-resolution
    SGM on half() of the raw images, Sobel-X then census at each scale, cost
    with invalid 63, the full-resolution cost averaged with the half-resolution
    cost at (y/2, x/2, d/2), 8 saturating weighted paths (transition form),
    argmin of stage 1 over every half-resolution pixel and of stage 2 at
    (2y, 2x)."""
    def stage_cost(a, b):
        return cost_closed(census_bruteforce(sobel_closed(a)), census_bruteforce(sobel_closed(b)), D, 63)

    def S_of(C):
        return sum(path_bruteforce(C, dx, dy, p1, 200, True).astype(np.int64) for dx, dy, p1 in MS_DIRS)

    C1 = stage_cost(half_closed(L), half_closed(R))
    disp1 = np.argmin(S_of(C1), axis=2).astype(np.uint8)   # first minimum: the lowest d
    C2 = stage_cost(L, R).astype(np.int64)
    H, W = L.shape
    ys, xs = np.arange(H)[:, None], np.arange(W)[None, :]
    C2 = ((C2 + C1[ys // 2, xs // 2][:, :, np.arange(D) // 2].astype(np.int64) + 1) >> 1).astype(np.uint8)
    disp = np.argmin(S_of(C2)[0::2, 0::2], axis=2).astype(np.uint8)
    return disp, disp1


@pytest.mark.parametrize("W,H,D", [(34, 22, 16), (34, 22, 32), (34, 22, 64), (34, 22, 128), (40, 64, 32),
                                   (40, 64, 64), (140, 18, 128), (36, 16, 256)])
def test_ms_oracle_matches_bruteforce_composition(W, H, D):
    """Pins orc_sgm_ms at every D the engine accepts, not only D=64 (the
    reference's multiscale oracle hard-codes D=64, multiscale/sgm_ms_ref.c:38)."""
    L, R, _ = sgm_amd.scene_planes(W, H, min(D - 1, W - 1), W * 7 + H + D, planes=(2, 4))
    want, want1 = ms_bruteforce(L, R, D)
    got, got1 = oracle.sgm_ms(L, R, D, want_stage1=True)
    np.testing.assert_array_equal(got1, want1)
    np.testing.assert_array_equal(got, want)


def test_ms_bruteforce_catches_a_merge_index_mistake():
    """The brute force is sharp enough to see a D-dependent merge mistake: the
    same composition merging c1[d] instead of c1[d/2] differs from it."""
    W, H, D = 40, 24, 32
    L, R, _ = sgm_amd.scene_planes(W, H, D - 1, 5, planes=(2, 4))
    want, _ = ms_bruteforce(L, R, D)
    C1 = cost_closed(census_bruteforce(sobel_closed(half_closed(L))), census_bruteforce(sobel_closed(half_closed(R))), D, 63)
    C2 = cost_closed(census_bruteforce(sobel_closed(L)), census_bruteforce(sobel_closed(R)), D, 63).astype(np.int64)
    ys, xs = np.arange(H)[:, None], np.arange(W)[None, :]
    bad = ((C2 + C1[ys // 2, xs // 2].astype(np.int64) + 1) >> 1).astype(np.uint8)
    S = sum(path_bruteforce(bad, dx, dy, p1, 200, True).astype(np.int64) for dx, dy, p1 in MS_DIRS)
    assert not np.array_equal(np.argmin(S[0::2, 0::2], axis=2).astype(np.uint8), want)


# ---------------------------------------------------------------- goldens
def test_generator_matches_reference_inputs(goldens):
    import hashlib
    for rec in goldens["reference_inputs"]:
        sp = rec["scene"]
        L, R, G = sgm_amd.scene_reference(sp["W"], sp["H"], sp["gen_D"], sp["seed"])
        hdr = b"P5\n%d %d\n255\n" % (sp["W"], sp["H"])
        for name, arr in (("left", L), ("right", R), ("gt", G)):
            assert hashlib.sha256(hdr + arr.tobytes()).hexdigest() == rec[name + "_sha256"], (sp, name)


@pytest.mark.slow
def test_oracle_reproduces_reference_goldens(goldens):
    for g in goldens["reference_goldens"]:
        L, R = make_scene(g["scene"])
        if g["mode"] == "plain":
            out = oracle.sgm_plain(L, R, g["D"], g["paths"])
        else:
            out = oracle.sgm_ms(L, R, g["D"])
        assert list(out.shape) == g["shape"]
        assert oracle.fnv1a64(out) == g["fnv"], g["name"]


def test_oracle_matches_compiled_reference(goldens):
    for g in goldens["computed"]:
        sp = g["scene"]
        if sp["W"] * sp["H"] * g["D"] > 640 * 480 * 64 * (2 if g["mode"] == "plain" else 1):
            continue  # full-size cases: test_oracle_matches_compiled_reference_full
        L, R = make_scene(sp)
        assert sgm_amd.fnv1a64(np.concatenate([L.ravel(), R.ravel()])) == g["input_fnv"], g["name"]
        out = oracle.sgm_plain(L, R, g["D"], g["paths"]) if g["mode"] == "plain" else oracle.sgm_ms(L, R, g["D"])
        assert oracle.fnv1a64(out) == g["fnv"], g["name"]


@pytest.mark.slow
def test_oracle_matches_compiled_reference_full(goldens):
    for g in goldens["computed"]:
        sp = g["scene"]
        if sp["W"] * sp["H"] * g["D"] <= 640 * 480 * 64 * (2 if g["mode"] == "plain" else 1):
            continue
        L, R = make_scene(sp)
        out = oracle.sgm_plain(L, R, g["D"], g["paths"]) if g["mode"] == "plain" else oracle.sgm_ms(L, R, g["D"])
        assert oracle.fnv1a64(out) == g["fnv"], g["name"]


def test_golden_discriminates(goldens):
    """The discrimination rule of the reference's checker
    (scripts/check_golden_discriminates.py:64-119): max disparity D-1 present,
    >= 8 distinct values, >= 0.5% of pixels in the top decile."""
    from sgm_amd.goldencheck import discriminates
    g = next(x for x in goldens["computed"] if x["name"] == "P_640x480_d64")
    L, R = make_scene(g["scene"])
    ok, stats = discriminates(oracle.sgm_plain(L, R, 64, 4), 64)
    assert ok, stats
    # a truncated search (d < 48) must fail it
    ok_trunc, _ = discriminates(np.minimum(oracle.sgm_plain(L, R, 64, 4), 47), 64)
    assert not ok_trunc


def _all_fixtures(goldens):
    for sec in ("reference_goldens", "computed", "ties", "batch"):
        for r in goldens[sec]:
            yield sec, r


def test_every_gate_fixture_discriminates(goldens):
    """Every golden and fixture that gates a cell passes the reference checker's
    rule (scripts/check_golden_discriminates.py:64-119, scripts/sweep.sh:37-38);
    the others are recorded with their role and why they are not gates."""
    from sgm_amd.goldencheck import discriminates  # noqa: F401
    gates = 0
    for sec, r in _all_fixtures(goldens):
        gc = r["goldencheck"]
        assert gc["role"] in ("gate", "shape", "ties"), r["name"]
        if gc["role"] == "gate":
            gates += 1
            assert gc["ok"], (sec, r["name"], gc)
            # the rule it passed: a search truncated below 0.9*D leaves no pixel
            # in the top decile, so it cannot reproduce this golden
            assert gc["max"] == r["D"] - 1 and gc["top_decile_pct"] >= 0.5, r["name"]
        elif gc["role"] == "shape":
            assert gc.get("why"), r["name"]
    assert gates >= 50


def test_fixture_discrimination_stats_recomputed(goldens):
    """The stored statistics are those of the map itself: recomputed from the
    oracle's output (which the fixture hash pins) for every fixture small
    enough to run here, and a truncated copy of each gate fails the rule."""
    from sgm_amd.goldencheck import discriminates
    n = 0
    for sec, r in _all_fixtures(goldens):
        sp = r["scene"]
        if sec == "reference_goldens" or sp["W"] * sp["H"] * r["D"] > 640 * 480 * 128:
            continue
        L, R = make_scene(sp)
        out = oracle.sgm_plain(L, R, r["D"], r["paths"]) if r["mode"] == "plain" else oracle.sgm_ms(L, R, r["D"])
        assert oracle.fnv1a64(out) == r["fnv"], r["name"]
        ok, st = discriminates(out, r["D"])
        gc = r["goldencheck"]
        assert ok == gc["ok"] and st["max"] == gc["max"] and st["distinct"] == gc["distinct"], r["name"]
        assert st["top_decile_px"] == gc["top_decile_px"], r["name"]
        if gc["role"] == "gate":
            trunc = np.minimum(out, int(0.9 * r["D"]) - 1)
            assert not discriminates(trunc, r["D"])[0], r["name"]
        n += 1
    assert n >= 15


def test_tie_fixtures(goldens):
    """Tie-rich scenes: the oracle's output hash and its count of tied-minimum
    pixels equal what the reference's own SGM_COUNT_TIES build reported
    (common/sgm_ref.c:125-138), and at least 1 % of pixels are tied."""
    seen = 0
    for r in goldens["ties"]:
        assert r["tie_pct"] >= 1.0, r["name"]
        sp = r["scene"]
        if sp["W"] * sp["H"] > 640 * 480:
            continue
        L, R = make_scene(sp)
        assert sgm_amd.fnv1a64(np.concatenate([L.ravel(), R.ravel()])) == r["input_fnv"], r["name"]
        S = oracle.path_sums(L, R, r["D"], r["paths"])
        assert oracle.tie_pixels(S) == r["tie_pixels"], r["name"]
        disp = oracle.sgm_plain(L, R, r["D"], r["paths"])
        np.testing.assert_array_equal(disp, np.argmin(S, axis=2))  # first minimum: lowest d
        assert oracle.fnv1a64(disp) == r["fnv"], r["name"]
        # the lowest-d rule is what the hash pins: taking the HIGHEST tied d
        # changes it on these scenes
        hi = (r["D"] - 1 - np.argmin(S[:, :, ::-1], axis=2)).astype(np.uint8)
        assert oracle.fnv1a64(hi) != r["fnv"], r["name"]
        seen += 1
    assert seen >= 2


def test_tie_count_closed_forms():
    S = np.array([[[3, 1, 1, 5], [2, 3, 4, 5], [7, 7, 7, 7]]], np.uint16)
    assert oracle.tie_pixels(S) == 2
