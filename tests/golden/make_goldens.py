"""make_goldens.py -- write tests/golden/goldens.json, the hash fixtures that pin
this repository's oracle and engine to the reference.

Run in a container that has the reference at /root/reference (or
$SGM_REFERENCE) after `oracle/build_ref.sh` built its binaries into
oracle/_ref/.  Three kinds of record:

  reference_goldens   FNV-1a 64 of the reference's own on-disk golden maps
                      (data/golden, data/shared, data/multiscale), with the
                      generator call that regenerates their input pair;
  reference_inputs    sha256 of the reference generator's left/right PGM files
                      for those calls (pins the generator restatement);
  computed            FNV of the COMPILED REFERENCE oracle (sgm_ref /
                      sgm_ms_ref built from its sources) on scenes from this
                      repository's generators, at sizes and D the reference
                      ships no golden for.  The tie-rich scenes also carry the
                      reference's own count of tied-minimum pixels (its
                      SGM_COUNT_TIES build, common/sgm_ref.c:125-138).
  batch               the 24 pairs of the batched workload.

Every golden and fixture also records the discrimination statistics of the
reference's checker (scripts/check_golden_discriminates.py:64-119) and
whether its scene can fail a truncated search at all (a scene whose
disparities stop below D-1, or a frame narrower than D, cannot).

Every input is regenerated from its recorded spec; only hashes are stored.
"""
import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "sgm-bench_amd"))
import sgm_amd  # noqa: E402
from sgm_amd.goldencheck import discriminates  # noqa: E402

REF = os.environ.get("SGM_REFERENCE", "/root/reference")
BIN = os.path.join(ROOT, "oracle", "_ref")


def fnv(a):
    return sgm_amd.fnv1a64(np.ascontiguousarray(a, np.uint8))


def run_plain(L, R, D, paths, tmp, ties=False):
    """The compiled reference's disparity map (and, with ties, its tie count)."""
    exe = os.path.join(BIN, "sgm_ref_d%d%s" % (D, "_ties" if ties else ("_p8" if paths == 8 else "")))
    sgm_amd.pgm_write(os.path.join(tmp, "l.pgm"), L)
    sgm_amd.pgm_write(os.path.join(tmp, "r.pgm"), R)
    out = os.path.join(tmp, "o.pgm")
    r = subprocess.run([exe, os.path.join(tmp, "l.pgm"), os.path.join(tmp, "r.pgm"), "-w", "0", "-n", "1",
                        "-o", out, "--no-roofline"], check=True, capture_output=True, cwd=tmp, text=True)
    disp = sgm_amd.pgm_read(out)
    if not ties:
        return disp
    m = re.search(r"TIES: (\d+) of (\d+) pixels", r.stderr)
    return disp, int(m.group(1))


def run_ms(L, R, tmp):
    exe = os.path.join(BIN, "sgm_ms_ref_omp")
    sgm_amd.pgm_write(os.path.join(tmp, "l.pgm"), L)
    sgm_amd.pgm_write(os.path.join(tmp, "r.pgm"), R)
    out = os.path.join(tmp, "o.pgm")
    subprocess.run([exe, os.path.join(tmp, "l.pgm"), os.path.join(tmp, "r.pgm"), "-o", out, "-n", "1"],
                   check=True, capture_output=True, cwd=tmp)
    return sgm_amd.pgm_read(out)


# Fixtures that pin something other than the top of the disparity range, and
# so are not truncation gates: tiny, ragged and extreme shapes (edge handling)
# and the 640x480 D=256 reference-generator cell, whose slabs stop at d=219
# (d256_640x480 gates that size instead).  Every other golden and fixture is
# a gate and must pass the discrimination check, as every cell of the
# reference's sweep must (scripts/sweep.sh:37-38).
NOT_GATES = {
    "odd_97x61_d64": "ragged 97x61 shape: edge handling",
    "odd_97x61_d64_p8": "ragged 97x61 shape, 8 paths: edge handling",
    "tiny_16x16_d16": "smallest frame the engine takes",
    "tiny_33x17_d32": "tiny ragged frame",
    "narrow_40x300_d64": "frame narrower than D",
    "ms_odd_98x62": "ragged multiscale frame",
    "grid_640x480_d256": "the reference generator's slabs stop at d=219 at this size",
}


def goldencheck(disp, D, name, role="gate"):
    """Discrimination statistics (sgm_amd.goldencheck restates the reference
    checker's rule), the fixture's role, and why a non-gate is not one."""
    ok, st = discriminates(disp, D)
    rec = dict(st, ok=bool(ok), role=role)
    if name in NOT_GATES:
        rec["role"], rec["why"] = "shape", NOT_GATES[name]
    return rec


def scene(spec):
    return sgm_amd.make_scene(spec)


def ref_scene(W, H, D, seed):
    return {"kind": "reference", "W": W, "H": H, "gen_D": D, "seed": seed}


def periodic(W, H, dmax, period, seed):
    return {"kind": "periodic", "W": W, "H": H, "dmax": dmax, "period": period, "seed": seed}


def planes(W, H, dmax, seed, planes=(8, 12), occl=0, noise=2, slanted=0):
    return {"kind": "planes", "W": W, "H": H, "dmax": dmax, "seed": seed, "planes": list(planes),
            "occl": occl, "noise": noise, "slanted": slanted}


def main():
    if not os.path.isdir(BIN):
        sys.exit("oracle/_ref missing: run oracle/build_ref.sh first")
    out = {"reference_goldens": [], "reference_inputs": [], "computed": []}
    # on-disk reference goldens and the generator calls that make their inputs
    disk = [
        ("A_synthetic", ref_scene(1920, 1080, 64, 1), "plain", 64, 4, "data/golden/synthetic.pgm"),
        ("shared_d64", ref_scene(1920, 1080, 256, 3), "plain", 64, 4, "data/shared/golden_d64.pgm"),
        ("shared_d128", ref_scene(1920, 1080, 256, 3), "plain", 128, 4, "data/shared/golden_d128.pgm"),
        ("shared_d256", ref_scene(1920, 1080, 256, 3), "plain", 256, 4, "data/shared/golden_d256.pgm"),
        ("B_multiscale", ref_scene(1920, 800, 64, 1), "ms", 64, 8, "data/multiscale/golden_ms.pgm"),
        ("C_multiscale", ref_scene(1920, 1080, 64, 1), "ms", 64, 8, "data/multiscale/golden_ms_1080.pgm"),
    ]
    for name, spec, kind, D, paths, path in disk:
        g = sgm_amd.pgm_read(os.path.join(REF, path))
        out["reference_goldens"].append({"name": name, "scene": spec, "mode": kind, "D": D,
                                         "paths": paths, "shape": list(g.shape), "fnv": fnv(g),
                                         "source": path, "goldencheck": goldencheck(g, D, name)})
    with tempfile.TemporaryDirectory() as tmp:
        for (W, H, D, seed) in [(1920, 1080, 64, 1), (1920, 1080, 256, 3), (1920, 800, 64, 1)]:
            subprocess.run([os.path.join(BIN, "gen_synthetic_d%d" % D), str(W), str(H), tmp, str(seed)],
                           check=True, capture_output=True)
            rec = {"scene": ref_scene(W, H, D, seed)}
            for n in ("left", "right", "gt"):
                rec[n + "_sha256"] = hashlib.sha256(open(os.path.join(tmp, n + ".pgm"), "rb").read()).hexdigest()
            out["reference_inputs"].append(rec)

        computed = [
            # plain, the configurations and the batch's plain members
            ("P_640x480_d64", planes(640, 480, 63, 11), "plain", 64, 4),
            ("A_1920x1080_d64", planes(1920, 1080, 63, 21, (20, 20), 50), "plain", 64, 4),
            ("A_1920x1080_d128", planes(1920, 1080, 127, 22, (20, 20), 50), "plain", 128, 4),
            ("batch_1280x720_d64", planes(1280, 720, 63, 101, (8, 12)), "plain", 64, 4),
            ("batch_1920x1080_d128", planes(1920, 1080, 127, 103, (10, 30), 50), "plain", 128, 4),
            ("learn_1920x1080_d128", planes(1920, 1080, 127, 7001, (10, 30), 0, 2, 1), "plain", 128, 4),
            # the reference generator at other sizes / D
            ("ref_640x480_d64", ref_scene(640, 480, 64, 1), "plain", 64, 4),
            ("ref_640x480_d128", ref_scene(640, 480, 128, 1), "plain", 128, 4),
            ("ref_1280x720_d64", ref_scene(1280, 720, 64, 1), "plain", 64, 4),
            ("ref_960x540_d128", ref_scene(960, 540, 128, 1), "plain", 128, 4),
            # D sweep over resolutions (the reference generator at each D): with the
            # cells above and the on-disk goldens, every D of 32..256 at 640x480,
            # 1280x720 and 1920x1080
            ("grid_640x480_d32", ref_scene(640, 480, 32, 1), "plain", 32, 4),
            ("grid_640x480_d256", ref_scene(640, 480, 256, 1), "plain", 256, 4),
            ("grid_1280x720_d32", ref_scene(1280, 720, 32, 1), "plain", 32, 4),
            ("grid_1280x720_d128", ref_scene(1280, 720, 128, 1), "plain", 128, 4),
            ("grid_1280x720_d256", ref_scene(1280, 720, 256, 1), "plain", 256, 4),
            ("grid_1920x1080_d32", ref_scene(1920, 1080, 32, 1), "plain", 32, 4),
            ("grid_1920x1080_d128", ref_scene(1920, 1080, 128, 1), "plain", 128, 4),
            ("grid_960x540_d32", ref_scene(960, 540, 32, 1), "plain", 32, 4),
            ("grid_960x540_d64", ref_scene(960, 540, 64, 1), "plain", 64, 4),
            # 8 paths (SGM_PATHS=8 in the reference oracle)
            ("p8_640x480_d64", planes(640, 480, 63, 12), "plain", 64, 8),
            ("p8_ref_640x480_d64", ref_scene(640, 480, 64, 1), "plain", 64, 8),
            ("p8_1920x1080_d64", planes(1920, 1080, 63, 23, (20, 20), 50), "plain", 64, 8),
            # small, ragged and extreme shapes
            ("odd_97x61_d64", planes(97, 61, 63, 5, (3, 5)), "plain", 64, 4),
            ("odd_97x61_d64_p8", planes(97, 61, 63, 5, (3, 5)), "plain", 64, 8),
            ("tiny_16x16_d16", planes(16, 16, 15, 6, (1, 2)), "plain", 16, 4),
            ("tiny_33x17_d32", planes(33, 17, 31, 8, (1, 3)), "plain", 32, 4),
            ("narrow_40x300_d64", planes(40, 300, 39, 9, (2, 4)), "plain", 64, 4),
            ("wide_1000x20_d256", planes(1000, 20, 255, 10, (4, 6)), "plain", 256, 4),
            ("d256_640x480", planes(640, 480, 255, 13, (8, 12)), "plain", 256, 4),
            # multiscale (the reference oracle is D=64 only)
            ("B_1920x800_ms", planes(1920, 800, 63, 31, (20, 20), 50), "ms", 64, 8),
            ("C_1920x1080_ms", planes(1920, 1080, 63, 41, (20, 20), 50), "ms", 64, 8),
            ("ms_ref_640x480", ref_scene(640, 480, 64, 1), "ms", 64, 8),
            ("ms_odd_98x62", planes(98, 62, 63, 14, (3, 5)), "ms", 64, 8),
        ]
        for name, spec, kind, D, paths in computed:
            L, R = scene(spec)
            disp = run_plain(L, R, D, paths, tmp) if kind == "plain" else run_ms(L, R, tmp)
            out["computed"].append({"name": name, "scene": spec, "mode": kind, "D": D, "paths": paths,
                                    "input_fnv": fnv(np.concatenate([L.ravel(), R.ravel()])),
                                    "fnv": fnv(disp), "goldencheck": goldencheck(disp, D, name)})
            print(name, fnv(disp), out["computed"][-1]["goldencheck"], flush=True)
        # tie-rich scenes: the reference's tie-counting build, cross-checked
        # against the oracle's count over its own path sums
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        out["ties"] = []
        for name, spec, D in [("ties_640x480_d64_p16", periodic(640, 480, 63, 16, 1), 64),
                              ("ties_640x480_d128_p8", periodic(640, 480, 127, 8, 2), 128),
                              ("ties_1280x720_d64_p24", periodic(1280, 720, 63, 24, 3), 64)]:
            L, R = scene(spec)
            disp, ties = run_plain(L, R, D, 4, tmp, ties=True)
            S = oracle.path_sums(L, R, D, 4)
            assert oracle.tie_pixels(S) == ties, (name, oracle.tie_pixels(S), ties)
            assert np.array_equal(np.argmin(S, axis=2).astype(np.uint8), disp), name
            out["ties"].append({"name": name, "scene": spec, "mode": "plain", "D": D, "paths": 4,
                                "input_fnv": fnv(np.concatenate([L.ravel(), R.ravel()])), "fnv": fnv(disp),
                                "tie_pixels": ties, "tie_pct": round(100.0 * ties / L.size, 3),
                                "goldencheck": goldencheck(disp, D, name, role="ties")})
            print(name, fnv(disp), "ties", ties, flush=True)
        # the batched workload: compiled reference for the plain pairs; the
        # oracle restatement for the D=128 multiscale pairs (the reference's
        # multiscale oracle is D=64 only, multiscale/sgm_ms_ref.c:38)
        out["batch"] = []
        for p in sgm_amd.batch_spec():
            L, R = scene(p["scene"])
            if p["mode"] == sgm_amd.PLAIN:
                disp, src = run_plain(L, R, p["D"], p["paths"], tmp), "reference"
            else:
                disp, src = oracle.sgm_ms(L, R, p["D"]), "oracle (no D=128 reference; pinned by brute force)"
            h = fnv(disp)
            out["batch"].append({"name": p["name"], "scene": p["scene"], "mode": "plain" if p["mode"] == 0 else "ms",
                                 "D": p["D"], "paths": p["paths"], "fnv": h, "source": src,
                                 "goldencheck": goldencheck(disp, p["D"], p["name"])})
            print(p["name"], h, src, flush=True)
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "goldens.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
