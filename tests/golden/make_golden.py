"""Writes tests/golden/golden.json from the CPU oracle and the generator.

The reference ships no golden vectors for this path; these fixtures pin the
build's own specification over time (LAB samples, generator output hashes,
small-run label and sum hashes), not parity with the reference.
Run:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import anisotropic_supervoxel_amd as asv  # noqa: E402
import oracle as orc  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    out = {}
    rng = np.random.default_rng(2024)
    rgb = rng.integers(0, 256, (64, 3))
    out["lab_samples"] = [[int(r), int(g), int(b), *orc.lab_of(r, g, b)] for r, g, b in rgb]
    synth = {}
    for (seed, W, H, D, E, vm, sg) in [(0, 160, 120, 30, 12, 2.0, 4.0), (5, 48, 40, 8, 4, 4.0, 6.0)]:
        v = asv.synth_video(seed, W, H, D, E, vm, sg)
        synth[f"{seed}-{W}x{H}x{D}-{E}-{vm}-{sg}"] = [sha(x) for x in v]
    out["synth_sha256"] = synth
    runs = {}
    for (seed, W, H, D, E, K, it) in [(11, 40, 32, 10, 5, 60, 4), (0, 160, 120, 30, 12, 500, 10)]:
        rgb, fwd, bwd, _ = asv.synth_video(seed, W, H, D, E, 2.0, 4.0)
        labels, sums, _, A = orc.run(rgb, fwd, bwd, K, it)
        runs[f"{seed}-{W}x{H}x{D}-K{K}-it{it}"] = {"A": A, "labels_sha256": sha(labels),
                                                    "sums_sha256": sha(sums),
                                                    "n_unlabelled": int((labels < 0).sum()),
                                                    "n_alive": int((sums[:A, 3] > 0).sum())}
    out["oracle_runs"] = runs
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote golden.json")


if __name__ == "__main__":
    main()
