"""Restatement of the reference's temporal-block program for the checker.

TEST INFRASTRUCTURE ONLY.  Follows streamGPU-AS/Render.cpp:198-277
(block intervals), :348-468 (prepare_data_seeds: seed budget, kept ++ new
seeds, block2all maps), :470-504 (block_segment: a plain run from those
seeds) and :507-569 (post_and_next: alive centroids, kept seeds with the z
shift, label stitching).  Each block's run is the C oracle's
orc_run_seeded; check_dead_seeds is oracle/post.py.  Grid placement within
a block follows streamGPU-AS/Image.cpp:26-119 (the plain rule).
"""
import ctypes

import numpy as np

import oracle as orc
import post


def plan_blocks(D, B, O):
    """streamGPU-AS/Render.cpp:198-241 with block_size B, overlap O."""
    S = B - O
    if D < B + O:
        return [(0, D)]
    out = [(0, B)]
    last_end = B
    done = False
    while last_end + S < D:
        if D - last_end - S < O:
            out.append((last_end - O, D))
            done = True
        else:
            out.append((last_end - O, last_end + S))
        last_end += S
    if not done:
        out.append((last_end - O, D))
    return out


def run_seeded(rgb, fwd, bwd, xyz, s, iters, compactness, angle_gate, tie_rule=0, prepared=0):
    lib = orc.load()
    lib.orc_run_seeded.restype = ctypes.c_int
    lib.orc_run_seeded.argtypes = [ctypes.c_void_p] * 5 + [ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 3
    D, H, W, _ = rgb.shape
    A = len(xyz)
    P = orc.OrcParams(W, H, D, A, iters, compactness, tie_rule, angle_gate, 0, prepared)
    labels = np.zeros((D, H, W), np.int32)
    sums = np.zeros((A + 1, 7), np.int64)
    xyz = np.ascontiguousarray(xyz, np.int32)
    r = lib.orc_run_seeded(ctypes.byref(P), orc._p(np.ascontiguousarray(rgb)), orc._p(np.ascontiguousarray(fwd)),
                           orc._p(np.ascontiguousarray(bwd)), orc._p(xyz), A, s, orc._p(labels), orc._p(sums), None)
    if r < 0:
        raise ValueError(f"orc_run_seeded failed: {r}")
    return labels, sums


def run_video(rgb, fwd, bwd, K, block_size=30, overlap=10, iters=20, compactness=6.0, angle_gate=0,
              seeds_per_100=None):
    """Whole video; returns (labels [D,H,W] global ids, GL rows; seeds placed)."""
    D, H, W, _ = rgb.shape
    B, O = block_size, overlap
    S = B - O
    blocks = plan_blocks(D, B, O)
    seeds100 = seeds_per_100 if seeds_per_100 else K * 100 // D
    labels_all = np.full((D, H, W), -7, np.int32)
    placed = 0
    kept = []            # [(x, y, z centre coords, old id)]
    prev_map = {}
    for b, (start, end) in enumerate(blocks):
        Db = end - start
        last = b == len(blocks) - 1
        budget = seeds100 * Db // 100 if (b == 0 and len(blocks) == 1) else seeds100 * (Db - O) // 100
        budget = max(budget, 1)
        new_xyz, s = orc.grid_seeds(W, H, Db, budget)
        block2all = {}
        xyz = []
        for i, (x, y, z, old) in enumerate(kept):
            xyz.append((int(np.floor(x)), int(np.floor(y)), int(np.floor(z))))
            block2all[i] = i if b == 0 else prev_map.setdefault(old, 0)
        added = placed
        for j, p in enumerate(new_xyz):
            xyz.append(tuple(int(v) for v in p))
            block2all[len(kept) + j] = j if b == 0 else added
            if b != 0:
                added += 1
        A = len(xyz)
        labels, sums = run_seeded(rgb[start:end], fwd[start:end], bwd[start:end], np.array(xyz), max(s, 1), iters,
                                  compactness, angle_gate)
        counts = list(sums[:, 3])
        flat, dead, nfin = post.check_dead_seeds(counts, A, labels.reshape(-1).tolist())
        labels = np.array(flat, np.int32).reshape(labels.shape)
        alive = []
        for i in range(A):
            n = int(sums[i, 3])
            if n < 1:
                continue
            alive.append((float(sums[i, 0] // (2 * n)) + 0.5, float(sums[i, 1] // (2 * n)) + 0.5,
                          float(sums[i, 2] // (2 * n)) + 0.5, i))
        placed += len(alive) - len(kept)
        nxt = []
        done = Db
        if not last:
            for (x, y, z, i) in alive:
                if z > B - O:
                    nz = z - S
                    if nz < 1:
                        nz = 1.5
                    nxt.append((x, y, nz, i))
            done = Db - O
        for f in range(done):
            labels_all[start + f] = np.vectorize(lambda l: block2all.get(int(l), 0), otypes=[np.int32])(labels[f])
        kept = nxt
        prev_map = block2all
    return labels_all, placed
