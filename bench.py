"""Benchmark: plain JFA + Lloyd supervoxel run on one MI355X.

One step = one whole plain run of the hot path over the resident clip:
LAB construction, then `iters` Lloyd iterations (seed stage with the
per-voxel metric evaluated at the seeds and their rings, P JFA passes,
reduction, finalisation), ending with the label volume in HBM.
Default workload: 640x480x200 synthetic video (seed 2, 32 ellipses,
|v| <= 4 px/frame, noise sigma 6), K = 10000 (9720 placed), 10 iterations,
reference JFA schedule 1,9,5,3,2,1,2,1 (P = 8), FP32 metric, int32 labels.

Prints one JSON line.  `value` is Mvoxel-iters/s = V * iters * steps / time.
`roofline` describes the dominant kernel (one full JFA pass, the marching
kernel; an iteration's first pass, and its second when sparse, evaluate only
the seeds' neighbourhoods and are timed separately): 12 algorithmic bytes per
voxel (int32 label in, packed LAB in, int32 label out) over the pass's
average launch time from HIP events on the launch stream.
`quality` is the undersegmentation error and boundary recall of the final
labels against the generator's ground-truth regions (outside the timing).
`cpu_baseline` times the CPU oracle on a bounded sample of the same video.
`served` is the streaming form: the plain per-window loop against the
served loop with continuous batching over concurrent streams, aggregate
Mvoxel-iters/s and p50/p99 window latency for each (see served()).
`batched` is 64 clips of uneven sizes run together against the plain
per-clip loop on the same resident contexts (see batched()).
"""
import argparse
import json
import os
import sys
import time

# Eight hardware queues, so that the served leg's eight worker streams do
# not share queues (HIP's default is four).  Must precede HIP start-up.
os.environ.setdefault("GPU_MAX_HW_QUEUES", "8")

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
WORKLOADS = {
    # name: (seed, W, H, D, ellipses, vmax, sigma, K, iters)
    "plain-640x480x200-K10000": (2, 640, 480, 200, 32, 4.0, 6.0, 10000, 10),
    "plain-320x240x85-K2000": (1, 320, 240, 85, 16, 2.0, 4.0, 2000, 10),
    "plain-160x120x30-K500": (0, 160, 120, 30, 12, 2.0, 4.0, 500, 10),
}


def load_traffic(workload):
    """HBM bytes per JFA launch measured by rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, scaled by a copy-kernel calibration), if committed."""
    path = os.path.join(ROOT, "profiles", "jfa_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get(workload, {}).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_baseline(seed, W, H, D, ell, vmax, sigma, K, iters, frames):
    """Oracle on the first `frames` frames of the same video, K scaled to
    keep the seed density (same grid step s and schedule)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import anisotropic_supervoxel_amd as asv
    import oracle as orc
    rgb, fwd, bwd, _ = asv.synth_video(seed, W, H, D, ell, vmax, sigma)
    rgb = np.ascontiguousarray(rgb[:frames])
    fwd = np.ascontiguousarray(fwd[:frames])
    bwd = np.ascontiguousarray(bwd[:frames])
    Ks = max(1, K * frames // D)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    orc.load()
    t0 = time.perf_counter()
    orc.run(rgb, fwd, bwd, Ks, iters)
    dt = time.perf_counter() - t0
    vox_iters = W * H * frames * iters
    return {"value": vox_iters / dt / 1e6, "unit": "Mvoxel-iters/s", "cores": threads, "kind": "port",
            "sample": f"CPU oracle (C, OpenMP) on frames 0..{frames - 1} of the same video "
                      f"({W}x{H}x{frames}, K={Ks}, {iters} iterations, same step s and schedule), "
                      f"{dt:.1f} s wall"}


def served(n_streams, actives, seed):
    """Served streaming workload: n_streams requests with Poisson arrivals at
    40/s (resolutions 320x240 / 640x480 / 1280x720, 60-300 frames, density
    1-4 supervoxels per 1000 pixels per frame), windows of 10 frames with
    stride 5, 5 Lloyd iterations per window, seeds carried across windows.
    Frames are generated up front into page-locked host memory.  Runs the
    plain per-window loop (one stream at a time, no overlap) and the served
    loop (each of `actives` concurrent-stream limits, admission by
    footprint), each with every stream's own launches on its own HIP stream
    (the default) and with the windows' Lloyd loops batched across the
    concurrent streams, and checks that every stream's stitched-label
    checksum is the same in every mode.  Both loops deliver each window's finished labels to
    host memory and keep the checksum on the device."""
    from anisotropic_supervoxel_amd import serve
    reqs = serve.make_requests(n_streams, rate_per_s=40.0, seed=seed)
    serve.attach_frames(reqs)
    out = {}
    hashes = {}
    modes = [("plain", dict(max_active=1, honor_arrivals=0, batch_windows=0))]
    modes += [(f"served{a}", dict(max_active=a, batch_windows=0)) for a in actives]
    modes += [(f"batched{a}", dict(max_active=a, batch_windows=1)) for a in actives]
    for name, kw in modes:
        res, w, wall = serve.run(serve.default_params(**kw), reqs)
        s = serve.summarize(res, w, wall)
        hashes[name] = [r["labels_hash"] for r in res]
        out[name] = {"mvoxel_iters_per_s": round(s["mvoxel_iters_per_s"], 1), "p50_window_ms": round(s["p50_window_ms"], 2),
                     "p99_window_ms": round(s["p99_window_ms"], 2), "wall_ms": round(wall, 1), "windows": s["windows"]}
    plain = out["plain"]["mvoxel_iters_per_s"]
    return {
        "workload": f"{n_streams} streams, Poisson 40/s, 320x240/640x480/1280x720, 60-300 frames, density 1-4 per "
                    "1000 px per frame, windows 10 stride 5, 5 Lloyd iterations",
        "plain_per_window_loop": out["plain"],
        "served": [dict(max_active=a, **out[f"served{a}"],
                        served_over_plain=round(out[f"served{a}"]["mvoxel_iters_per_s"] / plain, 3))
                   for a in actives],
        "served_batched_windows": [dict(max_active=a, **out[f"batched{a}"],
                                        served_over_plain=round(out[f"batched{a}"]["mvoxel_iters_per_s"] / plain, 3))
                                   for a in actives],
        "labels_identical_alone_and_in_mix": all(hashes[m] == hashes["plain"] for m in hashes),
        "voxel_iters": int(sum(r["voxel_iters"] for r in res)),
    }


def batched(n_clips, steps):
    """Batched workload: n_clips clips from the generator (video seeds
    100.., sizes {160x120, 320x240, 480x360, 640x480}, 20-120 frames,
    K 200-4000, 8-32 ellipses, 10 Lloyd iterations; see batch.py), resident
    on the device.  Times `steps` whole runs (LAB and 10 iterations
    of every clip) of the plain per-clip loop (asv_run clip after clip on the
    same contexts) and of the batched form (asv_batch_run: graphs, batched
    marching launches, autotuned shapes; tuning happens in an untimed first
    run), and checks every clip's final sums agree between the two."""
    import numpy as np
    import torch
    from anisotropic_supervoxel_amd import batch as B
    clips = B.make_clips(n_clips)
    b = B.Batch([B.clip_params(c) for c in clips])
    vox = 0
    for i, c in enumerate(clips):
        rgb, fwd, bwd, _ = B.clip_video(c)
        b.upload(i, rgb, fwd, bwd)
        vox += c["W"] * c["H"] * c["D"]
    torch.cuda.synchronize()
    iters = clips[0]["iters"]

    def plain():
        for cl in b.clips:
            cl.run()

    out = {}
    for name, fn in (("plain_per_clip_loop", plain), ("batched", b.run)):
        fn()                                   # warm-up (and, batched, autotune + capture)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        out[name] = {"ms_per_run": round(dt * 1e3, 2), "mvoxel_iters_per_s": round(vox * iters / dt / 1e6, 1)}
        out[name + "_sums"] = [b.sums(i) for i in range(n_clips)]
    same = all(np.array_equal(a, c) for a, c in zip(out.pop("plain_per_clip_loop_sums"), out.pop("batched_sums")))
    shapes, tune_ms = b.tuning()
    b.close()
    return {
        "workload": f"{n_clips} clips, video seeds 100-{99 + n_clips}, 160x120/320x240/480x360/640x480, "
                    "20-120 frames, K 200-4000, 8-32 ellipses, 10 Lloyd iterations",
        "voxels": int(vox),
        **out,
        "batched_over_plain": round(out["batched"]["mvoxel_iters_per_s"] / out["plain_per_clip_loop"]["mvoxel_iters_per_s"], 3),
        "sums_identical": bool(same),
        "jfa_launches_per_iteration": len(shapes),
        "batched_launches": sum(1 for s in shapes if s[2] > 0),
        "autotune_gpu_ms": round(tune_ms, 1),
    }


def cpu_baselines(config3_all_cores=None):
    """The CPU oracle (C, OpenMP) timed on the plain configurations, on all
    host cores (16 at most) and on one: configs 1 and 2 in full on all
    cores, config 1 in full on one; config 2 on its first 10 frames and
    config 3 on its first 5 on one core (K scaled to keep the grid step and
    schedule); config 3 on all cores is the cpu_baseline sample, passed in.
    Each row says what it ran.  The oracle evaluates the metric of every
    voxel, which dominates its single-thread time."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import anisotropic_supervoxel_amd as asv
    import oracle as orc
    allc = min(16, os.cpu_count() or 1)
    rows = []
    plan = [("plain-160x120x30-K500", None, allc), ("plain-160x120x30-K500", None, 1),
            ("plain-320x240x85-K2000", None, allc), ("plain-320x240x85-K2000", 10, 1),
            ("plain-640x480x200-K10000", 5, 1)]
    videos = {}
    for name, frames, threads in plan:
        seed, W, H, D, ell, vmax, sigma, K, iters = WORKLOADS[name]
        if name not in videos:
            videos[name] = asv.synth_video(seed, W, H, D, ell, vmax, sigma)
        rgb, fwd, bwd, _ = videos[name]
        f = frames or D
        Ks = K if f == D else max(1, K * f // D)
        orc.load().orc_set_threads(threads)
        t0 = time.perf_counter()
        orc.run(np.ascontiguousarray(rgb[:f]), np.ascontiguousarray(fwd[:f]), np.ascontiguousarray(bwd[:f]), Ks, iters)
        dt = time.perf_counter() - t0
        rows.append({"workload": name, "frames": f, "K": Ks, "cores": threads, "wall_s": round(dt, 2),
                     "mvoxel_iters_per_s": round(W * H * f * iters / dt / 1e6, 2),
                     "sample": "full clip" if f == D else f"first {f} of {D} frames, K scaled"})
        print(f"cpu baseline: {name} frames {f} cores {threads}: {dt:.1f} s", file=sys.stderr, flush=True)
    orc.load().orc_set_threads(allc)
    if config3_all_cores:
        c = config3_all_cores
        rows.insert(4, {"workload": "plain-640x480x200-K10000", "cores": c["cores"],
                        "mvoxel_iters_per_s": round(c["value"], 2), "sample": c["sample"]})
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="plain-640x480x200-K10000", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-frames", type=int, default=70)
    ap.add_argument("--no-cpu-table", action="store_true",
                    help="skip the oracle timings on configs 1-3 (all cores and one)")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--served-streams", type=int, default=24)
    ap.add_argument("--served-active", default="4,6,8", help="comma-separated concurrent-stream limits")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-clips", type=int, default=64)
    ap.add_argument("--batched-steps", type=int, default=2)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one 1 GiB torch copy (HBM counter calibration under rocprofv3 --pmc)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU sharding is a design note, not a measured path", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import anisotropic_supervoxel_amd as asv

    seed, W, H, D, ell, vmax, sigma, K, iters = WORKLOADS[args.workload]
    V = W * H * D
    rgb, fwd, bwd, regions = asv.synth_video(seed, W, H, D, ell, vmax, sigma)
    p = asv.default_params(W, H, D, K, lloyd_iters=iters)
    clip = asv.Clip(p)
    clip.upload(rgb, fwd, bwd)
    torch.cuda.synchronize()

    if args.calibrate_copy:
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    for _ in range(args.warmup):
        clip.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        clip.run()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = dt * 1000.0 / max(args.steps, 1)
    value = V * iters * args.steps / dt / 1e6

    # per-kernel timing on the launch stream (HIP events around each launch)
    ms, passes = clip.run_timed()
    jfa_avg_ms = float(ms[2]) / max(passes, 1)
    red_avg_ms = float(ms[3]) / iters
    jfa_bytes = 12.0 * V
    achieved = jfa_bytes / (jfa_avg_ms * 1e-3) / 1e9
    steps_list, mult = clip.schedule()
    # quality of the final labels (GL rows, flipped back to the video's image
    # rows) against the generator's ground-truth regions
    from anisotropic_supervoxel_amd import api as asv_api
    ue, br = asv_api.eval_metrics(np.ascontiguousarray(clip.labels()[:, ::-1, :]), regions)

    out = {
        "metric": "Mvoxel-iters/s per GPU (% HBM roofline); p50/p99 window latency, served vs plain",
        "value": round(value, 1),
        "unit": "Mvoxel-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": args.workload,
            "W": W, "H": H, "D": D, "voxels": V, "K": K, "seeds_placed": clip.A,
            "lloyd_iters": iters, "jfa_schedule": steps_list, "jfa_ratio_mult": mult,
            "tie_rule": "lowest-id", "step_includes": "LAB, seed stages (metric at the seeds), JFA, reduction, finalize",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_jfa_march (one full JFA pass)",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": load_traffic(args.workload),
            "algorithmic_bytes_per_launch": int(jfa_bytes),
            "avg_launch_ms": round(jfa_avg_ms, 4),
        },
        "breakdown_ms": {
            "total_timed_run": round(float(ms[0]), 3),
            "prepare": round(float(ms[1]), 3),
            "jfa_full_passes": round(float(ms[2]), 3),
            "jfa_seed_neighbourhood_passes": round(float(ms[5]), 3),
            "reduction_all": round(float(ms[3]), 3),
            "seed_and_finalize": round(float(ms[4]), 3),
            "reduction_avg": round(red_avg_ms, 4),
            "reduction_GBps": round(8.0 * V / (red_avg_ms * 1e-3) / 1e9, 1),
        },
        "quality": {
            "undersegmentation_error": round(ue, 5),
            "boundary_recall": round(br, 5),
            "against": "ground-truth regions of the synthetic video; final labels before post-processing",
        },
        "cpu_baseline": None,
        "served": None,
        "batched": None,
    }
    clip.close()
    if not args.no_batched:
        out["batched"] = batched(args.batched_clips, args.batched_steps)
    if not args.no_served:
        out["served"] = served(args.served_streams, [int(a) for a in args.served_active.split(",")], seed=5)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(seed, W, H, D, ell, vmax, sigma, K, iters, args.cpu_frames)
    if not args.no_cpu_table:
        out["cpu_baselines"] = cpu_baselines(out["cpu_baseline"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
