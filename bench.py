"""Benchmark of the ORB front-end on one MI355X.

Workload (the plain per-frame HIP path): a 32-frame 1280x720 sequence from
the seeded generator (each frame the previous one shifted by an integer in
[-4, 4]^2 plus fresh noise), 6 levels x1.2, FAST threshold 20/255, 2000
features per level; each frame goes through the feature-detect stage and is
matched against the previous frame with the reference matcher semantics
(LOD+-1 pools, 10% window, ratio 0.8, mutual check).  One step = the whole
sequence: 32 detects + 31 matches, inputs resident in HBM.

Prints one JSON line (rank 0).  Diagnostics go to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

W, H, N_FRAMES, CAP, LEVELS = 1280, 720, 32, 2000, 6
RATIO = 0.8
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def thr20():
    return float(np.float32(20) / np.float32(255))


def kernel_bytes(dims, cap, counts):
    """Algorithmic HBM bytes per launch of each kernel for one frame (or one
    frame pair for the matcher kernels)."""
    frame = dims[0][0] * dims[0][1]
    pyr = sum(w * h for w, h in dims[1:])
    words = sum(h * ((w + 63) // 64) for w, h in dims)
    feat_array = LEVELS * (8 + 48 * cap)
    kept = sum(min(c, cap) for c in counts)
    return {
        # u8 frame in, f32 levels 1..5 out, cull mask in, survivor bitmap out
        "k_detect": frame + 4 * pyr + 16 * words,
        # survivor bitmap in, keypoint list out, counts out
        "k_compact": 8 * words + 4 * kept + 8 * LEVELS,
        # levels in (u8 level 0, f32 others), keypoints in, feature array out
        "k_describe": frame + 4 * pyr + 4 * kept + feat_array,
        # two feature arrays in, one int32 per slot out
        "k_match_dir": 2 * feat_array + 4 * LEVELS * cap,
        "k_match_mutual": 2 * feat_array + 8 * LEVELS * cap,
    }


def cpu_baseline(frames, n_sample):
    """The oracle (scalar C++, one thread) on the first n_sample frames of
    the sequence: detect each and match consecutive pairs."""
    from oracle import oracle as O
    p = O.params(W, H, threshold=thr20(), cap=CAP)
    t0 = time.perf_counter()
    prev = None
    for t in range(n_sample):
        f, _ = O.detect_frame(frames[t], p)
        cur = O.load_features(f, LEVELS, CAP)
        if prev is not None:
            O.find_matches(prev[0], prev[1], cur[0], cur[1], W, H, RATIO)
        prev = cur
    dt = time.perf_counter() - t0
    return n_sample / dt, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-sample", type=int, default=8,
                    help="frames of the sequence the CPU baseline processes")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)

    import torch
    from uni_slam_amd import orbfe, synth

    dev = torch.device("cuda:0")
    frames, _ = synth.sequence(1, W, H, n_frames=N_FRAMES)
    d_frames = torch.from_numpy(frames).to(dev)
    ctx = orbfe.Context(W, H, features_per_level=CAP, fast_threshold=thr20(), lowe_ratio=RATIO)
    feats = [torch.zeros(ctx.feature_bytes, dtype=torch.uint8, device=dev)
             for _ in range(N_FRAMES)]
    matches = [torch.zeros(LEVELS * CAP, dtype=torch.int32, device=dev)
               for _ in range(N_FRAMES - 1)]
    stream = torch.cuda.current_stream()

    def step():
        for t in range(N_FRAMES):
            ctx.detect_async(d_frames[t], W, feats[t], stream=stream)
            if t > 0:
                ctx.match_pipeline_async(feats[t - 1], feats[t], matches[t - 1], stream=stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = 1000.0 * dt / args.steps
    fps = N_FRAMES * args.steps / dt
    n_matches = int(sum(int((m >= 0).sum()) for m in matches))
    log(f"plain: {ms_per_step:.3f} ms/step, {fps:.1f} frames/s, {n_matches} matches/step")

    # Per-kernel durations over the same work, with HIP events on the stream
    # the kernels run on.
    ctx.timing(True)
    ctx.timing_reset()
    for _ in range(max(2, args.steps // 4)):
        step()
    torch.cuda.synchronize()
    names = ["k_detect", "k_compact", "k_describe", "k_match_dir", "k_match_mutual"]
    times = {n: ctx.timing_read(n) for n in names}
    ctx.timing(False)
    counts = [c for c, _ in orbfe.parse_feature_array(feats[0].cpu().numpy(), LEVELS, CAP)]
    kb = kernel_bytes(ctx.dims, CAP, counts)
    per_kernel = {}
    for n, (launches, total) in times.items():
        if launches:
            avg_us = 1000.0 * total / launches
            per_kernel[n] = {"launches": launches, "avg_us": round(avg_us, 3),
                             "bytes": kb[n], "GBps": round(kb[n] / (avg_us * 1e-6) / 1e9, 2)}
    log("per-kernel:", json.dumps(per_kernel))
    dom = max(per_kernel, key=lambda n: per_kernel[n]["avg_us"] * per_kernel[n]["launches"])
    ach = per_kernel[dom]["GBps"]
    roofline = {"kernel": dom, "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 5), "traffic": None,
                "avg_us": per_kernel[dom]["avg_us"], "bytes_per_launch": kb[dom]}

    cpu = None
    if not args.no_cpu:
        cfps, cdt = cpu_baseline(frames, args.cpu_sample)
        cpu = {"value": round(cfps, 4), "unit": "frames/s", "cores": 1, "kind": "port",
               "sample": f"first {args.cpu_sample} frames of the sequence: detect each, "
                         f"match {args.cpu_sample - 1} consecutive pairs ({cdt:.1f} s)"}
        log(f"cpu baseline: {cfps:.3f} frames/s")

    out = {
        "metric": "frames/s (plain per-frame HIP path: detect + match vs previous frame)",
        "value": round(fps, 2), "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8/f32/u32",
        "data": "synthetic (seeded value noise + rectangles + N(0,3) noise)",
        "config": {"workload": "config 1: plain per-frame path, 1280x720, 32-frame shifted "
                               "sequence, cap 2000/level, threshold 20/255, ratio 0.8",
                   "frames_per_step": N_FRAMES, "resolution": [W, H], "cap": CAP,
                   "parallelism": "single GPU"},
        "matches_per_step": n_matches,
        "per_kernel": per_kernel,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
