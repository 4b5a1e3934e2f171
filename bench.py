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
PEAK_INT8_MFMA_POPS = 5.0  # MI355X dense int8 matrix-core peak (published)


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
        # two feature arrays' positions in, two cell indices out
        "k_match_index": 2 * 8 * kept + 2 * 4 * LEVELS * (577 + cap),
        # two feature arrays in, at most one 8-byte list entry per slot out
        "k_match_dir": 2 * feat_array + 8 * LEVELS * cap,
        # two feature arrays and the list entries in, prev_match out
        "k_match_mutual": 2 * feat_array + 12 * LEVELS * cap,
    }


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated against a copy of
    known size), or None when no summary covers it."""
    path = os.path.join(ROOT, "profiles", "traffic_plain_1280x720.json")
    try:
        with open(path) as f:
            t = json.load(f)["kernels"][kernel]
        return int(t["fetch_bytes"] + t["write_bytes"])
    except (OSError, KeyError, ValueError):
        return None


def valu_roofline(kernel, avg_us, peak_rate):
    """VALU roofline of `kernel` on this workload: its VALU instructions per
    launch (rocprofv3 SQ_INSTS_VALU, committed summary; the count is fixed by
    the workload) over the average launch duration measured in this run,
    against the sustained VALU issue rate probed on this device.  None when
    no summary covers the kernel."""
    path = os.path.join(ROOT, "profiles", "valu_1280x720.json")
    try:
        with open(path) as f:
            v = json.load(f)
        n = v["plain"][kernel]["valu_per_launch"]
        batched = v["batched"][kernel]["util"]
    except (OSError, KeyError, ValueError):
        return None
    ach = n / (avg_us * 1e-6)
    return {"valu_per_launch": n, "achieved_wave_instr_per_s": float(f"{ach:.4g}"),
            "peak_wave_instr_per_s": float(f"{peak_rate:.4g}"), "frac": round(ach / peak_rate, 4),
            "frac_in_32_frame_batch": batched, "source": "profiles/valu_1280x720.json"}


def cpu_baseline(frames, n_sample):
    """The oracle (scalar C++, one thread) on n_sample frames of the sequence
    (cycling through it): detect each and match consecutive pairs."""
    from oracle import oracle as O
    p = O.params(W, H, threshold=thr20(), cap=CAP)
    t0 = time.perf_counter()
    prev = None
    for t in range(n_sample):
        f, _ = O.detect_frame(frames[t % len(frames)], p)
        cur = O.load_features(f, LEVELS, CAP)
        if prev is not None:
            O.find_matches(prev[0], prev[1], cur[0], cur[1], W, H, RATIO)
        prev = cur
    dt = time.perf_counter() - t0
    return n_sample / dt, dt


def cpu_info():
    """Host CPU model, logical CPUs and the CPUs this process may run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"cpu_model": model, "nproc": os.cpu_count(), "cpus_available": avail}


def median_time(fn, runs=5):
    fn()  # warm-up
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def cpu_baseline_pair(synth):
    """The oracle on the 752x480 frame pair (detect both frames at cap 1000,
    match with ratio 0.8): pairs/s and frames/s, one thread, median of 5."""
    from oracle import oracle as O
    a, b, _ = synth.frame_pair(0, 752, 480)
    p = O.params(752, 480, threshold=thr20(), cap=1000)

    def run():
        fa, _ = O.detect_frame(a, p)
        fb, _ = O.detect_frame(b, p)
        A, ca = O.load_features(fa, LEVELS, 1000)
        B, cb = O.load_features(fb, LEVELS, 1000)
        O.find_matches(A, ca, B, cb, 752, 480, RATIO)
    t = median_time(run)
    return {"value": round(1.0 / t, 4), "unit": "frame pairs/s", "frames_per_s": round(2.0 / t, 4),
            "cores": 1, "kind": "port", "runs": 5, "seconds_per_pair": round(t, 4),
            "sample": "the 752x480 pair (detect both at cap 1000, match at ratio 0.8)"}


def cpu_baseline_matcher(synth):
    """The oracle's brute-force matcher on 1024 queries x 65536 train
    descriptors (a bounded sample of the 65536^2 workload), one thread,
    median of 5: pairs/s."""
    from oracle import oracle as O
    q, t, _, _ = synth.descriptors(65536, 1024, 65536)
    tm = median_time(lambda: O.match_bf(q, t))
    return {"value": round(1024 * 65536 / tm, 1), "unit": "pairs/s", "cores": 1, "kind": "port",
            "runs": 5, "sample": "1024 queries x 65536 train (%.3f s per run)" % tm}


# MI355X published dense FP32 vector rate: 157.3 TFLOP/s = 78.6 T FMA lanes/s,
# i.e. 2 wave64 VALU instructions per CU per clock at 2.4 GHz.  The integer
# rooflines use the sustained rate measured on the device itself
# (orbfe_probe_valu_rate), which bench.py reports beside it.
PUBLISHED_VALU_WAVE_INSTR_PER_S = 256 * 2 * 2.4e9


def capture_graph(torch, launch):
    """Capture launch(stream) into a HIP graph (torch's graph API drives
    hipStreamBeginCapture on the stream the kernels are enqueued on)."""
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        launch(side)  # warm-up outside capture
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        launch(torch.cuda.current_stream())
    return g


def timed(torch, fn, reps):
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def phase_batched_sequence(torch, orbfe, ctx, d_frames, feats, matches, reps):
    """The 1280x720 32-frame sequence as one prepared batch (32 detects) + one
    prepared match batch (31 pairs), captured in a HIP graph."""
    frames = [{"image": d_frames[t], "width": W, "height": H, "pitch": W, "cap": CAP,
               "features": feats[t]} for t in range(N_FRAMES)]
    pairs = [{"prev": feats[t - 1], "curr": feats[t], "prev_match": matches[t - 1],
              "width": W, "height": H, "cap": CAP} for t in range(1, N_FRAMES)]
    want = [f.clone() for f in feats] + [m.clone() for m in matches]  # the plain path's
    b = ctx.batch_create(frames)
    mb = ctx.match_batch_create(pairs)

    def launch(stream):
        b.detect_async(stream=stream)
        mb.match_async(stream=stream)
    g = capture_graph(torch, launch)
    for x in feats + matches:
        x.fill_(9)
    g.replay()
    torch.cuda.synchronize()
    exact = all(torch.equal(x, w) for x, w in zip(feats + matches, want))
    dt = timed(torch, g.replay, reps)
    return {"frames_per_s": round(N_FRAMES / dt, 1), "ms_per_step": round(dt * 1e3, 4),
            "launches_per_step": 5, "graph": True, "bit_exact_vs_plain": exact}, (b, mb, g)


def phase_mixed(torch, orbfe, synth, reps):
    """Mixed-resolution batch: 64 frames of random resolution and cap, each matched against
    its translated twin; plain per-frame launches vs one prepared batch per
    stage in a HIP graph, its launch shapes picked by the autotuner."""
    from uni_slam_amd import autotune
    items = synth.mixed_batch(2, n_frames=64)
    dev = torch.device("cuda:0")
    per_res = {}
    frames, pairs, plain_calls = [], [], []
    for it in items:
        h, w = it["a"].shape
        cap = it["cap"]
        key = (w, h, cap)
        if key not in per_res:
            per_res[key] = orbfe.Context(w, h, features_per_level=cap, fast_threshold=thr20(),
                                         lowe_ratio=RATIO)
        c = per_res[key]
        fa = torch.zeros(orbfe.feature_array_bytes(6, cap), dtype=torch.uint8, device=dev)
        fb = torch.zeros_like(fa)
        pm = torch.zeros(6 * cap, dtype=torch.int32, device=dev)
        da = torch.from_numpy(it["a"]).to(dev)
        db = torch.from_numpy(it["b"]).to(dev)
        frames += [{"image": da, "width": w, "height": h, "pitch": w, "cap": cap, "features": fa},
                   {"image": db, "width": w, "height": h, "pitch": w, "cap": cap, "features": fb}]
        pairs.append({"prev": fa, "curr": fb, "prev_match": pm, "width": w, "height": h,
                      "cap": cap})
        plain_calls.append((c, da, db, w, fa, fb, pm))
    stream = torch.cuda.current_stream()

    def plain():
        for c, da, db, w, fa, fb, pm in plain_calls:
            c.detect_async(da, w, fa, stream=stream)
            c.detect_async(db, w, fb, stream=stream)
            c.match_pipeline_async(fa, fb, pm, stream=stream)
    t_plain = timed(torch, plain, reps)
    want = [f["features"].clone() for f in frames] + [p["prev_match"].clone() for p in pairs]
    tune = autotune.tune_detect(frames, 752, 480, reps=max(3, reps), fast_threshold=thr20(),
                                lowe_ratio=RATIO)
    mtune = autotune.tune_match(pairs, 752, 480, reps=max(3, reps), fast_threshold=thr20(),
                                lowe_ratio=RATIO)
    ctx = orbfe.Context(752, 480, fast_threshold=thr20(), lowe_ratio=RATIO, **tune.config(),
                        **mtune.config())
    b = ctx.batch_create(frames)
    mb = ctx.match_batch_create(pairs)

    def launch(st):
        b.detect_async(stream=st)
        mb.match_async(stream=st)
    g = capture_graph(torch, launch)
    outs = [f["features"] for f in frames] + [p["prev_match"] for p in pairs]
    for x in outs:
        x.fill_(9)
    g.replay()
    torch.cuda.synchronize()
    exact = all(torch.equal(x, w) for x, w in zip(outs, want))
    t_batch = timed(torch, g.replay, reps)
    n = len(frames)
    return {"frames": n, "pairs": len(pairs), "bit_exact_vs_plain": exact,
            "plain_frames_per_s": round(n / t_plain, 1),
            "batched_frames_per_s": round(n / t_batch, 1),
            "speedup": round(t_plain / t_batch, 2),
            "plain_ms": round(t_plain * 1e3, 3), "batched_ms": round(t_batch * 1e3, 3),
            "tuned_shape": {"detect_tile_rows": tune.best[0], "describe_lanes": tune.best[1],
                            "match_lanes": mtune.best},
            "detect_stage_ms_by_shape": {f"{r}x{d}": round(t * 1e3, 4)
                                         for (r, d), t in sorted(tune.seconds.items())},
            "match_stage_ms_by_lanes": {str(n): round(t * 1e3, 4)
                                        for n, t in sorted(mtune.seconds.items())}}


def valu_peak(ctx):
    """Sustained VALU issue rate (wave64 instructions/s) and the integer lane
    rate it gives (Tops/s)."""
    rate = ctx.probe_valu_rate()
    return {"wave_instr_per_s": float(f"{rate:.4g}"), "int_Tops": round(rate * 64 / 1e12, 2),
            "published_wave_instr_per_s": PUBLISHED_VALU_WAVE_INSTR_PER_S,
            "source": "orbfe_probe_valu_rate (8 independent xor/add chains per lane, every CU)"}


def phase_matcher(torch, orbfe, synth, n, reps, int_peak_tops, nt=None):
    """Brute-force matcher stress: best/second-best of n query against nt
    (default n) train 256-bit descriptors."""
    nq, nt = n, (n if nt is None else nt)
    q, t, _, _ = synth.descriptors(nq * 7 + nt, nq, nt)
    dev = torch.device("cuda:0")
    dq = torch.from_numpy(q).to(dev)
    dt_ = torch.from_numpy(t).to(dev)
    idx = torch.empty(nq, dtype=torch.int32, device=dev)
    d1 = torch.empty_like(idx)
    d2 = torch.empty_like(idx)
    ctx = orbfe.Context(64, 48)
    stream = torch.cuda.current_stream()
    mfma_probe = ctx.probe_mfma_rate(stream=stream) / 1e15
    dt = timed(torch, lambda: ctx.match_bf_async(dq, nq, dt_, nt, idx, d1, d2, stream=stream),
               reps)
    ctx.timing(True)
    ctx.timing_reset()
    for _ in range(reps):
        ctx.match_bf_async(dq, nq, dt_, nt, idx, d1, d2, stream=stream)
    torch.cuda.synchronize()
    launches, ms = ctx.timing_read("k_bf_mfma")
    ctx.timing(False)
    kern = ms / launches * 1e-3
    pairs = nq * nt
    # k_bf_mfma forms every distance as an int8 dot product of two 256-sign
    # vectors on the matrix cores: 256 multiply-adds = 512 int8 ops per pair.
    mfma_pops = 512 * pairs / kern / 1e15
    return {"n_query": nq, "n_train": nt, "ms": round(dt * 1e3, 4),
            "pairs_per_s": round(pairs / dt, 1),
            "effective_GBps": round(32 * pairs / dt / 1e9, 1),
            "mfma_int8_Pops": round(mfma_pops, 3),
            "mfma_int8_peak_Pops": PEAK_INT8_MFMA_POPS,
            "mfma_frac": round(mfma_pops / PEAK_INT8_MFMA_POPS, 4),
            "mfma_int8_probe_Pops": round(mfma_probe, 3),
            "mfma_frac_of_probe": round(mfma_pops / mfma_probe, 4),
            "mfma_probe_source": "orbfe_probe_mfma_rate (2 independent v_mfma_i32_32x32x32_i8 "
                                 "chains per wave, 8 waves per SIMD, every CU)",
            "xor_popcount_equiv_Tops": round(16 * pairs / kern / 1e12, 2),
            "int_peak_Tops": round(int_peak_tops, 2),
            "k_bf_mfma_us": round(kern * 1e6, 2)}


def phase_matcher_grid(torch, orbfe, synth, reps):
    """Every N x M combination of the matcher stress workload (N, M in
    {4096, 16384, 65536}): pairs/s (matches/s) of the whole call."""
    out = {}
    for nq in (4096, 16384, 65536):
        for nt in (4096, 16384, 65536):
            r = phase_matcher(torch, orbfe, synth, nq, reps, 0.0, nt=nt)
            out[f"{nq}x{nt}"] = {"ms": r["ms"], "pairs_per_s": r["pairs_per_s"],
                                 "mfma_int8_Pops": r["mfma_int8_Pops"]}
    return out


def phase_readback(torch, orbfe, ctx, d_frames, feats, reps):
    """The readback of one detected 1280x720 frame into the host FeatureSet:
    before, the whole feature slab copied to pinned memory and converted on
    the CPU (orbfe_load_features, the FromGpuModel loop); after, the export
    kernel writing the wslam::Feature records straight into pinned memory
    (orbfe_export_features_async).  Microseconds per frame over the
    sequence, detect excluded; both give the same records."""
    stream = torch.cuda.current_stream()
    slab = torch.empty(ctx.feature_bytes, dtype=torch.uint8).pin_memory()
    out = torch.empty(LEVELS * CAP * 13, dtype=torch.int32).pin_memory()
    counts = torch.empty(LEVELS, dtype=torch.int32).pin_memory()
    for t in range(N_FRAMES):
        ctx.detect_async(d_frames[t], W, feats[t], stream=stream)
    torch.cuda.synchronize()

    def before():
        for t in range(N_FRAMES):
            slab.copy_(feats[t], non_blocking=True)
            torch.cuda.synchronize()
            rec, _ = orbfe.load_features(slab.numpy(), LEVELS, CAP)
        return rec

    def after():
        for t in range(N_FRAMES):
            ctx.export_features_async(feats[t], out, counts, stream=stream)
            torch.cuda.synchronize()
        n = int(counts.numpy().sum())
        return out.numpy().view(np.uint8)[:52 * n].view(orbfe.FEATURE_DTYPE)

    same = np.array_equal(before().view(np.uint8), after().view(np.uint8))
    tb = median_time(before, reps) / N_FRAMES
    ta = median_time(after, reps) / N_FRAMES
    return {"before_us_per_frame": round(tb * 1e6, 2), "after_us_per_frame": round(ta * 1e6, 2),
            "speedup": round(tb / ta, 2), "same_records": bool(same),
            "before": "feature slab D2H to pinned memory + CPU FromGpuModel loop",
            "after": "export kernel writes wslam::Feature records into pinned memory"}


def phase_ransac(torch, orbfe, reps):
    """RANSAC hypothesis scoring: 500 fundamental matrices (the reference's
    max_iterations) against 8000 correspondences in one launch
    (orbfe_ransac_score_async), beside the oracle's scalar loop on one
    thread; hypotheses x correspondences per second."""
    from oracle import oracle as O
    rng = np.random.default_rng(5)
    n, k = 8000, 500
    prev = rng.uniform(0, 1280, (n, 2))
    curr = prev + rng.normal(0, 3, (n, 2))
    F = rng.normal(size=(k, 9)) * 1e-3
    ctx = orbfe.Context(64, 48)
    dF = torch.from_numpy(F).cuda()
    dp = torch.from_numpy(prev).cuda()
    dc = torch.from_numpy(curr).cuda()
    counts = torch.empty(k, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()
    dt = timed(torch, lambda: ctx.ransac_score_async(dF, k, dp, dc, n, 3.84, counts,
                                                     stream=stream), reps)
    same = np.array_equal(counts.cpu().numpy().view(np.uint32),
                          O.sampson_counts(F, prev, curr, 3.84))
    tc = median_time(lambda: O.sampson_counts(F[:50], prev, curr, 3.84), 3) * (k / 50)
    return {"hypotheses": k, "correspondences": n, "gpu_ms": round(dt * 1e3, 4),
            "gpu_evals_per_s": round(k * n / dt, 1), "cpu_evals_per_s": round(k * n / tc, 1),
            "cpu_cores": 1, "counts_equal_oracle": bool(same)}


def phase_serving(torch, orbfe, synth, seconds, seed=4):
    """Serving: 32 streams at mixed resolutions, Poisson arrivals at
    20-60 fps each, continuous batching on two in-flight slots."""
    from uni_slam_amd import serve
    rng = np.random.Generator(np.random.PCG64(seed))
    specs, rates, pool = [], [], []
    for s in range(32):
        w, h = synth.RESOLUTIONS[int(rng.integers(0, 4))]
        specs.append(serve.StreamSpec(w, h, 2000))
        rates.append(float(rng.uniform(20, 60)))
        fr, _ = synth.sequence(1000 + s, w, h, n_frames=4)
        pool.append([torch.from_numpy(f).pin_memory() for f in fr])
    arrivals = serve.poisson_arrivals(rates, seconds, seed)
    srv = serve.Server(specs, fast_threshold=thr20(), lowe_ratio=RATIO, hbm_budget=2 << 30)
    wall, done = serve.run_schedule(srv, arrivals, pool)
    stats = serve.latency_stats(done)
    vga = [c for c in done if (specs[c.stream].width, specs[c.stream].height) == (752, 480)]
    out = {"streams": 32, "seconds": seconds, "frames": len(done),
           "frames_per_s": round(len(done) / wall, 1), "batches": srv.batches,
           "max_batch": srv.max_batch_seen,
           "latency_ms": {k: round(v, 3) for k, v in stats.items() if k.endswith("_ms")}}
    if vga:
        out["latency_752x480_ms"] = {k: round(v, 3) for k, v in serve.latency_stats(vga).items()
                                     if k.endswith("_ms")}
    # capacity: the same streams with every frame of a 2 s schedule already
    # waiting (arrivals compressed to zero), frames/s of the drained backlog
    srv2 = serve.Server(specs, fast_threshold=thr20(), lowe_ratio=RATIO, hbm_budget=2 << 30)
    arr2 = serve.poisson_arrivals(rates, 2.0, seed + 1)
    wall2, done2 = serve.run_schedule(srv2, arr2, pool, time_scale=0.0)
    out["saturated_frames_per_s"] = round(len(done2) / wall2, 1)
    # the plain per-frame path over the same backlog (upload + per-frame launches)
    wall_p = serve.plain_backlog(specs, arr2, pool, fast_threshold=thr20(), lowe_ratio=RATIO)
    out["plain_frames_per_s"] = round(len(arr2) / wall_p, 1)
    out["saturated_vs_plain"] = round(wall_p / wall2, 2)
    # parity: a short served run of all 32 streams (8 frames each, compressed
    # schedule) with outputs kept, against the plain per-frame path
    srv3 = serve.Server(specs, fast_threshold=thr20(), lowe_ratio=RATIO, hbm_budget=2 << 30,
                        keep_outputs=True)
    arr3 = [a for a in serve.poisson_arrivals(rates, 10.0, seed + 2) if a[2] < 8]
    serve.run_schedule(srv3, arr3, pool, time_scale=0.0)
    exact = len(srv3.outputs) == len(arr3)
    ctxs = {}
    for s, sp in enumerate(specs):
        key = (sp.width, sp.height)
        if key not in ctxs:
            ctxs[key] = orbfe.Context(sp.width, sp.height, features_per_level=sp.cap,
                                      fast_threshold=thr20(), lowe_ratio=RATIO)
        c = ctxs[key]
        prev = None
        for k in range(8):
            feats, _ = c.detect(pool[s][k % len(pool[s])].numpy())
            got_f, got_m = srv3.outputs[(s, k)]
            exact &= bool(np.array_equal(got_f, feats))
            if prev is not None:
                exact &= bool(np.array_equal(got_m, c.match_pipeline(prev, feats)[0]))
            prev = feats
    for c in ctxs.values():
        c.close()
    out["bit_exact_vs_plain"] = exact
    out["parity_sample"] = f"{len(arr3)} served frames of the 32 streams vs the plain path"
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-sample", type=int, default=2 * N_FRAMES,
                    help="frames the CPU baseline processes (cycling through the sequence)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the batched / mixed / matcher / serving phases")
    ap.add_argument("--serve-seconds", type=float, default=60.0,
                    help="length of the real-time serving schedule (the workload's 60 s)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 256 MiB device-to-device copy (counter calibration)")
    ap.add_argument("--workload", default="plain",
                    choices=["plain", "batched", "mixed", "matcher", "matcher_grid", "serving",
                             "readback", "ransac"],
                    help="plain: the full report; the others run only that phase and "
                         "print its JSON (for profiling one path)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = 256 << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n, "reps": 3}), flush=True)
        return
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

    if args.workload != "plain":
        if args.workload == "batched":
            # the plain path's results, which the batched form must reproduce
            for t in range(N_FRAMES):
                ctx.detect_async(d_frames[t], W, feats[t], stream=stream)
                if t > 0:
                    ctx.match_pipeline_async(feats[t - 1], feats[t], matches[t - 1],
                                             stream=stream)
            torch.cuda.synchronize()
            out, _keep = phase_batched_sequence(torch, orbfe, ctx, d_frames, feats, matches,
                                                max(2, args.steps))
        elif args.workload == "mixed":
            out = phase_mixed(torch, orbfe, synth, max(2, args.steps))
        elif args.workload == "matcher":
            peak = valu_peak(ctx)
            out = phase_matcher(torch, orbfe, synth, 65536, max(2, args.steps), peak["int_Tops"])
            out["valu_peak"] = peak
        elif args.workload == "matcher_grid":
            out = phase_matcher_grid(torch, orbfe, synth, max(2, args.steps))
        elif args.workload == "readback":
            out = phase_readback(torch, orbfe, ctx, d_frames, feats, 5)
        elif args.workload == "ransac":
            out = phase_ransac(torch, orbfe, max(2, args.steps))
        else:
            out = phase_serving(torch, orbfe, synth, args.serve_seconds)
        print(json.dumps({"workload": args.workload, **out}), flush=True)
        return

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
    mps = n_matches * args.steps / dt
    log(f"plain: {ms_per_step:.3f} ms/step, {fps:.1f} frames/s, {n_matches} matches/step, "
        f"{mps:.0f} matches/s")

    # Per-kernel durations over the same work, with HIP events on the stream
    # the kernels run on.
    ctx.timing(True)
    ctx.timing_reset()
    for _ in range(max(2, args.steps // 4)):
        step()
    torch.cuda.synchronize()
    names = ["k_detect", "k_compact", "k_describe", "k_match_index", "k_match_dir",
             "k_match_mutual"]
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
    peak = valu_peak(ctx)
    log("valu peak:", json.dumps(peak))
    dom = max(per_kernel, key=lambda n: per_kernel[n]["avg_us"] * per_kernel[n]["launches"])
    ach = per_kernel[dom]["GBps"]
    roofline = {"kernel": dom, "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 5),
                "traffic": measured_traffic(dom),
                "valu": valu_roofline(dom, per_kernel[dom]["avg_us"], peak["wave_instr_per_s"]),
                "limiter": "priced against HBM (the roofline of its algorithmic bytes); what "
                           "limits it is instruction issue and per-tile latency: f32 bilinear "
                           "and 9-tap smoothing arithmetic (roofline.valu: VALU issue against "
                           "the probed peak, live on one frame and from the profile in "
                           "32-frame batches); one frame's 2775 tiles fill the device 1.35 "
                           "times at 8 workgroups per CU (the wave-slot limit, 8 waves per "
                           "SIMD) (profiles/valu_1280x720.md)",
                "avg_us": per_kernel[dom]["avg_us"], "bytes_per_launch": kb[dom]}

    extra = {}
    if not args.no_extra:
        bseq, _keep = phase_batched_sequence(torch, orbfe, ctx, d_frames, feats, matches,
                                             max(5, args.steps))
        bseq["speedup_vs_plain"] = round(bseq["frames_per_s"] / fps, 2)
        extra["batched_sequence"] = bseq
        log("batched sequence:", json.dumps(bseq))
        extra["mixed_batch"] = phase_mixed(torch, orbfe, synth, max(3, args.steps // 4))
        log("mixed batch:", json.dumps(extra["mixed_batch"]))
        extra["valu_peak"] = peak
        extra["matcher"] = phase_matcher(torch, orbfe, synth, 65536, 3, peak["int_Tops"])
        log("matcher:", json.dumps(extra["matcher"]))
        extra["matcher_grid"] = phase_matcher_grid(torch, orbfe, synth, 3)
        log("matcher grid:", json.dumps(extra["matcher_grid"]))
        extra["readback"] = phase_readback(torch, orbfe, ctx, d_frames, feats, 5)
        log("readback:", json.dumps(extra["readback"]))
        extra["ransac"] = phase_ransac(torch, orbfe, 5)
        log("ransac:", json.dumps(extra["ransac"]))
        extra["serving"] = phase_serving(torch, orbfe, synth, args.serve_seconds)
        log("serving:", json.dumps(extra["serving"]))

    cpu = None
    if not args.no_cpu:
        cfps, cdt = cpu_baseline(frames, args.cpu_sample)
        cpu = {"value": round(cfps, 4), "unit": "frames/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_sample} frames of the 1280x720 sequence (cycled): detect "
                         f"each, match {args.cpu_sample - 1} consecutive pairs ({cdt:.1f} s)",
               "threads": 1, **cpu_info()}
        log(f"cpu baseline: {cfps:.3f} frames/s")
        extra["cpu_baseline_pair_752x480"] = cpu_baseline_pair(synth)
        log("cpu pair:", json.dumps(extra["cpu_baseline_pair_752x480"]))
        extra["cpu_baseline_matcher"] = cpu_baseline_matcher(synth)
        log("cpu matcher:", json.dumps(extra["cpu_baseline_matcher"]))

    out = {
        "metric": "frames/s (plain per-frame HIP path: detect + match vs previous frame)",
        "value": round(fps, 2), "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8/f32/u32",
        "data": "synthetic (seeded value noise + rectangles + N(0,3) noise)",
        "config": {"workload": "plain per-frame path: 1280x720, 32-frame shifted "
                               "sequence, cap 2000/level, threshold 20/255, ratio 0.8",
                   "frames_per_step": N_FRAMES, "resolution": [W, H], "cap": CAP,
                   "parallelism": "single GPU"},
        "matches_per_step": n_matches,
        "matches_per_s": round(mps, 1),
        "per_kernel": per_kernel,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
