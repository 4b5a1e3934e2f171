"""The benchmark workloads at their full size, checked bit for bit against
the oracle's SHA-256 hashes in tests/golden/workloads.json (written on the
CPU by tests/golden/make_workloads.py):

* the whole 32-frame 1280x720 sequence through the plain per-frame path and
  through the prepared batch replayed from a HIP graph;
* the 64-pair mixed-resolution batch (640x480 .. 1920x1080, caps 1000..4000,
  98 levels over their cap) as one prepared batch per stage in a HIP graph,
  replayed with eager work in between, in the default and a non-default
  launch shape, and through plain per-frame launches;
* the 32 camera streams of the serving workload through the server.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from tests.conftest import thr20
from uni_slam_amd import orbfe, synth

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
RATIO = 0.8


def _say(*a):
    print(*a, flush=True)


def sha(a) -> str:
    if hasattr(a, "cpu"):
        a = a.cpu().numpy()
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def workloads():
    with open(os.path.join(HERE, "workloads.json")) as f:
        return json.load(f)


def capture(torch, launch):
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        launch(side)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        launch(torch.cuda.current_stream())
    return g


def test_sequence_1280x720x32_plain_and_batched(torch_gpu):
    torch = torch_gpu
    gold = workloads()["seq_1280x720x32"]["frames"]
    W, H, cap, n = 1280, 720, 2000, 32
    frames, _ = synth.sequence(1, W, H, n_frames=n)
    dev = torch.device("cuda:0")
    d_frames = torch.from_numpy(frames).to(dev)
    with orbfe.Context(W, H, features_per_level=cap, fast_threshold=thr20(),
                       lowe_ratio=RATIO) as ctx:
        feats = [torch.zeros(ctx.feature_bytes, dtype=torch.uint8, device=dev) for _ in range(n)]
        pyrs = [torch.zeros(ctx.pyramid_floats, dtype=torch.float32, device=dev)
                for _ in range(n)]
        pms = [torch.zeros(6 * cap, dtype=torch.int32, device=dev) for _ in range(n - 1)]
        _say("plain path: 32 frames 1280x720, 31 matches")
        for t in range(n):
            ctx.detect_async(d_frames[t], W, feats[t], pyramid=pyrs[t])
            if t > 0:
                ctx.match_pipeline_async(feats[t - 1], feats[t], pms[t - 1])
        torch.cuda.synchronize()
        for t in range(n):
            assert sha(feats[t]) == gold[t]["features"], t
            assert sha(pyrs[t]) == gold[t]["pyramid"], t
            if t > 0:
                assert sha(pms[t - 1]) == gold[t]["prev_match"], t
                assert int((pms[t - 1] >= 0).sum()) == gold[t]["n_matches"], t
        for x in feats + pms:
            x.fill_(3)
        _say("prepared batch of the sequence in a HIP graph")
        b = ctx.batch_create([{"image": d_frames[t], "width": W, "height": H, "pitch": W,
                               "cap": cap, "features": feats[t]} for t in range(n)])
        mb = ctx.match_batch_create([{"prev": feats[t - 1], "curr": feats[t],
                                      "prev_match": pms[t - 1], "width": W, "height": H,
                                      "cap": cap} for t in range(1, n)])

        def launch(s):
            b.detect_async(stream=s)
            mb.match_async(stream=s)
        g = capture(torch, launch)
        for x in feats + pms:
            x.fill_(5)
        g.replay()
        torch.cuda.synchronize()
        for t in range(n):
            assert sha(feats[t]) == gold[t]["features"], t
            if t > 0:
                assert sha(pms[t - 1]) == gold[t]["prev_match"], t
        b.close()
        mb.close()


def _mixed_buffers(torch, items):
    dev = torch.device("cuda:0")
    frames, pairs = [], []
    for it in items:
        h, w = it["a"].shape
        cap = it["cap"]
        npyr = sum(x * y for x, y in orbfe.level_dims(w, h)[1:])
        side = []
        for img in (it["a"], it["b"]):
            side.append({"image": torch.from_numpy(img).to(dev), "width": w, "height": h,
                         "pitch": w, "cap": cap,
                         "features": torch.zeros(orbfe.feature_array_bytes(6, cap),
                                                 dtype=torch.uint8, device=dev),
                         "pyramid": torch.zeros(npyr, dtype=torch.float32, device=dev)})
        frames += side
        pairs.append({"prev": side[0]["features"], "curr": side[1]["features"],
                      "prev_match": torch.zeros(6 * cap, dtype=torch.int32, device=dev),
                      "width": w, "height": h, "cap": cap})
    return frames, pairs


def _check_mixed(gold, frames, pairs, pyramids=True):
    for i, g in enumerate(gold):
        fa, fb = frames[2 * i], frames[2 * i + 1]
        assert sha(fa["features"]) == g["features_a"], (i, g["width"], g["cap"])
        assert sha(fb["features"]) == g["features_b"], (i, g["width"], g["cap"])
        if pyramids:
            assert sha(fa["pyramid"]) == g["pyramid_a"], i
            assert sha(fb["pyramid"]) == g["pyramid_b"], i
        assert sha(pairs[i]["prev_match"]) == g["prev_match"], i


@pytest.mark.parametrize("shape", [{}, {"detect_tile_rows": 32, "describe_lanes": 8,
                                        "match_lanes": 32}])
def test_mixed_64_batch_in_graph_matches_golden(torch_gpu, shape):
    torch = torch_gpu
    gold = workloads()["mixed_64"]["pairs"]
    items = synth.mixed_batch(2, n_frames=64)
    frames, pairs = _mixed_buffers(torch, items)
    with orbfe.Context(752, 480, fast_threshold=thr20(), lowe_ratio=RATIO, **shape) as ctx:
        _say(f"mixed batch 64 pairs, shape {shape or 'default'}: prepare")
        b = ctx.batch_create(frames)
        mb = ctx.match_batch_create(pairs)

        def launch(s):
            b.detect_async(stream=s)
            mb.match_async(stream=s)
        g = capture(torch, launch)
        for rep in range(3):
            for f in frames:  # eager work between the replays
                f["features"].fill_(7)
                f["pyramid"].fill_(-1.0)
            for p in pairs:
                p["prev_match"].fill_(-9)
            torch.cuda.synchronize()
            _say(f"graph replay {rep}")
            g.replay()
            torch.cuda.synchronize()
            _check_mixed(gold, frames, pairs, pyramids=rep == 0)
        b.close()
        mb.close()


def test_mixed_64_plain_per_frame_matches_golden(torch_gpu):
    """Every frame of the mixed batch through its own plain context: the
    per-frame path at 640x480 and 1920x1080 with caps up to 4000, including
    the levels whose survivors exceed the cap."""
    torch = torch_gpu
    gold = workloads()["mixed_64"]["pairs"]
    items = synth.mixed_batch(2, n_frames=64)
    frames, pairs = _mixed_buffers(torch, items)
    ctxs = {}
    for i, it in enumerate(items):
        h, w = it["a"].shape
        key = (w, h, it["cap"])
        if key not in ctxs:
            ctxs[key] = orbfe.Context(w, h, features_per_level=it["cap"],
                                      fast_threshold=thr20(), lowe_ratio=RATIO)
        c = ctxs[key]
        fa, fb = frames[2 * i], frames[2 * i + 1]
        c.detect_async(fa["image"], w, fa["features"], pyramid=fa["pyramid"])
        c.detect_async(fb["image"], w, fb["features"], pyramid=fb["pyramid"])
        c.match_pipeline_async(fa["features"], fb["features"], pairs[i]["prev_match"])
    _say(f"plain per-frame: 128 frames, 64 pairs, {len(ctxs)} contexts")
    torch.cuda.synchronize()
    _check_mixed(gold, frames, pairs)
    for c in ctxs.values():
        c.close()


def test_streams_32_served_match_golden(torch_gpu):
    """The serving workload's 32 streams (all four resolutions), 8 frames each
    from their 4-frame pools, arriving on a compressed Poisson schedule:
    every served frame and match equals the oracle's."""
    torch = torch_gpu
    from uni_slam_amd import serve
    gold = workloads()["streams_32x4"]["streams"]
    rng = np.random.Generator(np.random.PCG64(4))
    specs, rates, pool = [], [], []
    for s in range(32):
        w, h = synth.RESOLUTIONS[int(rng.integers(0, 4))]
        specs.append(serve.StreamSpec(w, h, 2000))
        rates.append(float(rng.uniform(20, 60)))
        fr, _ = synth.sequence(1000 + s, w, h, n_frames=4)
        pool.append([torch.from_numpy(f).pin_memory() for f in fr])
    assert [(sp.width, sp.height) for sp in specs] == [(g["width"], g["height"]) for g in gold]
    per = 8
    arr = serve.poisson_arrivals(rates, 10.0, 11)
    arrivals = [a for a in arr if a[2] < per]
    assert len(arrivals) == 32 * per
    srv = serve.Server(specs, fast_threshold=thr20(), lowe_ratio=RATIO, keep_outputs=True,
                       hbm_budget=2 << 30)
    _say("serving 32 streams x 8 frames (compressed schedule)")
    wall, done = serve.run_schedule(srv, arrivals, pool, time_scale=0.0)
    assert len(done) == len(arrivals)
    assert srv.max_batch_seen > 8  # frames of many streams share batches
    for s in range(32):
        for k in range(per):
            f, m = srv.outputs[(s, k)]
            assert sha(f) == gold[s]["features"][k % 4], (s, k)
            if k > 0:
                assert sha(m) == gold[s]["prev_match"][k % 4], (s, k)
