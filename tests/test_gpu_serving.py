"""Batched-in-a-graph and served forms give the plain path's results."""
import numpy as np
import pytest

from tests.conftest import thr20
from uni_slam_amd import orbfe, synth

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def test_graph_replays_equal_eager(torch_gpu):
    torch = torch_gpu
    items = synth.mixed_batch(13, n_frames=4)
    dev = torch.device("cuda:0")
    ctx = orbfe.Context(752, 480, fast_threshold=thr20(), lowe_ratio=0.8)
    frames, pairs = [], []
    for it in items:
        h, w = it["a"].shape
        cap = it["cap"]
        fa = torch.zeros(orbfe.feature_array_bytes(6, cap), dtype=torch.uint8, device=dev)
        fb = torch.zeros_like(fa)
        pm = torch.zeros(6 * cap, dtype=torch.int32, device=dev)
        frames += [{"image": torch.from_numpy(it["a"]).to(dev), "width": w, "height": h,
                    "pitch": w, "cap": cap, "features": fa},
                   {"image": torch.from_numpy(it["b"]).to(dev), "width": w, "height": h,
                    "pitch": w, "cap": cap, "features": fb}]
        pairs.append({"prev": fa, "curr": fb, "prev_match": pm, "width": w, "height": h,
                      "cap": cap})
    b = ctx.batch_create(frames)
    mb = ctx.match_batch_create(pairs)
    _say("eager prepared batch")
    b.detect_async()
    mb.match_async()
    torch.cuda.synchronize()
    want_f = [f["features"].clone() for f in frames]
    want_m = [p["prev_match"].clone() for p in pairs]
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        b.detect_async(stream=side)
        mb.match_async(stream=side)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        cs = torch.cuda.current_stream()
        b.detect_async(stream=cs)
        mb.match_async(stream=cs)
    for rep in range(4):
        _say(f"graph replay {rep}")
        for f in frames:
            f["features"].fill_(7)  # eager work between replays
        for p in pairs:
            p["prev_match"].fill_(-9)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for i, f in enumerate(frames):
            assert torch.equal(f["features"], want_f[i]), (rep, i)
        for i, p in enumerate(pairs):
            assert torch.equal(p["prev_match"], want_m[i]), (rep, i)
    # and the eager results are the plain path's
    f0 = frames[0]
    with orbfe.Context(f0["width"], f0["height"], features_per_level=f0["cap"],
                       fast_threshold=thr20()) as c:
        plain, _ = c.detect(items[0]["a"])
    assert np.array_equal(want_f[0].cpu().numpy(), plain)


def test_served_results_equal_plain(torch_gpu):
    torch = torch_gpu
    from uni_slam_amd import serve
    specs = [serve.StreamSpec(752, 480, 600), serve.StreamSpec(640, 480, 500),
             serve.StreamSpec(1280, 720, 900), serve.StreamSpec(752, 480, 700)]
    n_per = 5
    pool = []
    for s, sp in enumerate(specs):
        fr, _ = synth.sequence(200 + s, sp.width, sp.height, n_frames=n_per)
        pool.append([torch.from_numpy(f).pin_memory() for f in fr])
    # all streams' frames arrive in a burst so batches overlap and mix
    arrivals = sorted((0.001 * k + 0.0001 * s, s, k) for s in range(len(specs))
                      for k in range(n_per))
    srv = serve.Server(specs, fast_threshold=thr20(), lowe_ratio=0.8, keep_outputs=True,
                       max_batch=3, hbm_budget=64 << 20)
    _say("serving 4 streams x 5 frames")
    wall, done = serve.run_schedule(srv, arrivals, pool, time_scale=0.0)
    assert len(done) == len(arrivals)
    assert srv.batches >= 7  # batches of at most 3 frames
    for s, sp in enumerate(specs):
        with orbfe.Context(sp.width, sp.height, features_per_level=sp.cap,
                           fast_threshold=thr20(), lowe_ratio=0.8) as c:
            _say(f"plain stream {s} {sp.width}x{sp.height}")
            prev = None
            for k in range(n_per):
                feats, _ = c.detect(pool[s][k].numpy())
                got_f, got_m = srv.outputs[(s, k)]
                assert np.array_equal(got_f, feats), (s, k)
                if prev is not None:
                    pm, _ = c.match_pipeline(prev, feats)
                    assert np.array_equal(got_m, pm), (s, k)
                prev = feats


@pytest.mark.parametrize("per_stream", [1, 2, 3])
def test_admission_fifo_in_stream_order_within_budget(torch_gpu, per_stream):
    """Admission takes the oldest pending frames, at most per_stream per stream
    and in each stream's order, within the HBM budget; results stay the plain
    path's (a stream's consecutive frames may share a batch)."""
    torch = torch_gpu
    import collections
    from uni_slam_amd import serve
    specs = [serve.StreamSpec(640, 480, 400), serve.StreamSpec(752, 480, 500),
             serve.StreamSpec(640, 480, 600), serve.StreamSpec(1280, 720, 700)]
    n_per = 5
    pool = []
    for s, sp in enumerate(specs):
        fr, _ = synth.sequence(300 + s, sp.width, sp.height, n_frames=n_per)
        pool.append([torch.from_numpy(f).pin_memory() for f in fr])
    scratch = [serve.frame_scratch_bytes(sp.width, sp.height, sp.cap) for sp in specs]
    budget = 3 * max(scratch) + min(scratch)  # room for three to four frames at a time
    # stream 0's frames arrive first (a backlog), then the others interleaved
    arrivals = sorted((0.0001 * k if s == 0 else 0.001 * (k + 1) + 0.0001 * s, s, k)
                      for s in range(len(specs)) for k in range(n_per))
    srv = serve.Server(specs, fast_threshold=thr20(), lowe_ratio=0.8, keep_outputs=True,
                       max_batch=4, hbm_budget=budget, log_batches=True,
                       frames_per_stream=per_stream)
    _say(f"serving 4 streams x {n_per} frames, {per_stream} per stream and batch, "
         f"budget {budget}")
    wall, done = serve.run_schedule(srv, arrivals, pool, time_scale=0.0)
    assert len(done) == len(arrivals)
    log = srv.batch_log
    assert len(log) == srv.batches
    order_of = {(s, k): i for i, (_, s, k) in enumerate(arrivals)}
    next_seq = [0] * len(specs)
    admitted = set()
    shared = 0
    for batch in log:
        cnt = collections.Counter(s for s, _, _ in batch)
        assert max(cnt.values()) <= per_stream, batch
        shared += max(cnt.values()) > 1
        assert len(batch) <= 4
        assert sum(scratch[s] for s, _, _ in batch) <= budget, batch
        orders = [o for _, _, o in batch]
        assert orders == sorted(orders), batch                # oldest first
        for s, k, o in batch:
            assert k == next_seq[s], (batch, next_seq)        # each stream in order
            assert o == order_of[(s, k)]
            next_seq[s] += 1
            admitted.add((s, k))
        # FIFO: an older frame left out waits for an earlier frame of its
        # stream, or its stream reached its per-batch limit
        for (s2, k2), o2 in order_of.items():
            if o2 < max(orders) and (s2, k2) not in admitted:
                assert k2 > next_seq[s2] or cnt[s2] == per_stream, (batch, (s2, k2))
    if per_stream > 1:
        assert shared > 0  # some batch held consecutive frames of one stream
    assert next_seq == [n_per] * len(specs)
    for s, sp in enumerate(specs):
        with orbfe.Context(sp.width, sp.height, features_per_level=sp.cap,
                           fast_threshold=thr20(), lowe_ratio=0.8) as c:
            prev = None
            for k in range(n_per):
                feats, _ = c.detect(pool[s][k].numpy())
                got_f, got_m = srv.outputs[(s, k)]
                assert np.array_equal(got_f, feats), (s, k)
                if prev is not None:
                    pm, _ = c.match_pipeline(prev, feats)
                    assert np.array_equal(got_m, pm), (s, k)
                prev = feats
