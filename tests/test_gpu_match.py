"""GPU parity of the matchers: the pipeline matcher (FindMatches semantics)
against the oracle, and the brute-force matcher against the oracle / numpy."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import np_ref
from tests.conftest import thr20
from uni_slam_amd import orbfe, synth

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _oracle_pipeline(fa, fb, W, H, cap, ratio):
    A, ca = O.load_features(fa, 6, cap)
    B, cb = O.load_features(fb, 6, cap)
    pm, n = O.find_matches(A, ca, B, cb, W, H, ratio)
    return O.prev_match_to_padded(pm, ca, cap), n


@pytest.mark.parametrize("ratio,lanes", [(0.85, 16), (0.8, 16), (0.8, 32), (0.8, 64)])
def test_match_pipeline_pair_752x480(torch_gpu, ratio, lanes):
    a, b, _ = synth.frame_pair(0, 752, 480)
    with orbfe.Context(752, 480, features_per_level=1000, fast_threshold=thr20(),
                       lowe_ratio=ratio, match_lanes=lanes) as ctx:
        _say(f"detect + match 752x480 pair, ratio {ratio}, {lanes} lanes per query")
        fa, _ = ctx.detect(a)
        fb, _ = ctx.detect(b)
        got, n = ctx.match_pipeline(fa, fb)
    want, wn = _oracle_pipeline(fa, fb, 752, 480, 1000, ratio)
    assert n == wn and n > 100
    assert np.array_equal(got, want)


def test_match_pipeline_sequence_1280x720_device(torch_gpu):
    torch = torch_gpu
    frames, _ = synth.sequence(11, 1280, 720, n_frames=3)
    with orbfe.Context(1280, 720, features_per_level=2000, fast_threshold=thr20(),
                       lowe_ratio=0.8) as ctx:
        feats = []
        for t in range(3):
            _say(f"detect_async 1280x720 frame {t}")
            d_img = torch.from_numpy(frames[t]).cuda()
            d_f = torch.zeros(ctx.feature_bytes, dtype=torch.uint8, device="cuda")
            ctx.detect_async(d_img, 1280, d_f)
            feats.append(d_f)
        outs = []
        for t in range(1, 3):
            _say(f"match_pipeline_async 1280x720 frames {t - 1}->{t}")
            pm = torch.full((6 * ctx.cap,), -7, dtype=torch.int32, device="cuda")
            ctx.match_pipeline_async(feats[t - 1], feats[t], pm)
            outs.append(pm)
        torch.cuda.synchronize()
        for t in range(1, 3):
            fa = feats[t - 1].cpu().numpy()
            fb = feats[t].cpu().numpy()
            want, wn = _oracle_pipeline(fa, fb, 1280, 720, 2000, 0.8)
            got = outs[t - 1].cpu().numpy()
            assert np.array_equal(got, want), t
            assert wn > 100


@pytest.mark.parametrize("seed", range(4))
def test_match_pipeline_synthetic_sets(torch_gpu, seed):
    """Hand-built feature arrays with clustered positions and near-duplicate
    descriptors (ties, ratio rejections, windows, tiny pools)."""
    W, H, cap = 160, 120, 40
    rng = np.random.default_rng(100 + seed)
    dims = orbfe.level_dims(W, H)
    base = rng.integers(0, 2**32, (30, 8), dtype=np.uint64).astype(np.uint32)

    def make():
        buf = np.zeros(orbfe.feature_array_bytes(6, cap), np.uint8)
        for lod, (w, h) in enumerate(dims):
            n = int(rng.integers(0, cap + 10))
            blk = buf[lod * (8 + 48 * cap):(lod + 1) * (8 + 48 * cap)]
            blk[:4].view(np.uint32)[0] = n
            k = min(n, cap)
            ys = np.sort(rng.integers(0, h, k))
            xs = rng.integers(0, w, k)
            o = np.lexsort((xs, ys))
            recs = blk[8:].view(orbfe.GPU_FEATURE_DTYPE)[:k]
            recs["x"] = xs[o].astype(np.float32)
            recs["y"] = ys[o].astype(np.float32)
            recs["strength"] = 12
            src = base[rng.integers(0, 30, k)]
            flips = rng.random((k, 256)) < rng.choice([0.0, 0.03, 0.3], k)[:, None]
            recs["descriptor"] = src ^ np.packbits(flips.astype(np.uint8), axis=1,
                                                   bitorder="little").view(np.uint32)
        return buf

    fa, fb = make(), make()
    for ratio, lanes in ((0.85, 16), (1.0, 16), (1.0, 32), (0.85, 64)):
        with orbfe.Context(W, H, features_per_level=cap, lowe_ratio=ratio,
                           match_lanes=lanes) as ctx:
            _say(f"match synthetic sets seed {seed} ratio {ratio} lanes {lanes}")
            got, n = ctx.match_pipeline(fa, fb)
        want, wn = _oracle_pipeline(fa, fb, W, H, cap, ratio)
        assert np.array_equal(got, want)
        # and the oracle agrees with the independent masked brute force
        A = [x[1] for x in orbfe.parse_feature_array(fa, 6, cap)]
        B = [x[1] for x in orbfe.parse_feature_array(fb, 6, cap)]
        bf = np_ref.find_matches_bf(A, B, W, H, ratio)
        flat = np.full(6 * cap, -1, np.int64)
        for lod in range(6):
            flat[lod * cap:lod * cap + len(bf[lod])] = bf[lod]
        assert np.array_equal(want, flat)


def test_match_pipeline_batch_equals_plain(torch_gpu):
    torch = torch_gpu
    batch = synth.mixed_batch(9, n_frames=6)
    ctx = orbfe.Context(752, 480, fast_threshold=thr20(), lowe_ratio=0.8)
    pairs, plain = [], []
    for i, item in enumerate(batch):
        h, w = item["a"].shape
        cap = item["cap"]
        with orbfe.Context(w, h, features_per_level=cap, fast_threshold=thr20(),
                           lowe_ratio=0.8) as c1:
            _say(f"plain pair {i} {w}x{h} cap {cap}")
            fa, _ = c1.detect(item["a"])
            fb, _ = c1.detect(item["b"])
            plain.append(c1.match_pipeline(fa, fb)[0])
        pairs.append({"prev": torch.from_numpy(fa).cuda(), "curr": torch.from_numpy(fb).cuda(),
                      "prev_match": torch.zeros(6 * cap, dtype=torch.int32, device="cuda"),
                      "width": w, "height": h, "cap": cap})
    _say("dynamic match batch of 6 pairs")
    ctx.match_batch_async(pairs)
    torch.cuda.synchronize()
    for i, p in enumerate(pairs):
        assert np.array_equal(p["prev_match"].cpu().numpy(), plain[i]), i
        p["prev_match"].zero_()
    _say("prepared match batch of 6 pairs")
    mb = ctx.match_batch_create(pairs)
    mb.match_async()
    torch.cuda.synchronize()
    for i, p in enumerate(pairs):
        assert np.array_equal(p["prev_match"].cpu().numpy(), plain[i]), i
    mb.close()
    ctx.close()


def _bf_gpu(torch, ctx, q, t):
    nq, nt = len(q), len(t)
    dq = torch.from_numpy(q).cuda()
    dt = torch.from_numpy(t if nt else np.zeros((1, 8), np.uint32)).cuda()
    idx = torch.zeros(max(nq, 1), dtype=torch.int32, device="cuda")
    d1 = torch.zeros_like(idx)
    d2 = torch.zeros_like(idx)
    ctx.match_bf_async(dq, nq, dt, nt, idx, d1, d2)
    torch.cuda.synchronize()
    u = lambda x: x.cpu().numpy().view(np.uint32)[:nq]
    return u(idx), u(d1), u(d2)


@pytest.mark.parametrize("nq,nt", [(1, 0), (1, 1), (3, 2), (1000, 257), (4096, 4096),
                                   (513, 70000)])
def test_match_bf_against_oracle(torch_gpu, nq, nt):
    q, t, _, _ = synth.descriptors(nq + 7 * nt, nq, nt)
    with orbfe.Context(64, 48) as ctx:
        _say(f"match_bf {nq}x{nt}")
        idx, d1, d2 = _bf_gpu(torch_gpu, ctx, q, t)
    ri, r1, r2 = O.match_bf(q, t)
    assert np.array_equal(idx, ri) and np.array_equal(d1, r1) and np.array_equal(d2, r2)


@pytest.mark.parametrize("nt,dups", [(600, [17, 333, 590]),
                                     (20000, [31, 32, 8191, 8192, 19999])])
def test_match_bf_ties_pick_lowest_index(torch_gpu, nt, dups):
    # duplicates of train 5 in other 32-row tiles and (nt = 20000) in other
    # train splits of at most 8192 rows
    rng = np.random.default_rng(4)
    t = rng.integers(0, 2**32, (nt, 8), dtype=np.uint64).astype(np.uint32)
    t[dups] = t[5]
    q = t[[5] + dups].copy()
    with orbfe.Context(64, 48) as ctx:
        _say(f"match_bf ties {nt}")
        idx, d1, d2 = _bf_gpu(torch_gpu, ctx, q, t)
    k = len(dups) + 1
    assert list(idx) == [5] * k and list(d1) == [0] * k and list(d2) == [0] * k


def test_match_bf_extreme_distances(torch_gpu):
    # distances 0 and 256 (all bits differ); one train row only for a query
    # of all ones besides its complement
    ones = np.full(8, 0xFFFFFFFF, np.uint32)
    zeros = np.zeros(8, np.uint32)
    t = np.stack([ones, ones, zeros, ones])
    q = np.stack([zeros, ones])
    with orbfe.Context(64, 48) as ctx:
        _say("match_bf extremes")
        idx, d1, d2 = _bf_gpu(torch_gpu, ctx, q, t)
        i1, e1, e2 = _bf_gpu(torch_gpu, ctx, q[:1], t[:1])
    assert list(idx) == [2, 0] and list(d1) == [0, 0] and list(d2) == [256, 0]
    assert list(i1) == [0] and list(e1) == [256] and list(e2) == [0xFFFFFFFF]


def _oracle_bf_threaded(q, t, threads=8):
    """O.match_bf over query chunks on a thread pool (ctypes releases the GIL
    during the call, so the chunks run in parallel)."""
    from concurrent.futures import ThreadPoolExecutor
    chunks = np.array_split(np.arange(len(q)), threads)
    with ThreadPoolExecutor(threads) as ex:
        parts = list(ex.map(lambda c: O.match_bf(q[c], t), chunks))
    return tuple(np.concatenate([p[k] for p in parts]) for k in range(3))


@pytest.mark.parametrize("nq", [4096, 16384, 65536])
@pytest.mark.parametrize("nt", [4096, 16384, 65536])
def test_match_bf_stress_sizes_against_oracle(torch_gpu, nq, nt):
    """Every N x M combination of the matcher stress workload (30% planted
    near-duplicates with p_flip 0.05, the rest uniform), compared in full
    with the oracle's brute force: index, best and second-best distance."""
    q, t, planted, src = synth.descriptors(nq * 7 + nt, nq, nt)
    with orbfe.Context(64, 48) as ctx:
        _say(f"match_bf stress {nq}x{nt}")
        idx, d1, d2 = _bf_gpu(torch_gpu, ctx, q, t)
    ri, r1, r2 = _oracle_bf_threaded(q, t)
    assert np.array_equal(idx, ri) and np.array_equal(d1, r1) and np.array_equal(d2, r2)
    pl = np.nonzero(planted)[0]
    assert np.mean(idx[pl] == src[pl]) > 0.999
