"""GPU parity of the feature-detect stage: liborbfe (HIP, through the C ABI)
against the oracle, bit for bit, on the same seeded inputs."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import thr20
from uni_slam_amd import orbfe, synth

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _oracle(img, W, H, cap, thr, border_zero=0):
    p = O.params(W, H, threshold=thr, cap=cap, border_zero=border_zero)
    return O.detect_frame(img, p, with_pyramid=True)


def _assert_same(gpu_feats, gpu_pyr, ref_feats, ref_pyr, levels, cap):
    if gpu_pyr is not None:
        assert np.array_equal(gpu_pyr.view(np.uint32), ref_pyr.view(np.uint32)), "pyramid"
    if not np.array_equal(gpu_feats, ref_feats):
        g = orbfe.parse_feature_array(gpu_feats, levels, cap)
        r = orbfe.parse_feature_array(ref_feats, levels, cap)
        for lod in range(levels):
            assert g[lod][0] == r[lod][0], f"level {lod} count {g[lod][0]} != {r[lod][0]}"
            for i, (a, b) in enumerate(zip(g[lod][1], r[lod][1])):
                assert a.tobytes() == b.tobytes(), f"level {lod} feature {i}: {a} != {b}"
        raise AssertionError("feature arrays differ outside the records")


@pytest.mark.parametrize("W,H,seed", [(64, 48, 0), (100, 37, 1), (131, 97, 2), (17, 13, 3),
                                      (200, 150, 4)])
@pytest.mark.parametrize("border_zero", [0, 1])
def test_detect_small_frames(torch_gpu, W, H, seed, border_zero):
    a, _, _ = synth.frame_pair(seed, W, H)
    cap = max(W * H // 250, 8)
    for thr in (thr20(), 0.15):
        _say(f"detect {W}x{H} seed={seed} thr={thr:.4f} border_zero={border_zero}")
        with orbfe.Context(W, H, features_per_level=cap, fast_threshold=thr,
                           border_zero=border_zero) as ctx:
            feats, pyr = ctx.detect(a, with_pyramid=True)
        rf, rp = _oracle(a, W, H, cap, thr, border_zero)
        _assert_same(feats, pyr, rf, rp, 6, cap)


def test_detect_pair_752x480(torch_gpu):
    a, b, _ = synth.frame_pair(0, 752, 480)
    with orbfe.Context(752, 480, features_per_level=1000, fast_threshold=thr20()) as ctx:
        for name, img in (("A", a), ("B", b)):
            _say(f"detect pair frame {name} 752x480 cap 1000")
            feats, pyr = ctx.detect(img, with_pyramid=True)
            rf, rp = _oracle(img, 752, 480, 1000, thr20())
            _assert_same(feats, pyr, rf, rp, 6, 1000)


def test_detect_reference_default_cap_and_threshold(torch_gpu):
    a, _, _ = synth.frame_pair(7, 752, 480)
    _say("detect 752x480 reference defaults (threshold 0.15, cap W*H/250)")
    with orbfe.Context(752, 480) as ctx:
        assert ctx.cap == 1443
        feats, pyr = ctx.detect(a, with_pyramid=True)
    rf, rp = _oracle(a, 752, 480, 1443, 0.15)
    _assert_same(feats, pyr, rf, rp, 6, 1443)


def test_detect_sequence_1280x720(torch_gpu):
    frames, _ = synth.sequence(11, 1280, 720, n_frames=2)
    with orbfe.Context(1280, 720, features_per_level=2000, fast_threshold=thr20()) as ctx:
        for t in range(2):
            _say(f"detect sequence frame {t} 1280x720 cap 2000")
            feats, pyr = ctx.detect(frames[t], with_pyramid=True)
            rf, rp = _oracle(frames[t], 1280, 720, 2000, thr20())
            _assert_same(feats, pyr, rf, rp, 6, 2000)


def test_detect_async_device_pointers_and_pitch(torch_gpu):
    torch = torch_gpu
    W, H, pitch = 320, 240, 384
    a, _, _ = synth.frame_pair(21, W, H)
    padded = np.full((H, pitch), 255, np.uint8)
    padded[:, :W] = a
    with orbfe.Context(W, H, features_per_level=300, fast_threshold=thr20()) as ctx:
        _say(f"detect_async {W}x{H} pitch {pitch}")
        d_img = torch.from_numpy(padded).cuda()
        d_feat = torch.zeros(ctx.feature_bytes, dtype=torch.uint8, device="cuda")
        d_pyr = torch.zeros(ctx.pyramid_floats, dtype=torch.float32, device="cuda")
        stream = torch.cuda.current_stream()
        ctx.detect_async(d_img, pitch, d_feat, d_pyr, stream=stream)
        torch.cuda.synchronize()
        feats = d_feat.cpu().numpy()
        pyr = d_pyr.cpu().numpy()
        # no pyramid export: same features
        d_feat2 = torch.zeros_like(d_feat)
        ctx.detect_async(d_img, pitch, d_feat2, None, stream=stream)
        torch.cuda.synchronize()
        assert torch.equal(d_feat, d_feat2)
    rf, rp = _oracle(a, W, H, 300, thr20())
    _assert_same(feats, pyr, rf, rp, 6, 300)


def test_detect_is_deterministic(torch_gpu):
    a, _, _ = synth.frame_pair(3, 752, 480)
    with orbfe.Context(752, 480, features_per_level=1000, fast_threshold=thr20()) as ctx:
        _say("detect determinism 752x480 x3")
        outs = [ctx.detect(a, with_pyramid=True) for _ in range(3)]
    for f, p in outs[1:]:
        assert np.array_equal(f, outs[0][0]) and np.array_equal(p, outs[0][1])


def test_detect_batch_mixed_resolutions_equals_plain(torch_gpu):
    torch = torch_gpu
    batch = synth.mixed_batch(5, n_frames=8)
    ctx = orbfe.Context(752, 480, fast_threshold=thr20())
    frames, plain = [], []
    for i, item in enumerate(batch):
        h, w = item["a"].shape
        cap = item["cap"]
        img = torch.from_numpy(item["a"]).cuda()
        fb = orbfe.feature_array_bytes(6, cap)
        npyr = sum(x * y for x, y in orbfe.level_dims(w, h)[1:])
        frames.append({"image": img, "width": w, "height": h, "pitch": w, "cap": cap,
                       "features": torch.zeros(fb, dtype=torch.uint8, device="cuda"),
                       "pyramid": torch.zeros(npyr, dtype=torch.float32, device="cuda")})
        with orbfe.Context(w, h, features_per_level=cap, fast_threshold=thr20()) as c1:
            _say(f"plain frame {i} {w}x{h} cap {cap}")
            plain.append(c1.detect(item["a"], with_pyramid=True))
    _say("dynamic batch of 8 mixed frames")
    ctx.detect_batch_async(frames)
    torch.cuda.synchronize()
    for i, f in enumerate(frames):
        assert np.array_equal(f["features"].cpu().numpy(), plain[i][0]), i
        assert np.array_equal(f["pyramid"].cpu().numpy(), plain[i][1]), i
        f["features"].zero_()
        f["pyramid"].zero_()
    _say("prepared batch of 8 mixed frames")
    b = ctx.batch_create(frames)
    b.detect_async()
    torch.cuda.synchronize()
    for i, f in enumerate(frames):
        assert np.array_equal(f["features"].cpu().numpy(), plain[i][0]), i
    b.close()
    ctx.close()
    # spot-check one frame against the oracle
    item = batch[0]
    h, w = item["a"].shape
    rf, rp = _oracle(item["a"], w, h, item["cap"], thr20())
    _assert_same(plain[0][0], plain[0][1], rf, rp, 6, item["cap"])


@pytest.mark.parametrize("rows,lanes", [(8, 8), (8, 32), (32, 8), (32, 16), (16, 32)])
def test_launch_shapes_match_oracle(torch_gpu, rows, lanes):
    """Every detect tile height and describe lane count gives the oracle's
    features and pyramid, on frames smaller and larger than one tile."""
    for (W, H, seed, bz) in ((17, 13, 3, 0), (131, 97, 2, 1), (752, 480, 5, 0)):
        a, _, _ = synth.frame_pair(seed, W, H)
        cap = max(W * H // 250, 8)
        _say(f"detect {W}x{H} tile_rows={rows} lanes={lanes} border_zero={bz}")
        with orbfe.Context(W, H, features_per_level=cap, fast_threshold=thr20(),
                           border_zero=bz, detect_tile_rows=rows, describe_lanes=lanes) as ctx:
            feats, pyr = ctx.detect(a, with_pyramid=True)
        rf, rp = _oracle(a, W, H, cap, thr20(), bz)
        _assert_same(feats, pyr, rf, rp, 6, cap)


@pytest.mark.parametrize("radius,min_run", [(0, 9), (2, 9), (5, 9), (8, 12), (3, 7), (3, 5)])
def test_orientation_radius_and_min_run_match_oracle(torch_gpu, radius, min_run):
    """Radii up to 3 read the staged neighbourhood in the describe kernel,
    larger ones the level itself; min_run below 9 turns off the detect
    kernel's two-test quick reject.  Both sides of each switch agree with
    the oracle."""
    for (W, H, seed, bz) in ((131, 97, 2, 0), (200, 150, 4, 1)):
        a, _, _ = synth.frame_pair(seed, W, H)
        cap = max(W * H // 250, 8)
        _say(f"detect {W}x{H} radius={radius} min_run={min_run} border_zero={bz}")
        with orbfe.Context(W, H, features_per_level=cap, fast_threshold=thr20(),
                           border_zero=bz, orientation_radius=radius,
                           fast_min_run=min_run) as ctx:
            feats, pyr = ctx.detect(a, with_pyramid=True)
        p = O.params(W, H, threshold=thr20(), cap=cap, border_zero=bz, radius=radius,
                     min_run=min_run)
        rf, rp = O.detect_frame(a, p, with_pyramid=True)
        _assert_same(feats, pyr, rf, rp, 6, cap)


@pytest.mark.parametrize("W,H,cap,flat", [(752, 480, 1000, False), (1280, 720, 300, False),
                                          (640, 480, 4000, False), (200, 150, 120, True)])
def test_export_features_zero_copy_equals_load_features(torch_gpu, W, H, cap, flat):
    """orbfe_export_features_async writes the FeatureSet (min(count, cap)
    wslam::Feature records per level, counts) straight into pinned host
    memory: equal to the oracle's loadFeatures of the oracle's feature array,
    with levels over their cap (cap 300) and with empty levels (flat frame)."""
    torch = torch_gpu
    a, _, _ = synth.frame_pair(7, W, H)
    if flat:
        a[:] = 128
    with orbfe.Context(W, H, features_per_level=cap, fast_threshold=thr20()) as ctx:
        img = torch.from_numpy(a).cuda()
        feats = torch.zeros(ctx.feature_bytes, dtype=torch.uint8, device="cuda")
        out = torch.full((6 * cap * 13,), 0xAB, dtype=torch.int32).pin_memory()
        counts = torch.full((6,), -1, dtype=torch.int32).pin_memory()
        _say(f"export {W}x{H} cap {cap} flat={flat}")
        ctx.detect_async(img, W, feats)
        ctx.export_features_async(feats, out, counts)
        torch.cuda.synchronize()
        got_counts = [int(c) for c in counts.numpy()]
        got = out.numpy().view(np.uint8).view(orbfe.FEATURE_DTYPE)[:sum(got_counts)]
        host = feats.cpu().numpy()
    rf, _ = O.detect_frame(a, O.params(W, H, threshold=thr20(), cap=cap))
    assert np.array_equal(host, rf)
    want, want_counts = O.load_features(rf, 6, cap)
    assert got_counts == want_counts
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))
    if flat:
        assert sum(got_counts) == 0
    else:
        assert min(want_counts) > 0
