"""The launch-shape autotuner: every shape it tries gives the default
shape's features on a mixed batch (the tuner itself raises otherwise), and
the context it recommends reproduces them."""
import pytest

from tests.conftest import thr20
from uni_slam_amd import autotune, orbfe, synth

pytestmark = pytest.mark.gpu


def test_tuner_variants_identical_and_best_reproduces(torch_gpu):
    torch = torch_gpu
    dev = torch.device("cuda:0")
    items = synth.mixed_batch(21, n_frames=5)
    frames = []
    for it in items:
        h, w = it["a"].shape
        cap = it["cap"]
        for img in (it["a"], it["b"]):
            frames.append({"image": torch.from_numpy(img).to(dev), "width": w, "height": h,
                           "pitch": w, "cap": cap,
                           "features": torch.zeros(orbfe.feature_array_bytes(6, cap),
                                                   dtype=torch.uint8, device=dev)})
    print(f"tuning {len(frames)} frames over {len(autotune.SHAPES)} shapes", flush=True)
    res = autotune.tune_detect(frames, 752, 480, reps=2, fast_threshold=thr20())
    assert set(res.seconds) == set(autotune.SHAPES)
    assert res.best in autotune.SHAPES
    assert all(t > 0 for t in res.seconds.values())
    want = [f["features"].clone() for f in frames]
    # plain per-frame launches under the recommended shape agree
    for f, wf in zip(frames, want):
        with orbfe.Context(f["width"], f["height"], features_per_level=f["cap"],
                           fast_threshold=thr20(), **res.config()) as c:
            got, _ = c.detect(f["image"].cpu().numpy())
        assert (got == wf.cpu().numpy()).all()


def test_match_tuner_variants_identical_and_best_reproduces(torch_gpu):
    torch = torch_gpu
    dev = torch.device("cuda:0")
    items = synth.mixed_batch(22, n_frames=4)
    pairs, plain = [], []
    for it in items:
        h, w = it["a"].shape
        cap = it["cap"]
        with orbfe.Context(w, h, features_per_level=cap, fast_threshold=thr20(),
                           lowe_ratio=0.8) as c:
            fa, _ = c.detect(it["a"])
            fb, _ = c.detect(it["b"])
            plain.append(c.match_pipeline(fa, fb)[0])
        pairs.append({"prev": torch.from_numpy(fa).to(dev), "curr": torch.from_numpy(fb).to(dev),
                      "prev_match": torch.zeros(6 * cap, dtype=torch.int32, device=dev),
                      "width": w, "height": h, "cap": cap})
    print(f"tuning {len(pairs)} pairs over {len(autotune.MATCH_LANES)} shapes", flush=True)
    res = autotune.tune_match(pairs, 752, 480, reps=2, fast_threshold=thr20(), lowe_ratio=0.8)
    assert set(res.seconds) == set(autotune.MATCH_LANES)
    assert res.best in autotune.MATCH_LANES
    for p, want in zip(pairs, plain):
        assert (p["prev_match"].cpu().numpy() == want).all()
