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
