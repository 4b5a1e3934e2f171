"""Writes the golden fixtures of tests/golden/ from the oracle.

* golden.json: SHA-256 of the seeded inputs and of the oracle's outputs at
  the benchmark configurations (config-0 pair at 752x480 cap 1000 threshold
  20/255; frame 0 and 1 of the shifted 1280x720 sequence at cap 2000; the
  pipeline match of each pair with ratio 0.8), plus per-level counts.
* small.npz: a 96x64 pair with every output in full (feature arrays,
  pyramids, match map), small enough to diff by eye.

Run: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from uni_slam_amd import orbfe, synth  # noqa: E402

THR20 = float(np.float32(20) / np.float32(255))


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def run_pair(a, b, W, H, cap, ratio):
    p = O.params(W, H, threshold=THR20, cap=cap)
    fa, pa = O.detect_frame(a, p, with_pyramid=True)
    fb, pb = O.detect_frame(b, p, with_pyramid=True)
    A, ca = O.load_features(fa, 6, cap)
    B, cb = O.load_features(fb, 6, cap)
    pm, n = O.find_matches(A, ca, B, cb, W, H, ratio)
    pm = O.prev_match_to_padded(pm, ca, cap)
    return fa, pa, fb, pb, pm, n


def entry(a, b, W, H, cap, ratio):
    fa, pa, fb, pb, pm, n = run_pair(a, b, W, H, cap, ratio)
    counts = [c for c, _ in orbfe.parse_feature_array(fa, 6, cap)]
    return {
        "width": W, "height": H, "cap": cap, "threshold_u8": 20, "ratio": ratio,
        "input_a": sha(a), "input_b": sha(b),
        "features_a": sha(fa), "features_b": sha(fb),
        "pyramid_a": sha(pa), "pyramid_b": sha(pb),
        "prev_match": sha(pm), "n_matches": n, "counts_a": counts,
    }


def main():
    out = {}
    a, b, shift = synth.frame_pair(0, 752, 480)
    out["pair_752x480"] = entry(a, b, 752, 480, 1000, 0.8)
    out["pair_752x480"]["shift"] = list(shift)
    frames, _ = synth.sequence(1, 1280, 720, n_frames=2)
    out["seq_1280x720_frames01"] = entry(frames[0], frames[1], 1280, 720, 2000, 0.8)
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)

    a, b, _ = synth.frame_pair(42, 96, 64)
    fa, pa, fb, pb, pm, n = run_pair(a, b, 96, 64, 24, 0.8)
    np.savez_compressed(os.path.join(HERE, "small.npz"), a=a, b=b, features_a=fa,
                        features_b=fb, pyramid_a=pa, pyramid_b=pb, prev_match=pm,
                        n_matches=np.int64(n))
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
