"""Writes tests/golden/workloads.json: SHA-256 of the oracle's outputs on the
full benchmark workloads, for the GPU tests that run those workloads.

* seq_1280x720x32: all 32 frames of the shifted 1280x720 sequence
  (synth.sequence(1, 1280, 720, 32)), cap 2000: every frame's feature array
  and pyramid, every consecutive pair's prev_match (ratio 0.8);
* mixed_64: the mixed-resolution batch of 64 frames and their translated
  twins (synth.mixed_batch(2, 64): 640x480 / 752x480 / 1280x720 / 1920x1080,
  caps 1000..4000): feature array and pyramid of both frames, and the pair's
  prev_match;
* streams_32x4: the 32 camera streams of the serving workload (resolution
  drawn with PCG64(4), cap 2000, frames synth.sequence(1000 + s, w, h, 4)),
  every frame's feature array and the match of each frame against the one
  before it in the stream's cyclic frame pool (frame k against k - 1 mod 4).

Every entry also records the per-level survivor counts, so a reader sees
which levels exceed their cap.  Runs the oracle on all CPU cores (about a
minute on 8).  Run: python tests/golden/make_workloads.py
"""
import hashlib
import json
import multiprocessing as mp
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from uni_slam_amd import synth  # noqa: E402

THR20 = float(np.float32(20) / np.float32(255))
RATIO = 0.8
LEVELS = 6


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def serving_streams(seed: int = 4, n_streams: int = 32):
    """(width, height) of each stream of the serving workload, as bench.py
    draws them."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(n_streams):
        w, h = synth.RESOLUTIONS[int(rng.integers(0, 4))]
        rng.uniform(20, 60)  # the stream's frame rate, drawn in the same order
        out.append((w, h))
    return out


def _detect(img, W, H, cap):
    p = O.params(W, H, threshold=THR20, cap=cap)
    f, pyr = O.detect_frame(img, p, with_pyramid=True)
    counts = [int(v) for v in f.view(np.uint32)[0::(8 + 48 * cap) // 4][:LEVELS]]
    return f, pyr, counts


def _match(fa, fb, W, H, cap):
    A, ca = O.load_features(fa, LEVELS, cap)
    B, cb = O.load_features(fb, LEVELS, cap)
    pm, n = O.find_matches(A, ca, B, cb, W, H, RATIO)
    return O.prev_match_to_padded(pm, ca, cap), n


def job_detect(args):
    img, W, H, cap = args
    f, pyr, counts = _detect(img, W, H, cap)
    return f, {"features": sha(f), "pyramid": sha(pyr), "counts": counts}


def job_match(args):
    fa, fb, W, H, cap = args
    pm, n = _match(fa, fb, W, H, cap)
    return {"prev_match": sha(pm), "n_matches": n}


def main():
    O.lib()
    seq_frames, _ = synth.sequence(1, 1280, 720, n_frames=32)
    items = synth.mixed_batch(2, n_frames=64)
    streams = serving_streams()
    stream_frames = [synth.sequence(1000 + s, w, h, n_frames=4)[0]
                     for s, (w, h) in enumerate(streams)]
    with mp.Pool(min(8, os.cpu_count() or 1)) as pool:
        # sequence: every frame, then every consecutive pair
        det = pool.map(job_detect, [(f, 1280, 720, 2000) for f in seq_frames])
        mat = pool.map(job_match, [(det[t - 1][0], det[t][0], 1280, 720, 2000)
                                   for t in range(1, 32)])
        seq = []
        for t in range(32):
            e = dict(det[t][1])
            if t > 0:
                e.update(mat[t - 1])
            seq.append(e)
        # mixed batch: frame a and its twin b, then the pair
        jobs = []
        for it in items:
            h, w = it["a"].shape
            jobs += [(it["a"], w, h, it["cap"]), (it["b"], w, h, it["cap"])]
        det = pool.map(job_detect, jobs)
        mat = pool.map(job_match, [(det[2 * i][0], det[2 * i + 1][0]) + jobs[2 * i][1:]
                                   for i in range(len(items))])
        mixed = []
        for i, it in enumerate(items):
            h, w = it["a"].shape
            a, b = det[2 * i][1], det[2 * i + 1][1]
            mixed.append({"width": w, "height": h, "cap": it["cap"],
                          "features_a": a["features"], "pyramid_a": a["pyramid"],
                          "counts_a": a["counts"], "features_b": b["features"],
                          "pyramid_b": b["pyramid"], "counts_b": b["counts"], **mat[i]})
        # serving streams: 4 frames each, frame k against k - 1 of the cyclic pool
        jobs = [(stream_frames[s][k], w, h, 2000) for s, (w, h) in enumerate(streams)
                for k in range(4)]
        det = pool.map(job_detect, jobs)
        mat = pool.map(job_match, [(det[4 * s + (k - 1) % 4][0], det[4 * s + k][0], w, h, 2000)
                                   for s, (w, h) in enumerate(streams) for k in range(4)])
        srv = []
        for s, (w, h) in enumerate(streams):
            d = [det[4 * s + k][1] for k in range(4)]
            m = [mat[4 * s + k] for k in range(4)]
            srv.append({"width": w, "height": h, "cap": 2000,
                        "features": [x["features"] for x in d],
                        "counts": [x["counts"] for x in d],
                        "prev_match": [x["prev_match"] for x in m],
                        "n_matches": [x["n_matches"] for x in m]})
    out = {
        "threshold_u8": 20, "ratio": RATIO,
        "seq_1280x720x32": {"cap": 2000, "frames": seq},
        "mixed_64": {"pairs": mixed},
        "streams_32x4": {"streams": srv},
    }
    with open(os.path.join(HERE, "workloads.json"), "w") as f:
        json.dump(out, f, indent=0, sort_keys=True)
    over = sum(1 for e in mixed for side in "ab" for c in e["counts_" + side] if c > e["cap"])
    print(f"wrote workloads.json: 32 sequence frames, 64 mixed pairs ({over} levels over "
          f"their cap), 32 streams x 4 frames")


if __name__ == "__main__":
    main()
