"""The C++ pipeline host code (compute/ engine + wslam passes above the C
ABI): the stage machinery on the CPU, and the whole feature path through
Compute/HipStage on the GPU against the oracle."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import ROOT, thr20
from uni_slam_amd import synth

PKG = os.path.join(ROOT, "uni_slam_amd")


def test_stage_machinery_selftest():
    r = subprocess.run([os.path.join(PKG, "stage_selftest")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    assert "stage selftest ok" in r.stdout


def _feature_key(f):
    # wslam::Feature's member-wise order (models.hpp:109)
    return (int(f["x"]), int(f["y"]), int(f["lod"]), int(f["strength"]),
            float(f["orientation"]), tuple(int(d) for d in f["descriptor"]))


@pytest.mark.gpu
def test_pipeline_driver_matches_oracle(tmp_path, torch_gpu):
    W, H, N, cap, ratio = 752, 480, 4, 1000, 0.8
    frames, _ = synth.sequence(3, W, H, n_frames=N)
    raw = tmp_path / "frames.raw"
    frames.tofile(raw)
    out = tmp_path / "out.txt"
    print(f"wslam_pipeline {W}x{H} x{N}", flush=True)
    r = subprocess.run([os.path.join(PKG, "wslam_pipeline"), str(raw), str(W), str(H), str(N),
                        str(cap), repr(thr20()), repr(ratio), str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = out.read_text().splitlines()
    p = O.params(W, H, threshold=thr20(), cap=cap)
    sets = []
    for t in range(N):
        fa, _ = O.detect_frame(frames[t], p)
        sets.append(O.load_features(fa, 6, cap))
    want = []
    for t in range(N):
        want.append(f"frame {t} " + " ".join(str(c) for c in sets[t][1]))
        if t == 0:
            continue
        (A, ca), (B, cb) = sets[t - 1], sets[t]
        pm, _ = O.find_matches(A, ca, B, cb, W, H, ratio)
        offs_b = np.concatenate([[0], np.cumsum(cb)])
        offs_a = np.concatenate([[0], np.cumsum(ca)])
        for l in range(6):
            pairs = []
            for i in range(ca[l]):
                m = int(pm[offs_a[l] + i])
                if m < 0:
                    continue
                b = B[offs_b[m >> 24] + (m & 0xFFFFFF)]
                a = A[offs_a[l] + i]
                pairs.append((_feature_key(b), b, a))
            pairs.sort(key=lambda z: z[0])
            for _, b, a in pairs:
                want.append(f"m {t} {l} {b['x']} {b['y']} {b['lod']} {a['x']} {a['y']}")
    assert lines == want
