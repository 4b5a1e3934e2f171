"""The committed golden fixtures (tests/golden/, written by make_golden.py):
the generator and the oracle must reproduce them on the CPU, and the HIP
path must reproduce the same hashes on the GPU."""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import thr20
from uni_slam_amd import orbfe, synth

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def golden():
    with open(os.path.join(HERE, "golden.json")) as f:
        return json.load(f)


def inputs(name):
    if name == "pair_752x480":
        a, b, _ = synth.frame_pair(0, 752, 480)
    else:
        frames, _ = synth.sequence(1, 1280, 720, n_frames=2)
        a, b = frames[0], frames[1]
    return a, b


@pytest.mark.parametrize("name", ["pair_752x480", "seq_1280x720_frames01"])
def test_generator_and_oracle_reproduce_golden(name):
    g = golden()[name]
    a, b = inputs(name)
    assert sha(a) == g["input_a"] and sha(b) == g["input_b"]
    p = O.params(g["width"], g["height"], threshold=thr20(), cap=g["cap"])
    fa, pa = O.detect_frame(a, p, with_pyramid=True)
    fb, _ = O.detect_frame(b, p)
    assert sha(fa) == g["features_a"] and sha(pa) == g["pyramid_a"]
    assert sha(fb) == g["features_b"]
    A, ca = O.load_features(fa, 6, g["cap"])
    B, cb = O.load_features(fb, 6, g["cap"])
    pm, n = O.find_matches(A, ca, B, cb, g["width"], g["height"], g["ratio"])
    assert n == g["n_matches"]
    assert sha(O.prev_match_to_padded(pm, ca, g["cap"])) == g["prev_match"]


def test_oracle_reproduces_small_fixture():
    z = np.load(os.path.join(HERE, "small.npz"))
    p = O.params(96, 64, threshold=thr20(), cap=24)
    fa, pa = O.detect_frame(z["a"], p, with_pyramid=True)
    assert np.array_equal(fa, z["features_a"]) and np.array_equal(pa, z["pyramid_a"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["pair_752x480", "seq_1280x720_frames01"])
def test_gpu_reproduces_golden(torch_gpu, name):
    g = golden()[name]
    a, b = inputs(name)
    print(f"golden {name} on the GPU", flush=True)
    with orbfe.Context(g["width"], g["height"], features_per_level=g["cap"],
                       fast_threshold=thr20(), lowe_ratio=g["ratio"]) as ctx:
        fa, pa = ctx.detect(a, with_pyramid=True)
        fb, pb = ctx.detect(b, with_pyramid=True)
        pm, n = ctx.match_pipeline(fa, fb)
    assert sha(fa) == g["features_a"] and sha(pa) == g["pyramid_a"]
    assert sha(fb) == g["features_b"] and sha(pb) == g["pyramid_b"]
    assert sha(pm) == g["prev_match"] and n == g["n_matches"]


@pytest.mark.gpu
def test_gpu_reproduces_small_fixture(torch_gpu):
    z = np.load(os.path.join(HERE, "small.npz"))
    print("golden small fixture on the GPU", flush=True)
    with orbfe.Context(96, 64, features_per_level=24, fast_threshold=thr20(),
                       lowe_ratio=0.8) as ctx:
        fa, pa = ctx.detect(z["a"], with_pyramid=True)
        fb, _ = ctx.detect(z["b"])
        pm, n = ctx.match_pipeline(fa, fb)
    assert np.array_equal(fa, z["features_a"]) and np.array_equal(pa, z["pyramid_a"])
    assert np.array_equal(fb, z["features_b"])
    assert np.array_equal(pm, z["prev_match"]) and n == int(z["n_matches"])
