"""Pins of the oracle (the CPU restatement the HIP path is checked against).

Every oracle function is checked against something that does not share its
code: closed forms and special cases, invariants, independent float64 numpy
restatements (tests/np_ref.py), the corner-cull counts of the reference's
dispatch geometry, and the reference's own BRIEF table.
"""
import math
import os
import re

import numpy as np
import pytest

from oracle import oracle as O
from tests import np_ref
from tests.conftest import REFERENCE, thr20
from uni_slam_amd import synth


# --- input contract ------------------------------------------------------------
def test_lod0_is_correctly_rounded_division():
    img = np.arange(256, dtype=np.uint8).reshape(16, 16)
    got = O.lod0(img)
    want = np.arange(256, dtype=np.float32) / np.float32(255)
    assert np.array_equal(got.ravel().view(np.uint32), want.view(np.uint32))
    assert got.ravel()[0] == 0.0 and got.ravel()[255] == 1.0


@pytest.mark.parametrize("W,H,dims", [
    (752, 480, [(752, 480), (626, 400), (522, 333), (435, 277), (362, 231), (302, 192)]),
])
def test_level_dims_known(W, H, dims):
    assert O.level_dims(W, H) == dims


@pytest.mark.parametrize("W,H", synth.RESOLUTIONS)
def test_level_dims_pow_equals_repeated_multiply(W, H):
    # common.cpp:38-45 (pow) vs common.hpp:79-93 (repeated multiply)
    rep = []
    s = 1.0
    for _ in range(6):
        rep.append((int(W / s), int(H / s)))
        s *= 1.2
    assert O.level_dims(W, H) == rep


# --- pyramid ---------------------------------------------------------------------
def test_pyramid_constant_image_is_constant():
    for v in (0, 1, 77, 255):
        lod0 = O.lod0(np.full((48, 64), v, np.uint8))
        for (w, h) in O.level_dims(64, 48)[1:]:
            lvl = O.pyramid_level(lod0, w, h)
            assert np.all(lvl == lod0[0, 0])


def test_pyramid_matches_float64_bilinear():
    rng = np.random.default_rng(1)
    img = rng.integers(0, 256, (97, 131), dtype=np.uint8)
    lod0 = O.lod0(img)
    for (w, h) in O.level_dims(131, 97)[1:]:
        got = O.pyramid_level(lod0, w, h)
        want = np_ref.pyramid_level64(lod0, w, h)
        # f32 texel coordinates carry ~1e-7 relative error: at a gradient of
        # one unit per texel that is ~1e-5; a half-texel slip would be ~0.5.
        assert np.max(np.abs(got - want)) < 5e-5


def test_pyramid_horizontal_ramp_closed_form():
    W, H = 200, 40
    img = np.tile(np.arange(W, dtype=np.uint8), (H, 1))
    lod0 = O.lod0(img)
    w, h = 166, 33
    got = O.pyramid_level(lod0, w, h)
    # Interior texels: value = (uv*W - 0.5)/255 on the ramp I(x) = x/255.
    u = ((np.arange(w) + 0.5) / w) * W - 0.5
    inner = (u >= 0) & (u <= W - 1)
    want = u / 255.0
    assert np.max(np.abs(got[:, inner] - want[None, inner])) < 5e-5
    assert np.all(got[:, 0] >= 0)


# --- detect ----------------------------------------------------------------------
def test_detect_bright_dot_scores_17():
    img = np.zeros((32, 32), np.float32)
    img[16, 16] = 1.0
    S = O.detect(img, thr20())
    assert S[16, 16] == 17
    # ring pixels see one differing neighbour at most: no corner
    assert S.sum() == 17


def test_detect_step_edge_is_not_a_corner():
    img = np.zeros((32, 32), np.float32)
    img[:, 16:] = 1.0
    S = O.detect(img, thr20())
    assert S[10:22, 10:22].max() == 0  # an edge differs on <= 7 ring pixels


def test_detect_ring_does_not_wrap():
    # Differences on ring indices 12..15 and 0..4 are 9 in a circular run
    # but runs of 4 and 5 in the shader's non-wrapping scan: no corner.
    img = np.full((32, 32), 0.5, np.float32)
    cx, cy = 16, 16
    for i in list(range(12, 16)) + list(range(0, 5)):
        img[cy + np_ref.CIRCLE_Y[i], cx + np_ref.CIRCLE_X[i]] = 0.0
    S = O.detect(img, thr20())
    assert S[cy, cx] == 0
    # the same 9 differences contiguous without wrap (0..8) give streak 10
    img = np.full((32, 32), 0.5, np.float32)
    for i in range(0, 9):
        img[cy + np_ref.CIRCLE_Y[i], cx + np_ref.CIRCLE_X[i]] = 0.0
    assert O.detect(img, thr20())[cy, cx] == 10


def test_detect_threshold_is_strict():
    t = np.float32(0.25)
    img = np.full((16, 16), 0.5, np.float32)
    img[4:12, 4:12] = 0.5 + t  # |diff| == t exactly is not different
    assert O.detect(img, float(t)).max() == 0


def test_detect_far_edge_clamp_and_zero_mode():
    # Pixel (0, 8): ring offsets at x-1..x-3 wrap to the last column.
    img = np.zeros((16, 20), np.float32)
    img[:, -1] = 1.0
    S = O.detect(img, thr20())
    # ring indices 9..15 have dx < 0 -> 7 reads of column w-1 (value 1)
    assert S[8, 0] == 0  # run of 7 is below 9
    img[:, 3] = 1.0  # add the dx=+3 side (indices 3,4,5)
    S = O.detect(img, thr20())
    assert np.array_equal(S, np_ref.fast_scores(img, thr20()))
    Sz = O.detect(img, thr20(), border_zero=1)
    assert np.array_equal(Sz, np_ref.fast_scores(img, thr20(), border_zero=True))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_detect_matches_vectorised_restatement(seed):
    a, _, _ = synth.frame_pair(seed, 96, 64)
    lvl = O.lod0(a)
    for bz in (0, 1):
        assert np.array_equal(O.detect(lvl, thr20(), 9, bz),
                              np_ref.fast_scores(lvl, thr20(), 9, bool(bz)))


# --- cull ------------------------------------------------------------------------
def _cull_all(S_levels, W0, H0, dims):
    N0 = W0 * H0
    S_all = np.zeros((len(dims), N0), np.uint32)
    for l, (w, h) in enumerate(dims):
        S_all[l, :w * h] = S_levels[l].ravel()
    return O.cull(S_all, W0, H0, dims)


def test_cull_mask_counts_of_the_dispatch_geometry():
    # All-equal positive scores: survivors are exactly the statically
    # reachable positions of a 752x480 frame.  Rows >= 1 of LOD0 keep
    # 122,359 positions and levels 1..5 keep 268,137 (counts derived
    # independently of this code from the dispatch geometry).  Row 0 keeps
    # the first maximum of each 3-wide window at a multiple of 9 (84 windows
    # for w = 752); every such start is a multiple of 3 and so is copied by
    # main_vertical.
    W0, H0 = 752, 480
    dims = O.level_dims(W0, H0)
    S = [np.full((h, w), 10, np.uint32) for (w, h) in dims]
    O_all = _cull_all(S, W0, H0, dims)
    counts = [int((O_all[l, :w * h] > 0).sum()) for l, (w, h) in enumerate(dims)]
    row0 = O_all[0, :W0]
    assert int((row0 > 0).sum()) == len(range(0, W0, 9))
    assert np.array_equal(np.nonzero(row0)[0], np.arange(0, W0, 9))
    assert counts[0] - len(range(0, W0, 9)) == 122359
    assert sum(counts[1:]) == 268137


@pytest.mark.parametrize("W0,H0", [(752, 480), (64, 48), (131, 97), (640, 480)])
def test_cull_simulation_equals_closed_form(W0, H0):
    rng = np.random.default_rng(W0 * 7 + H0)
    dims = O.level_dims(W0, H0)
    S = []
    for (w, h) in dims:
        s = rng.integers(10, 18, (h, w)).astype(np.uint32)
        s[rng.random((h, w)) < 0.6] = 0
        S.append(s)
    O_all = _cull_all(S, W0, H0, dims)
    for l, (w, h) in enumerate(dims):
        got = O_all[l, :w * h].reshape(h, w)
        assert np.array_equal(got, np_ref.cull_closed_form(S[l], W0, H0, w, h)), l
        assert not O_all[l, w * h:].any()


# --- generate ----------------------------------------------------------------------
def test_generate_flat_images_give_zero_descriptor_and_angle():
    for v in (0.0, 1.0):
        lvl = np.full((40, 40), v, np.float32)
        Ob = np.zeros((40, 40), np.uint32)
        Ob[20, 20] = 12
        n, f = O.generate(lvl, Ob, 10)
        assert n == 1
        assert f[0]["orientation"] == 0.0
        assert not f[0]["descriptor"].any()
        assert (f[0]["x"], f[0]["y"], f[0]["strength"]) == (20.0, 20.0, 12)


def test_generate_raster_order_and_cap():
    rng = np.random.default_rng(3)
    lvl = rng.random((30, 50)).astype(np.float32)
    Ob = np.zeros((30, 50), np.uint32)
    pts = [(5, 7), (40, 3), (2, 20), (33, 20), (10, 25)]
    for x, y in pts:
        Ob[y, x] = 11
    n, f = O.generate(lvl, Ob, 3)
    assert n == 5 and len(f) == 3
    order = sorted(pts, key=lambda p: (p[1], p[0]))
    assert [(int(a), int(b)) for a, b in zip(f["x"], f["y"])] == order[:3]


def test_generate_matches_float64_brief():
    a, _, _ = synth.frame_pair(5, 160, 120)
    lvl = O.lod0(a)
    h, w = lvl.shape
    pat = _pattern()
    rng = np.random.default_rng(0)
    Ob = np.zeros((h, w), np.uint32)
    pts = [(int(x), int(y)) for x, y in zip(rng.integers(0, w, 40), rng.integers(0, h, 40))]
    pts += [(0, 0), (w - 1, h - 1), (1, h // 2), (w // 2, 2)]
    for x, y in pts:
        Ob[y, x] = 12
    n, feats = O.generate(lvl, Ob, 100)
    checked = 0
    for f in feats:
        x, y = int(f["x"]), int(f["y"])
        theta = float(f["orientation"])
        want, bound = np_ref.orientation64(lvl, x, y)
        assert abs(math.remainder(theta - want, 2 * math.pi)) < bound
        c, s = math.cos(theta), math.sin(theta)
        for i in range(256):
            ax, ay, bx, by = (int(v) for v in pat[i])
            rot = [ax * c - ay * s, ax * s + ay * c, bx * c - by * s, bx * s + by * c]
            if any(abs(abs(r - math.floor(r)) - 0.5) < 1e-4 for r in rot):
                continue  # rounding boundary: f32 and f64 may round differently
            r = [round(v) for v in rot]
            A = (np_ref.u32(x + r[0]), np_ref.u32(y + r[1]))
            B = (np_ref.u32(x + r[2]), np_ref.u32(y + r[3]))
            ga = np_ref.gaussian_tap9_64(lvl, *A)
            gb = np_ref.gaussian_tap9_64(lvl, *B)
            if abs(ga - gb) < 1e-5:
                continue
            bit = (int(f["descriptor"][i // 32]) >> (i % 32)) & 1
            assert bit == (1 if ga > gb else 0), (x, y, i)
            checked += 1
    assert checked > 256 * len(feats) * 0.9


def _pattern():
    from uni_slam_amd.brief import brief_pattern
    return brief_pattern()


@pytest.mark.skipif(not os.path.exists(REFERENCE), reason="reference checkout not present")
def test_brief_table_equals_reference():
    src = open(os.path.join(REFERENCE, "wslam/src/feature_detect/brief_tests.inc")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    ref = np.array([int(t) for t in re.findall(r"-?\d+", src)], np.int64).reshape(256, 4)
    assert np.array_equal(_pattern(), ref)


# --- own trig -------------------------------------------------------------------------
def _ulp(x):
    return float(np.spacing(np.float32(abs(x)) if x != 0 else np.float32(1e-38)))


def test_atan2_accuracy_and_special_cases():
    rng = np.random.default_rng(7)
    ys = rng.normal(size=20000).astype(np.float32) * np.float32(3.0)
    xs = rng.normal(size=20000).astype(np.float32) * np.float32(3.0)
    worst = 0.0
    for y, x in zip(ys, xs):
        got = O.atan2f(float(y), float(x))
        want = math.atan2(float(y), float(x))
        worst = max(worst, abs(got - want) / _ulp(want))
    assert worst <= 4.0
    assert O.atan2f(0.0, 0.0) == 0.0
    assert O.atan2f(0.0, -0.0) == pytest.approx(math.pi, abs=1e-6)
    assert O.atan2f(1.0, 0.0) == pytest.approx(math.pi / 2, abs=1e-7)
    assert O.atan2f(-1.0, 0.0) == pytest.approx(-math.pi / 2, abs=1e-7)
    assert O.atan2f(0.0, -1.0) == pytest.approx(math.pi, abs=1e-6)
    assert O.atan2f(1.0, 1.0) == pytest.approx(math.pi / 4, abs=1e-7)


def test_sincos_accuracy():
    xs = np.linspace(-math.pi, math.pi, 20001).astype(np.float32)
    worst = 0.0
    for x in xs:
        s, c = O.sincosf(float(x))
        worst = max(worst, abs(s - math.sin(float(x))), abs(c - math.cos(float(x))))
        assert s * s + c * c == pytest.approx(1.0, abs=1e-6)
    assert worst < 2e-7
    assert O.sincosf(0.0) == (0.0, 1.0)


# --- matcher ----------------------------------------------------------------------------
def test_hamming_bounded():
    rng = np.random.default_rng(11)
    for _ in range(2000):
        a = rng.integers(0, 2**32, 8, dtype=np.uint64).astype(np.uint32)
        b = rng.integers(0, 2**32, 8, dtype=np.uint64).astype(np.uint32)
        full = int(np_ref.popcount_u32((a ^ b)[None])[0])
        lower = int(np_ref.popcount_u32((a ^ b)[None, :4])[0])
        cap = int(rng.integers(0, 300))
        got = O.hamming_bounded(a, b, cap)
        assert got == (cap if lower >= cap else full)


@pytest.mark.parametrize("nq,nt", [(1, 0), (1, 1), (5, 2), (64, 300), (200, 129)])
def test_match_bf_against_numpy(nq, nt):
    q, t, _, _ = synth.descriptors(nq * 31 + nt, nq, nt)
    idx, d1, d2 = O.match_bf(q, t)
    if nt == 0:
        assert np.all(idx == 0xFFFFFFFF) and np.all(d1 == 0xFFFFFFFF)
        return
    D = np_ref.hamming_matrix(q, t)
    want_idx = D.argmin(1)  # first minimum = lowest index
    assert np.array_equal(idx, want_idx)
    assert np.array_equal(d1, D.min(1))
    if nt >= 2:
        assert np.array_equal(d2, np.sort(D, 1)[:, 1])
    else:
        assert np.all(d2 == 0xFFFFFFFF)


def _small_sets(seed, W0=160, H0=120, per_level=(30, 25, 20, 15, 10, 6)):
    """Feature sets with clustered positions and related descriptors so that
    windows, ties and ratio rejections all occur."""
    rng = np.random.default_rng(seed)
    base = rng.integers(0, 2**32, (60, 8), dtype=np.uint64).astype(np.uint32)
    sets_a, sets_b = [], []
    dims = O.level_dims(W0, H0)
    for l, n in enumerate(per_level):
        w, h = dims[l]
        for sets in (sets_a, sets_b):
            ys = np.sort(rng.integers(0, h, n))
            xs = rng.integers(0, w, n)
            order = np.lexsort((xs, ys))
            arr = np.zeros(n, O.FEATURE_DTYPE)
            arr["x"], arr["y"], arr["lod"] = xs[order], ys[order], l
            src = base[rng.integers(0, 60, n)]
            flips = rng.random((n, 256)) < rng.choice([0.0, 0.02, 0.2], n)[:, None]
            fw = np.packbits(flips.astype(np.uint8), axis=1, bitorder="little").view(np.uint32)
            arr["descriptor"] = src ^ fw
            sets.append(arr)
    return sets_a, sets_b


@pytest.mark.parametrize("seed", range(6))
def test_find_matches_equals_masked_brute_force(seed):
    W0, H0 = 160, 120
    A, B = _small_sets(seed, W0, H0)
    flatA = np.concatenate(A)
    flatB = np.concatenate(B)
    for ratio in (0.85, 0.8, 1.0):
        got, n = O.find_matches(flatA, [len(a) for a in A], flatB, [len(b) for b in B],
                                W0, H0, ratio)
        want = np.concatenate(np_ref.find_matches_bf(A, B, W0, H0, ratio))
        assert np.array_equal(got, want)
        assert n == int((want >= 0).sum())


def test_find_matches_pool_of_one_gives_nothing():
    A, B = _small_sets(1, per_level=(1, 0, 0, 0, 0, 0))
    got, n = O.find_matches(np.concatenate(A), [1, 0, 0, 0, 0, 0], np.concatenate(B),
                            [1, 0, 0, 0, 0, 0], 160, 120)
    assert n == 0 and got[0] == -1


def test_load_features_clamps_count_and_converts():
    cap = 4
    feats = np.zeros(6 * (8 + 48 * cap), np.uint8)
    blk = feats[:8 + 48 * cap]
    blk[:4].view(np.uint32)[0] = 9  # more survivors than slots
    recs = blk[8:].view(O.GPU_FEATURE_DTYPE)
    recs["x"] = [1.0, 2.0, 3.0, 4.0]
    recs["y"] = [5.0, 6.0, 7.0, 8.0]
    recs["strength"] = 13
    out, counts = O.load_features(feats, 6, cap)
    assert counts == [4, 0, 0, 0, 0, 0]
    assert list(out["x"]) == [1, 2, 3, 4] and list(out["lod"]) == [0] * 4
