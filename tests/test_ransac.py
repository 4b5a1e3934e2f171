"""The oracle's RANSAC restatement (ransac_cpu.cpp): Sampson scoring, the
8-point fit and the model search, pinned by closed forms and by two-view
geometry with a known fundamental matrix (tests/ransac_data.py)."""
import numpy as np

from oracle import oracle as O
from tests.ransac_data import normalized, two_view

F_SHIFT_X = np.array([0, 0, 0, 0, 0, -1, 0, 1, 0], np.float64)  # [e_x]_x: y' = y


def test_sampson_closed_form_for_a_horizontal_shift():
    """F = [e_x]_x (epipolar lines horizontal): d = (y - y')^2 / 2 exactly,
    inlier iff strictly below the threshold."""
    rng = np.random.default_rng(0)
    n = 4000
    prev = rng.uniform(0, 1000, (n, 2))
    curr = prev + rng.normal(0, 2.0, (n, 2))
    curr[:8, 1] = prev[:8, 1] + np.array([2.0, -2.0, 1.999, -1.999, 0, 2.001, 3, 0.5])
    for thr in (2.0, 3.84, 0.5):
        counts, mask = O.sampson_counts(F_SHIFT_X, prev, curr, thr, with_mask=True)
        d = (prev[:, 1] - curr[:, 1]) ** 2 / 2
        want = d < thr
        assert np.array_equal(mask[0].astype(bool), want) and counts[0] == want.sum()
    # boundary: d == 2 exactly is not an inlier at threshold 2
    _, mask = O.sampson_counts(F_SHIFT_X, prev[:8], curr[:8], 2.0, with_mask=True)
    assert list(mask[0]) == [0, 0, 1, 1, 1, 0, 0, 1]


def test_sampson_scale_invariance_and_degenerate_model():
    p, c, _, F = two_view(3, 1000, 0.3, 0.3)
    F = normalized(F)
    base = O.sampson_counts(F, p, c, 3.84, with_mask=True)
    for s in (2.0 ** 10, 2.0 ** -6, -1.0):  # exact scalings leave every distance as is
        got = O.sampson_counts(F * s, p, c, 3.84, with_mask=True)
        assert np.array_equal(got[0], base[0]) and np.array_equal(got[1], base[1])
    # the zero matrix: denominator 0 < 1e-12 -> infinite distance, no inlier
    assert O.sampson_counts(np.zeros(9), p, c, 1e9)[0] == 0


def test_fit_recovers_the_true_fundamental_matrix():
    p, c, inl, F = two_view(4, 400, 0.0)
    for sl in (slice(0, 8), slice(100, 140), slice(None)):
        got = O.fit_fundamental(p[sl], c[sl])
        assert got is not None
        assert np.abs(normalized(got) - normalized(F)).max() < 1e-9, sl
        assert abs(np.linalg.det(got)) < 1e-9 * np.linalg.norm(got) ** 3  # rank 2
    # fewer than 8 correspondences: no fit
    assert O.fit_fundamental(p[:7], c[:7]) is None


def test_ransac_finds_the_model_among_outliers():
    sets = [two_view(5, 300, 0.3), two_view(6, 600, 0.5), two_view(7, 5, 0.0)]
    out = O.ransac([(p, c) for p, c, _, _ in sets])
    for (p, c, inl, F), (st, Fr, mask) in zip(sets[:2], out[:2]):
        assert st["found"] and st["total"] == len(p) and st["inliers"] == mask.sum()
        assert 0 < st["iterations"] <= 500
        m = mask.astype(bool)
        assert m[inl].all()  # every true correspondence kept
        # what else was kept lies near the true model's epipolar lines (the
        # final model is refit on the kept set, so not exactly within 3.84)
        extra = m & ~inl
        d_true = O.sampson_counts(F, p[extra], c[extra], 4 * 3.84, with_mask=True)[1][0]
        assert d_true.all()
        # the refit model explains the true correspondences (refit on the kept
        # set, which may hold a few outliers near the epipolar lines)
        assert O.sampson_counts(Fr, p[inl], c[inl], 1.0)[0] >= 0.9 * inl.sum()
    assert out[0][0]["iterations"] < 500  # 30% outliers: the adaptive bound stops early
    st, _, mask = out[2]
    assert st == {"total": 5, "iterations": 0, "inliers": 0, "found": False} and not mask.any()
