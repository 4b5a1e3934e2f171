"""Two-view correspondences with a known fundamental matrix, for the RANSAC
tests: random 3-D points seen by two pinhole cameras (the second rotated and
translated), LOD-0 pixel coordinates, a share replaced by outliers."""
import numpy as np


def _rot(rng, max_angle):
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    a = rng.uniform(-max_angle, max_angle)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(a) * K + (1 - np.cos(a)) * K @ K


def two_view(seed, n, outlier_frac=0.3, noise_px=0.0, width=1280, height=720):
    """Returns (prev (n, 2), curr (n, 2), inlier mask, F_true (3, 3)) with
    x_curr^T F_true x_prev = 0 for the inliers."""
    rng = np.random.default_rng(seed)
    f = 0.9 * width
    Kc = np.array([[f, 0, width / 2], [0, f, height / 2], [0, 0, 1.0]])
    R = _rot(rng, 0.15)
    t = np.array([1.0, rng.uniform(-0.3, 0.3), rng.uniform(-0.2, 0.2)])
    X = np.stack([rng.uniform(-3, 3, n), rng.uniform(-2, 2, n), rng.uniform(4, 12, n)], 1)
    x1 = (Kc @ X.T).T
    x1 = x1[:, :2] / x1[:, 2:]
    X2 = (R @ X.T).T + t
    x2 = (Kc @ X2.T).T
    x2 = x2[:, :2] / x2[:, 2:]
    if noise_px:
        x1 += rng.normal(0, noise_px, x1.shape)
        x2 += rng.normal(0, noise_px, x2.shape)
    inl = rng.random(n) >= outlier_frac
    out = ~inl
    x2[out] = np.stack([rng.uniform(0, width, out.sum()), rng.uniform(0, height, out.sum())], 1)
    tx = np.array([[0, -t[2], t[1]], [t[2], 0, -t[0]], [-t[1], t[0], 0]])
    Ki = np.linalg.inv(Kc)
    F = Ki.T @ tx @ R @ Ki
    return x1, x2, inl, F


def normalized(F):
    """F scaled to unit Frobenius norm with its largest-magnitude entry positive."""
    F = np.asarray(F, np.float64).reshape(3, 3)
    F = F / np.linalg.norm(F)
    return F * np.sign(F.flat[np.argmax(np.abs(F))])
