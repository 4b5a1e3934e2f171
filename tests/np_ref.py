"""Independent numpy restatements used to pin the oracle.

These are written from the WGSL / C++ of the reference without sharing code
with the oracle: float64 bilinear sampling and BRIEF, a closed-form corner
cull mask (the oracle simulates the two dispatch grids instead), and a masked
brute-force FindMatches.  They check structure (indices, offsets, signs,
operand order, bit order); bit-exactness of the f32 arithmetic is the
oracle's job and is checked against the GPU.
"""
from __future__ import annotations

import math

import numpy as np

CIRCLE_X = np.array([0, 1, 2, 3, 3, 3, 2, 1, 0, -1, -2, -3, -3, -3, -2, -1])
CIRCLE_Y = np.array([-3, -3, -2, -1, 0, 1, 2, 3, 3, 3, 2, 1, 0, -1, -2, -3])


def tint_clamp(c: np.ndarray, n: int) -> np.ndarray:
    """min(u32(c), n-1): negative coordinates wrap to the far edge."""
    u = np.asarray(c).astype(np.int64) & 0xFFFFFFFF
    return np.minimum(u, n - 1)


def bilinear64(tex: np.ndarray, u: np.ndarray, v: np.ndarray) -> np.ndarray:
    """Vulkan linear, clamp-to-edge sampling in float64 at normalised (u, v)."""
    h, w = tex.shape
    uu = np.asarray(u, np.float64) * w - 0.5
    vv = np.asarray(v, np.float64) * h - 0.5
    i0 = np.floor(uu)
    j0 = np.floor(vv)
    a = uu - i0
    b = vv - j0
    x0 = np.clip(i0, 0, w - 1).astype(np.int64)
    x1 = np.clip(i0 + 1, 0, w - 1).astype(np.int64)
    y0 = np.clip(j0, 0, h - 1).astype(np.int64)
    y1 = np.clip(j0 + 1, 0, h - 1).astype(np.int64)
    t = tex.astype(np.float64)
    return ((1 - a) * (1 - b) * t[y0, x0] + a * (1 - b) * t[y0, x1] +
            (1 - a) * b * t[y1, x0] + a * b * t[y1, x1])


def pyramid_level64(lod0: np.ndarray, w: int, h: int) -> np.ndarray:
    xs = (np.arange(w) + 0.5) / w
    ys = (np.arange(h) + 0.5) / h
    U, V = np.meshgrid(xs, ys)
    return bilinear64(lod0, U, V)


def fast_scores(level: np.ndarray, thr: float, min_run: int = 9, border_zero: bool = False):
    """detect_corners.wgsl in float32 compares (exact), vectorised."""
    h, w = level.shape
    lv = level.astype(np.float32)
    ys, xs = np.mgrid[0:h, 0:w]
    center = lv
    curr = np.ones((h, w), np.int64)
    streak = np.ones((h, w), np.int64)
    for i in range(16):
        cx = xs + CIRCLE_X[i]
        cy = ys + CIRCLE_Y[i]
        if border_zero:
            inside = (cx >= 0) & (cx < w) & (cy >= 0) & (cy < h)
            val = np.where(inside, lv[np.clip(cy, 0, h - 1), np.clip(cx, 0, w - 1)],
                           np.float32(0))
        else:
            val = lv[tint_clamp(cy, h), tint_clamp(cx, w)]
        diff = np.abs(center - val) > np.float32(thr)
        curr = np.where(diff, curr + 1, 1)
        streak = np.maximum(streak, curr)
    return np.where(streak > min_run, streak, 0).astype(np.uint32)


def cull_mask(W0: int, H0: int, w: int, h: int):
    """Closed form of the static part of cull_corners.wgsl for one level:
    (hor, ver) boolean arrays over p < w*h."""
    n = w * h
    N0 = W0 * H0
    gxh = 64 * (((W0 + 2) // 3 + 63) // 64)
    gx = np.arange(gxh)[None, :]
    gy = np.arange(H0)[:, None]
    s = (9 * gx + gy * w).ravel()
    hor = np.zeros(n, bool)
    hor[s[s < n]] = True
    gxv = 64 * ((W0 + 63) // 64)
    gyv = (H0 + 2) // 3
    gx = np.arange(gxv)[None, :]
    gy = np.arange(gyv)[:, None]
    s = (3 * gx + 3 * gy * w).ravel()
    s = s[s < n]
    ver = np.zeros(n, bool)
    ver[s] = True
    steps = np.minimum((N0 - s) // w, 3)
    for k in range(3):
        p = s[steps > k] + k * w
        ver[p[p < n]] = True
    return hor, ver


def cull_closed_form(S: np.ndarray, W0: int, H0: int, w: int, h: int) -> np.ndarray:
    """Survivor strengths of one level from its scores S (h, w)."""
    hor, ver = cull_mask(W0, H0, w, h)
    flat = S.ravel().astype(np.int64)
    n = w * h
    out = np.zeros(n, np.int64)
    rest = np.arange(w, n)
    keep = hor[rest] & ver[rest] & (flat[rest] > 0)
    out[rest[keep]] = flat[rest[keep]]
    for s in range(0, w, 9):
        e = min(s + 3, w)
        win = flat[s:e]
        m = s + int(np.argmax(win))  # first maximum
        if flat[m] > 0 and ver[m]:
            out[m] = flat[m]
    return out.reshape(h, w)


# --- BRIEF / orientation in float64 ------------------------------------------
OFFSETS = (-1.2, 0.0, 1.2)
WEIGHTS = (0.3125, 0.375, 0.3125)


def gaussian_tap9_64(level: np.ndarray, px: int, py: int) -> float:
    h, w = level.shape
    uvx = (px + 0.5) / w
    uvy = (py + 0.5) / h
    total = 0.0
    for j in range(3):
        for i in range(3):
            total += float(bilinear64(level, uvx + OFFSETS[i] / w, uvy + OFFSETS[j] / h)) \
                * WEIGHTS[i] * WEIGHTS[j]
    return total


def orientation64(level: np.ndarray, x: int, y: int, r: int = 3):
    """(angle, bound): the intensity-centroid angle in float64 and a bound on
    how far an f32 accumulation of the same moments can move it."""
    h, w = level.shape
    m10 = m01 = 0.0
    mag = 0.0
    for dy in range(-r, r + 1):
        for dx in range(-r, r + 1):
            if dx * dx + dy * dy > r * r:
                continue
            v = float(level[int(tint_clamp(y + dy, h)), int(tint_clamp(x + dx, w))])
            m10 += dx * v
            m01 += dy * v
            mag += (abs(dx) + abs(dy)) * abs(v)
    norm = math.hypot(m10, m01)
    bound = 1e-5 + (32 * mag * 2.0 ** -24 / norm if norm > 0 else math.inf)
    return math.atan2(m01, m10), bound


def u32(v: int) -> int:
    return v & 0xFFFFFFFF


# --- matcher -------------------------------------------------------------------
def popcount_u32(a: np.ndarray) -> np.ndarray:
    """Popcount summed over the last axis of a u32 array."""
    return np.bitwise_count(a.astype(np.uint32)).sum(-1, dtype=np.int64)


def hamming_matrix(q: np.ndarray, t: np.ndarray, chunk: int = 16) -> np.ndarray:
    """(len(q), len(t)) Hamming distances of 256-bit descriptors."""
    out = np.empty((len(q), len(t)), np.int64)
    for i in range(0, len(q), chunk):
        x = q[i:i + chunk, None, :] ^ t[None, :, :]
        out[i:i + chunk] = np.bitwise_count(x).sum(-1, dtype=np.int64)
    return out


def to_lod0(x, y, lod):
    s = np.float32(1.0)
    for _ in range(int(lod)):
        s = np.float32(s * np.float32(1.2))
    return np.float32(np.float32(x) * s), np.float32(np.float32(y) * s)


def find_matches_in_lod_bf(qs, pool, W0, H0, ratio):
    """Masked brute force of FindMatchesInLod: for each query the candidate
    with the unique minimum distance inside the window, if it passes the
    ratio test.  Returns {query index: pool index}."""
    dx = np.float32(np.float32(0.1) * np.float32(W0))
    dy = np.float32(np.float32(0.1) * np.float32(H0))
    res = {}
    if len(pool) < 2:
        return res
    pxy = [to_lod0(f["x"], f["y"], f["lod"]) for f in pool]
    for qi, q in enumerate(qs):
        qx, qy = to_lod0(q["x"], q["y"], q["lod"])
        lo = np.float32(qx - dx)
        hi = np.float32(qx + dx)
        ds = []
        for ci, c in enumerate(pool):
            cx, cy = pxy[ci]
            if cx < lo or cx > hi or abs(np.float32(qy - cy)) > dy:
                continue
            d = int(sum(bin(int(a) ^ int(b)).count("1")
                        for a, b in zip(q["descriptor"], c["descriptor"])))
            ds.append((d, ci))
        if not ds:
            continue
        ds.sort()
        best, bi = ds[0]
        second = ds[1][0] if len(ds) > 1 else 0xFFFFFFFF
        if np.float32(best) < np.float32(ratio) * np.float32(second):
            res[qi] = bi
    return res


def find_matches_bf(prev_sets, curr_sets, W0, H0, ratio):
    """FindMatches on lists of per-level structured arrays; returns the
    prev_match list per level ((lod << 24) | index or -1)."""
    L = len(prev_sets)

    def pool(sets, l):
        items = [(l, i) for i in range(len(sets[l]))]
        if l > 0:
            items += [(l - 1, i) for i in range(len(sets[l - 1]))]
        if l + 1 < L:
            items += [(l + 1, i) for i in range(len(sets[l + 1]))]
        return items

    def rec(sets, l, i):
        f = sets[l][i]
        return {"x": int(f["x"]), "y": int(f["y"]), "lod": l, "descriptor": f["descriptor"]}

    b_to_a = {}
    for l in range(L):
        items = pool(prev_sets, l)
        res = find_matches_in_lod_bf([rec(curr_sets, l, i) for i in range(len(curr_sets[l]))],
                                     [rec(prev_sets, *it) for it in items], W0, H0, ratio)
        for qi, pi in res.items():
            b_to_a[(l, qi)] = items[pi]
    out = []
    for l in range(L):
        items = pool(curr_sets, l)
        res = find_matches_in_lod_bf([rec(prev_sets, l, i) for i in range(len(prev_sets[l]))],
                                     [rec(curr_sets, *it) for it in items], W0, H0, ratio)
        pm = np.full(len(prev_sets[l]), -1, np.int64)
        for qi, ci in res.items():
            b = items[ci]
            if b_to_a.get(b) == (l, qi):
                pm[qi] = (b[0] << 24) | b[1]
        out.append(pm)
    return out
