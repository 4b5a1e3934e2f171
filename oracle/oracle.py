"""ctypes bindings of the oracle (liborbfe_oracle.so) -- TEST INFRASTRUCTURE.

The oracle is the CPU restatement of the reference path used as the checker of
the HIP kernels and as the CPU baseline of the benchmark.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liborbfe_oracle.so")

GPU_FEATURE_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("strength", "<u4"),
                              ("orientation", "<f4"), ("descriptor", "<u4", (8,))])
FEATURE_DTYPE = np.dtype([("x", "<u4"), ("y", "<u4"), ("lod", "<u4"), ("strength", "<u4"),
                          ("orientation", "<f4"), ("descriptor", "<u4", (8,))])


class Params(C.Structure):
    _fields_ = [("frame_width", C.c_uint32), ("frame_height", C.c_uint32),
                ("levels", C.c_uint32), ("scale_factor", C.c_double),
                ("fast_threshold", C.c_float), ("fast_min_run", C.c_uint32),
                ("orientation_radius", C.c_int32), ("cap", C.c_uint32),
                ("border_zero", C.c_uint32)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, u32, f32 = C.c_void_p, C.c_uint32, C.c_float
        L.oracle_level_dims.argtypes = [u32, u32, u32, C.c_double, vp, vp]
        L.oracle_lod0.argtypes = [vp, C.c_size_t, u32, u32, vp]
        L.oracle_pyramid_level.argtypes = [vp, u32, u32, u32, u32, vp]
        L.oracle_detect.argtypes = [vp, u32, u32, f32, u32, u32, vp]
        L.oracle_cull.argtypes = [vp, u32, u32, u32, vp, vp, vp]
        L.oracle_generate.argtypes = [vp, u32, u32, vp, u32, C.c_int32, u32, vp]
        L.oracle_generate.restype = u32
        L.oracle_detect_frame.argtypes = [vp, C.c_size_t, C.POINTER(Params), vp, vp]
        L.oracle_load_features.argtypes = [vp, u32, u32, vp, vp]
        L.oracle_atan2f.argtypes = [f32, f32]
        L.oracle_atan2f.restype = f32
        L.oracle_sincosf.argtypes = [f32, C.POINTER(f32), C.POINTER(f32)]
        L.oracle_hamming_bounded.argtypes = [vp, vp, u32]
        L.oracle_hamming_bounded.restype = u32
        L.oracle_find_matches.argtypes = [vp, vp, vp, vp, u32, u32, u32, f32, vp, vp]
        L.oracle_match_bf.argtypes = [vp, u32, vp, u32, vp, vp, vp]
        L.oracle_sampson_counts.argtypes = [vp, u32, vp, vp, u32, C.c_double, vp, vp]
        L.oracle_fit_fundamental.argtypes = [vp, vp, u32, vp]
        L.oracle_fit_fundamental.restype = C.c_int
        L.oracle_ransac.argtypes = [vp, u32, vp, vp, u32, vp, vp, vp]
        _lib = L
    return _lib


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def params(width, height, levels=6, scale=1.2, threshold=0.15, min_run=9, radius=3,
           cap=0, border_zero=0) -> Params:
    p = Params()
    p.frame_width, p.frame_height, p.levels = width, height, levels
    p.scale_factor, p.fast_threshold, p.fast_min_run = scale, threshold, min_run
    p.orientation_radius = radius
    p.cap = cap if cap else width * height // 250
    p.border_zero = border_zero
    return p


def level_dims(width, height, levels=6, scale=1.2):
    w = np.zeros(levels, np.uint32)
    h = np.zeros(levels, np.uint32)
    lib().oracle_level_dims(width, height, levels, scale, _p(w), _p(h))
    return [(int(a), int(b)) for a, b in zip(w, h)]


def lod0(img: np.ndarray) -> np.ndarray:
    img = np.ascontiguousarray(img, np.uint8)
    out = np.empty(img.shape, np.float32)
    lib().oracle_lod0(_p(img), img.strides[0], img.shape[1], img.shape[0], _p(out))
    return out


def pyramid_level(level0: np.ndarray, w: int, h: int) -> np.ndarray:
    level0 = np.ascontiguousarray(level0, np.float32)
    out = np.empty((h, w), np.float32)
    lib().oracle_pyramid_level(_p(level0), level0.shape[1], level0.shape[0], w, h, _p(out))
    return out


def detect(level: np.ndarray, threshold: float, min_run: int = 9, border_zero: int = 0):
    level = np.ascontiguousarray(level, np.float32)
    out = np.empty(level.shape, np.uint32)
    lib().oracle_detect(_p(level), level.shape[1], level.shape[0], threshold, min_run,
                        border_zero, _p(out))
    return out


def cull(S_all: np.ndarray, W0: int, H0: int, dims) -> np.ndarray:
    """S_all: (levels, W0*H0) u32 blocks; returns O_all of the same shape."""
    S_all = np.ascontiguousarray(S_all, np.uint32)
    L = S_all.shape[0]
    dw = np.array([d[0] for d in dims], np.uint32)
    dh = np.array([d[1] for d in dims], np.uint32)
    out = np.empty_like(S_all)
    lib().oracle_cull(_p(S_all), W0, H0, L, _p(dw), _p(dh), _p(out))
    return out


def generate(level: np.ndarray, O: np.ndarray, cap: int, radius: int = 3, border_zero: int = 0):
    level = np.ascontiguousarray(level, np.float32)
    O = np.ascontiguousarray(O, np.uint32)
    out = np.zeros(cap, GPU_FEATURE_DTYPE)
    n = lib().oracle_generate(_p(level), level.shape[1], level.shape[0], _p(O), cap, radius,
                              border_zero, _p(out))
    return int(n), out[:min(int(n), cap)]


def detect_frame(img: np.ndarray, p: Params, with_pyramid: bool = False):
    """Returns (feature array bytes as u8 array, pyramid floats or None)."""
    img = np.ascontiguousarray(img, np.uint8)
    dims = level_dims(p.frame_width, p.frame_height, p.levels, p.scale_factor)
    feats = np.zeros(p.levels * (8 + 48 * p.cap), np.uint8)
    pyr = None
    if with_pyramid:
        pyr = np.zeros(sum(w * h for w, h in dims[1:]), np.float32)
    rc = lib().oracle_detect_frame(_p(img), img.strides[0], C.byref(p),
                                   None if pyr is None else _p(pyr), _p(feats))
    assert rc == 0
    return feats, pyr


def load_features(feats: np.ndarray, levels: int, cap: int):
    out = np.empty(levels * cap, FEATURE_DTYPE)
    counts = np.zeros(levels, np.uint32)
    lib().oracle_load_features(_p(feats), levels, cap, _p(out), _p(counts))
    return out[:int(counts.sum())].copy(), [int(c) for c in counts]


def find_matches(prev, prev_counts, curr, curr_counts, W0, H0, ratio=0.85, levels=6):
    prev = np.ascontiguousarray(prev, FEATURE_DTYPE)
    curr = np.ascontiguousarray(curr, FEATURE_DTYPE)
    pc = np.array(prev_counts, np.uint32)
    cc = np.array(curr_counts, np.uint32)
    out = np.full(max(len(prev), 1), -1, np.int32)
    n = np.zeros(1, np.uint32)
    rc = lib().oracle_find_matches(_p(prev), _p(pc), _p(curr), _p(cc), levels, W0, H0, ratio,
                                   _p(out), _p(n))
    assert rc == 0
    return out[:len(prev)], int(n[0])


def match_bf(q: np.ndarray, t: np.ndarray):
    q = np.ascontiguousarray(q, np.uint32)
    t = np.ascontiguousarray(t, np.uint32)
    nq, nt = q.shape[0], t.shape[0]
    idx = np.empty(nq, np.uint32)
    d1 = np.empty(nq, np.uint32)
    d2 = np.empty(nq, np.uint32)
    lib().oracle_match_bf(_p(q), nq, _p(t) if nt else None, nt, _p(idx), _p(d1), _p(d2))
    return idx, d1, d2


def atan2f(y: float, x: float) -> float:
    return float(lib().oracle_atan2f(y, x))


def sincosf(x: float):
    s, c = C.c_float(), C.c_float()
    lib().oracle_sincosf(x, C.byref(s), C.byref(c))
    return float(s.value), float(c.value)


def hamming_bounded(a, b, cap: int) -> int:
    a = np.ascontiguousarray(a, np.uint32)
    b = np.ascontiguousarray(b, np.uint32)
    return int(lib().oracle_hamming_bounded(_p(a), _p(b), cap))


def prev_match_to_padded(prev_match_flat: np.ndarray, counts, cap: int, levels: int = 6):
    """Oracle prev_match (flattened per FeatureSet) -> the library's
    levels*cap layout with -1 padding."""
    out = np.full(levels * cap, -1, np.int32)
    off = 0
    for l in range(levels):
        n = counts[l]
        out[l * cap:l * cap + n] = prev_match_flat[off:off + n]
        off += n
    return out


def sampson_counts(F: np.ndarray, prev: np.ndarray, curr: np.ndarray, threshold: float,
                   with_mask: bool = False):
    """Inlier counts (and masks) of k fundamental matrices F (k, 9) over
    correspondences prev, curr (n, 2), float64."""
    F = np.ascontiguousarray(F, np.float64).reshape(-1, 9)
    prev = np.ascontiguousarray(prev, np.float64)
    curr = np.ascontiguousarray(curr, np.float64)
    k, n = F.shape[0], prev.shape[0]
    counts = np.zeros(k, np.uint32)
    mask = np.zeros((k, n), np.uint8) if with_mask else None
    lib().oracle_sampson_counts(_p(F), k, _p(prev), _p(curr), n, threshold, _p(counts),
                                None if mask is None else _p(mask))
    return (counts, mask) if with_mask else counts


def fit_fundamental(prev: np.ndarray, curr: np.ndarray):
    prev = np.ascontiguousarray(prev, np.float64)
    curr = np.ascontiguousarray(curr, np.float64)
    F = np.zeros(9, np.float64)
    ok = lib().oracle_fit_fundamental(_p(prev), _p(curr), prev.shape[0], _p(F))
    return F.reshape(3, 3) if ok else None


def ransac(sets, max_iterations: int = 500):
    """RansacCPU over consecutive sets [(prev (n, 2), curr (n, 2)), ...] with
    one generator: per set (stats dict, F (3, 3), inlier mask)."""
    sizes = np.array([len(p) for p, _ in sets], np.uint32)
    prev = np.ascontiguousarray(np.concatenate([p for p, _ in sets]), np.float64)
    curr = np.ascontiguousarray(np.concatenate([c for _, c in sets]), np.float64)
    stats = np.zeros((len(sets), 4), np.uint64)
    F = np.zeros((len(sets), 9), np.float64)
    mask = np.zeros(len(prev), np.uint8)
    lib().oracle_ransac(_p(sizes), len(sets), _p(prev), _p(curr), max_iterations, _p(stats),
                        _p(F), _p(mask))
    out, off = [], 0
    for i, n in enumerate(sizes):
        st = {"total": int(stats[i, 0]), "iterations": int(stats[i, 1]),
              "inliers": int(stats[i, 2]), "found": bool(stats[i, 3])}
        out.append((st, F[i].reshape(3, 3), mask[off:off + n].copy()))
        off += n
    return out
