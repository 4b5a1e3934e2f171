"""ctypes bindings of liborbfe (include/orbfe.h).

The host-side mirror of the reference's feature-detect stage and matcher for
Python callers (tests, the benchmark, the serving loop).  The library must be
built (``python __graft_entry__.py`` or ``make -C uni_slam_amd/csrc``); there
is no fallback: loading fails loudly when it is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liborbfe.so")

OK = 0
EINVAL, EHIP, ENOMEM, ERANGE = -1, -2, -3, -4  # ORBFE_* error codes of orbfe.h
MAX_LEVELS = 8
NOALLOC_SCRATCH_BYTES = 16 << 20  # ORBFE_BF_NOALLOC_SCRATCH_BYTES

GPU_FEATURE_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("strength", "<u4"),
                              ("orientation", "<f4"), ("descriptor", "<u4", (8,))])
FEATURE_DTYPE = np.dtype([("x", "<u4"), ("y", "<u4"), ("lod", "<u4"), ("strength", "<u4"),
                          ("orientation", "<f4"), ("descriptor", "<u4", (8,))])
assert GPU_FEATURE_DTYPE.itemsize == 48 and FEATURE_DTYPE.itemsize == 52


class Config(C.Structure):
    _fields_ = [("frame_width", C.c_uint32), ("frame_height", C.c_uint32),
                ("levels", C.c_uint32), ("scale_factor", C.c_double),
                ("fast_threshold", C.c_float), ("fast_min_run", C.c_uint32),
                ("orientation_radius", C.c_int32), ("features_per_level", C.c_uint32),
                ("border_zero", C.c_uint32), ("lowe_ratio", C.c_float),
                ("detect_tile_rows", C.c_uint32), ("describe_lanes", C.c_uint32),
                ("match_lanes", C.c_uint32)]


class Frame(C.Structure):
    _fields_ = [("image", C.c_void_p), ("pitch", C.c_size_t), ("width", C.c_uint32),
                ("height", C.c_uint32), ("cap", C.c_uint32), ("reserved", C.c_uint32),
                ("pyramid", C.c_void_p), ("features", C.c_void_p)]


class MatchPair(C.Structure):
    _fields_ = [("prev_features", C.c_void_p), ("curr_features", C.c_void_p),
                ("prev_match", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32),
                ("cap", C.c_uint32), ("reserved", C.c_uint32)]


_lib = None

# (name, restype, argtypes) of every entry point declared in include/orbfe.h.
_SIGNATURES = [
    ("orbfe_config_default", C.c_int, [C.POINTER(Config), C.c_uint32, C.c_uint32]),
    ("orbfe_create", C.c_int, [C.POINTER(Config), C.c_int, C.POINTER(C.c_void_p)]),
    ("orbfe_destroy", None, [C.c_void_p]),
    ("orbfe_last_error", C.c_char_p, [C.c_void_p]),
    ("orbfe_level_dims", C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    ("orbfe_feature_array_bytes", C.c_size_t, [C.c_void_p]),
    ("orbfe_pyramid_floats", C.c_size_t, [C.c_void_p]),
    ("orbfe_cap", C.c_uint32, [C.c_void_p]),
    ("orbfe_detect_async", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p,
                                     C.c_void_p, C.c_void_p]),
    ("orbfe_detect", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p]),
    ("orbfe_detect_batch_async", C.c_int, [C.c_void_p, C.POINTER(Frame), C.c_uint32,
                                           C.c_void_p]),
    ("orbfe_batch_create", C.c_int, [C.c_void_p, C.POINTER(Frame), C.c_uint32,
                                     C.POINTER(C.c_void_p)]),
    ("orbfe_batch_detect_async", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("orbfe_batch_destroy", None, [C.c_void_p]),
    ("orbfe_load_features", C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p,
                                      C.POINTER(C.c_uint32)]),
    ("orbfe_export_features_async", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_void_p]),
    ("orbfe_match_pipeline_async", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                             C.c_void_p]),
    ("orbfe_match_pipeline_batch_async", C.c_int, [C.c_void_p, C.POINTER(MatchPair),
                                                   C.c_uint32, C.c_void_p]),
    ("orbfe_match_batch_create", C.c_int, [C.c_void_p, C.POINTER(MatchPair), C.c_uint32,
                                           C.POINTER(C.c_void_p)]),
    ("orbfe_match_batch_async", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("orbfe_match_batch_destroy", None, [C.c_void_p]),
    ("orbfe_match_pipeline", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.POINTER(C.c_uint32)]),
    ("orbfe_match_bf_async", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p,
                                       C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p]),
    ("orbfe_match_bf_noalloc_async", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p,
                                               C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                               C.c_void_p]),
    ("orbfe_ransac_score_async", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p,
                                           C.c_void_p, C.c_uint32, C.c_double, C.c_void_p,
                                           C.c_void_p, C.c_void_p]),
    ("orbfe_timing_enable", C.c_int, [C.c_void_p, C.c_int]),
    ("orbfe_timing_read", C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_uint32),
                                    C.POINTER(C.c_double)]),
    ("orbfe_timing_reset", C.c_int, [C.c_void_p]),
    ("orbfe_probe_valu_rate", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_double)]),
    ("orbfe_probe_mfma_rate", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_double)]),
]

ENTRY_POINTS = [s[0] for s in _SIGNATURES]


def lib():
    """The loaded liborbfe (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"liborbfe is not built ({LIB_PATH} missing); run "
                               "`python __graft_entry__.py` or `make -C uni_slam_amd/csrc`")
        handle = C.CDLL(LIB_PATH)
        for name, res, args in _SIGNATURES:
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


class OrbfeError(RuntimeError):
    pass


def default_config(width: int, height: int, **overrides) -> Config:
    cfg = Config()
    rc = lib().orbfe_config_default(C.byref(cfg), width, height)
    if rc != OK:
        raise OrbfeError(f"orbfe_config_default failed ({rc})")
    for k, v in overrides.items():
        setattr(cfg, k, v)
    return cfg


def _ptr(x) -> Optional[int]:
    """Device/host address of a torch tensor, numpy array, int or None."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    if isinstance(x, np.ndarray):
        return x.ctypes.data
    raise TypeError(f"cannot take the address of {type(x)}")


def _stream(stream) -> Optional[int]:
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


def parse_feature_array(buf, levels: int, cap: int):
    """Split a feature array (bytes-like, gpumodels::FeatureArray layout) into
    per-level (count, features[:min(count, cap)]) pairs."""
    raw = np.frombuffer(memoryview(buf).cast("B"), dtype=np.uint8)
    stride = 8 + 48 * cap
    out = []
    for lod in range(levels):
        block = raw[lod * stride:(lod + 1) * stride]
        count = int(block[:4].view(np.uint32)[0])
        feats = block[8:].view(GPU_FEATURE_DTYPE)[:min(count, cap)]
        out.append((count, feats))
    return out


def feature_array_bytes(levels: int, cap: int) -> int:
    return levels * (8 + 48 * cap)


class Context:
    """One liborbfe context (one frame size, one device, one stream at a time)."""

    def __init__(self, width: int, height: int, device: int = 0, **overrides):
        self.cfg = default_config(width, height, **overrides)
        h = C.c_void_p()
        rc = lib().orbfe_create(C.byref(self.cfg), device, C.byref(h))
        if rc != OK:
            raise OrbfeError(f"orbfe_create({width}x{height}) failed with code {rc}")
        self._h = h
        self.width, self.height = width, height
        self.levels = self.cfg.levels
        self.cap = int(lib().orbfe_cap(h))
        self.feature_bytes = int(lib().orbfe_feature_array_bytes(h))
        self.pyramid_floats = int(lib().orbfe_pyramid_floats(h))
        w = (C.c_uint32 * MAX_LEVELS)()
        hh = (C.c_uint32 * MAX_LEVELS)()
        self._check(lib().orbfe_level_dims(h, w, hh))
        self.dims = [(int(w[i]), int(hh[i])) for i in range(self.levels)]

    def close(self):
        if getattr(self, "_h", None):
            lib().orbfe_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc != OK:
            msg = lib().orbfe_last_error(self._h)
            raise OrbfeError(f"liborbfe error {rc}: {msg.decode() if msg else ''}")

    # -- device-pointer (async) forms ---------------------------------------
    def detect_async(self, image, pitch: int, features, pyramid=None, stream=None):
        self._check(lib().orbfe_detect_async(self._h, _ptr(image), pitch, _ptr(pyramid),
                                             _ptr(features), _stream(stream)))

    def detect_batch_async(self, frames: Sequence[dict], stream=None):
        arr = _frames(frames)
        self._check(lib().orbfe_detect_batch_async(self._h, arr, len(frames), _stream(stream)))

    def batch_create(self, frames: Sequence[dict]) -> "Batch":
        arr = _frames(frames)
        h = C.c_void_p()
        self._check(lib().orbfe_batch_create(self._h, arr, len(frames), C.byref(h)))
        return Batch(self, h)

    def export_features_async(self, features, out, counts, stream=None):
        """FeatureSet of a device feature array into pinned host buffers
        (out: levels*cap FEATURE_DTYPE records, counts: levels u32)."""
        self._check(lib().orbfe_export_features_async(self._h, _ptr(features), _ptr(out),
                                                      _ptr(counts), _stream(stream)))

    def match_pipeline_async(self, prev_features, curr_features, prev_match, stream=None):
        self._check(lib().orbfe_match_pipeline_async(self._h, _ptr(prev_features),
                                                     _ptr(curr_features), _ptr(prev_match),
                                                     _stream(stream)))

    def match_batch_async(self, pairs: Sequence[dict], stream=None):
        arr = _pairs(pairs)
        self._check(lib().orbfe_match_pipeline_batch_async(self._h, arr, len(pairs),
                                                           _stream(stream)))

    def match_batch_create(self, pairs: Sequence[dict]) -> "MatchBatch":
        arr = _pairs(pairs)
        h = C.c_void_p()
        self._check(lib().orbfe_match_batch_create(self._h, arr, len(pairs), C.byref(h)))
        return MatchBatch(self, h)

    def match_bf_async(self, query, nq: int, train, nt: int, idx, d1, d2, stream=None):
        self._check(lib().orbfe_match_bf_async(self._h, _ptr(query), nq, _ptr(train), nt,
                                               _ptr(idx), _ptr(d1), _ptr(d2), _stream(stream)))

    def match_bf_noalloc_async(self, query, nq: int, train, nt: int, idx, d1, d2,
                               stream=None):
        """match_bf_async without any allocation (static device scratch)."""
        self._check(lib().orbfe_match_bf_noalloc_async(self._h, _ptr(query), nq, _ptr(train),
                                                       nt, _ptr(idx), _ptr(d1), _ptr(d2),
                                                       _stream(stream)))

    def ransac_score_async(self, F, k: int, prev, curr, n: int, threshold: float, counts,
                           mask=None, stream=None):
        """Inlier counts (and optionally masks) of k fundamental matrices over
        n correspondences (device tensors: F (k, 9) f64, prev/curr (n, 2) f64)."""
        self._check(lib().orbfe_ransac_score_async(self._h, _ptr(F), k, _ptr(prev), _ptr(curr),
                                                   n, threshold, _ptr(counts), _ptr(mask),
                                                   _stream(stream)))

    # -- host-memory forms ---------------------------------------------------
    def detect(self, image: np.ndarray, with_pyramid: bool = False):
        """Host frame in, (feature array bytes, pyramid floats or None) out."""
        image = np.ascontiguousarray(image, dtype=np.uint8)
        assert image.shape == (self.height, self.width)
        feats = np.empty(self.feature_bytes, dtype=np.uint8)
        pyr = np.empty(self.pyramid_floats, dtype=np.float32) if with_pyramid else None
        self._check(lib().orbfe_detect(self._h, image.ctypes.data, image.strides[0],
                                       None if pyr is None else pyr.ctypes.data,
                                       feats.ctypes.data))
        return feats, pyr

    def match_pipeline(self, prev_features: np.ndarray, curr_features: np.ndarray):
        prev_features = np.ascontiguousarray(prev_features)
        curr_features = np.ascontiguousarray(curr_features)
        out = np.empty(self.levels * self.cap, dtype=np.int32)
        n = C.c_uint32()
        self._check(lib().orbfe_match_pipeline(self._h, prev_features.ctypes.data,
                                               curr_features.ctypes.data, out.ctypes.data,
                                               C.byref(n)))
        return out, int(n.value)

    # -- timing --------------------------------------------------------------
    def timing(self, enable: bool):
        self._check(lib().orbfe_timing_enable(self._h, 1 if enable else 0))

    def timing_reset(self):
        self._check(lib().orbfe_timing_reset(self._h))

    def timing_read(self, name: str):
        n = C.c_uint32()
        ms = C.c_double()
        self._check(lib().orbfe_timing_read(self._h, name.encode(), C.byref(n), C.byref(ms)))
        return int(n.value), float(ms.value)

    def probe_valu_rate(self, stream=None) -> float:
        """Sustained VALU issue rate of the device (wave64 instructions/s)."""
        r = C.c_double()
        self._check(lib().orbfe_probe_valu_rate(self._h, _stream(stream), C.byref(r)))
        return float(r.value)

    def probe_mfma_rate(self, stream=None) -> float:
        """Sustained int8 matrix-core rate of the device (int8 ops/s,
        v_mfma_i32_32x32x32_i8 chains)."""
        r = C.c_double()
        self._check(lib().orbfe_probe_mfma_rate(self._h, _stream(stream), C.byref(r)))
        return float(r.value)


class Batch:
    def __init__(self, ctx: Context, h):
        self.ctx, self._h = ctx, h

    def detect_async(self, stream=None):
        self.ctx._check(lib().orbfe_batch_detect_async(self.ctx._h, self._h, _stream(stream)))

    def close(self):
        if self._h:
            lib().orbfe_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MatchBatch:
    def __init__(self, ctx: Context, h):
        self.ctx, self._h = ctx, h

    def match_async(self, stream=None):
        self.ctx._check(lib().orbfe_match_batch_async(self.ctx._h, self._h, _stream(stream)))

    def close(self):
        if self._h:
            lib().orbfe_match_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _frames(frames):
    """ctypes array of Frame from dicts (image, width, height, pitch, cap,
    pyramid, features); a ctypes Frame array is passed through as is."""
    if isinstance(frames, C.Array):
        return frames
    arr = (Frame * max(len(frames), 1))()
    for i, f in enumerate(frames):
        arr[i].image = _ptr(f["image"])
        arr[i].pitch = f.get("pitch", f["width"])
        arr[i].width = f["width"]
        arr[i].height = f["height"]
        arr[i].cap = f.get("cap", 0)
        arr[i].pyramid = _ptr(f.get("pyramid"))
        arr[i].features = _ptr(f["features"])
    return arr


def _pairs(pairs):
    """ctypes array of MatchPair from dicts (prev, curr, prev_match, width,
    height, cap); a ctypes MatchPair array is passed through as is."""
    if isinstance(pairs, C.Array):
        return pairs
    arr = (MatchPair * max(len(pairs), 1))()
    for i, p in enumerate(pairs):
        arr[i].prev_features = _ptr(p["prev"])
        arr[i].curr_features = _ptr(p["curr"])
        arr[i].prev_match = _ptr(p["prev_match"])
        arr[i].width = p["width"]
        arr[i].height = p["height"]
        arr[i].cap = p.get("cap", 0)
    return arr


def load_features(features: np.ndarray, levels: int, cap: int):
    """FromGpuModel over a host feature array: (FeatureSet records, counts)."""
    features = np.ascontiguousarray(features)
    out = np.empty(levels * cap, dtype=FEATURE_DTYPE)
    counts = (C.c_uint32 * levels)()
    rc = lib().orbfe_load_features(features.ctypes.data, levels, cap, out.ctypes.data, counts)
    if rc != OK:
        raise OrbfeError(f"orbfe_load_features failed ({rc})")
    n = sum(counts)
    return out[:n].copy(), [int(c) for c in counts]


def level_dims(width: int, height: int, levels: int = 6, scale: float = 1.2):
    """Level extents uint32(W0 / pow(scale, l)) (wslam/src/common.cpp:38-45)."""
    out = []
    for lod in range(levels):
        f = 1.0 if lod == 0 else scale ** lod
        out.append((int(width / f), int(height / f)))
    return out


def fast_threshold_u8(value: int) -> float:
    """A FAST threshold given on the 0..255 scale as the f32 fraction the
    shader compares against (value / 255 in f32)."""
    return float(np.float32(value) / np.float32(255.0))
