"""Served form of the front-end: Python binding of liborbfe's continuous-
batching server (orbfe_server_* of include/orbfe.h).

The server itself is native: per-stream slot rings resident in HBM, FIFO
admission over all streams within an HBM budget, batched detect + match per
in-flight slot, device-side waits between a stream's consecutive frames, and
the real-time schedule replay loop (orbfe_server_run) all live in
uni_slam_amd/csrc/orbfe_server.cpp.  This module turns Python schedules and
torch pinned frames into its C structures, and computes latency percentiles.

Results are bit-identical to the plain per-frame path: the batched kernels
compute each frame exactly as the per-frame launch does.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import orbfe

LEVELS = 6


class StreamSpec(C.Structure):
    """orbfe_stream_spec: width, height, cap (0 = width*height/250)."""
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("cap", C.c_uint32),
                ("reserved", C.c_uint32)]

    def __init__(self, width: int, height: int, cap: int = 0):
        super().__init__(width, height, cap, 0)


class ServerConfig(C.Structure):
    _fields_ = [("detect", orbfe.Config), ("hbm_budget", C.c_size_t),
                ("max_batch", C.c_uint32), ("inflight", C.c_uint32),
                ("frames_per_stream", C.c_uint32), ("log_batches", C.c_uint32)]


class Arrival(C.Structure):
    _fields_ = [("t", C.c_double), ("stream", C.c_uint32), ("seq", C.c_uint32)]


class Completion(C.Structure):
    _fields_ = [("stream", C.c_uint32), ("seq", C.c_uint32), ("t_arrival", C.c_double),
                ("t_done", C.c_double)]


class BatchEntry(C.Structure):
    _fields_ = [("batch", C.c_uint32), ("stream", C.c_uint32), ("seq", C.c_uint32),
                ("order", C.c_uint32)]


_SIGS = [
    ("orbfe_frame_scratch_bytes", C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                            C.POINTER(C.c_size_t)]),
    ("orbfe_server_config_default", C.c_int, [C.POINTER(ServerConfig)]),
    ("orbfe_server_create", C.c_int, [C.POINTER(ServerConfig), C.POINTER(StreamSpec), C.c_uint32,
                                      C.c_int, C.POINTER(C.c_void_p)]),
    ("orbfe_server_destroy", None, [C.c_void_p]),
    ("orbfe_server_last_error", C.c_char_p, [C.c_void_p]),
    ("orbfe_server_clock", C.c_double, [C.c_void_p]),
    ("orbfe_server_submit", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_double, C.c_void_p,
                                      C.c_void_p, C.POINTER(C.c_uint32)]),
    ("orbfe_server_pump", C.c_int, [C.c_void_p, C.POINTER(Completion), C.c_uint32,
                                    C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    ("orbfe_server_idle", C.c_int, [C.c_void_p]),
    ("orbfe_server_stats", C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]),
    ("orbfe_server_batch_log", C.c_int, [C.c_void_p, C.POINTER(BatchEntry), C.c_size_t,
                                         C.POINTER(C.c_size_t)]),
    ("orbfe_server_run", C.c_int, [C.c_void_p, C.POINTER(Arrival), C.c_uint32,
                                   C.POINTER(C.c_void_p), C.c_uint32, C.c_double,
                                   C.POINTER(Completion), C.POINTER(C.c_double)]),
]
ENTRY_POINTS = [s[0] for s in _SIGS]
_bound = False


def _lib():
    global _bound
    L = orbfe.lib()
    if not _bound:
        for name, res, args in _SIGS:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _bound = True
    return L


def frame_scratch_bytes(width: int, height: int, cap: int, ctx: Optional[orbfe.Context] = None
                        ) -> int:
    """Device scratch one in-flight frame takes in a batch (detect and match
    stages), as liborbfe lays it out (orbfe_frame_scratch_bytes)."""
    own = ctx is None
    if own:
        ctx = orbfe.Context(width, height, features_per_level=cap)
    try:
        n = C.c_size_t()
        ctx._check(_lib().orbfe_frame_scratch_bytes(ctx._h, width, height, cap, C.byref(n)))
        return int(n.value)
    finally:
        if own:
            ctx.close()


@dataclass
class Completed:
    stream: int
    seq: int
    t_arrival: float
    t_done: float


class Server:
    """A liborbfe server for the given streams.  frames_per_stream: how many
    consecutive frames of one stream a batch may hold (each stream keeps
    2 * frames_per_stream + 2 slots resident).  With keep_outputs, every
    frame's feature array and match table are copied to pinned host memory
    by the time it completes (outputs[(stream, seq)])."""

    def __init__(self, specs: Sequence[StreamSpec], device: int = 0,
                 hbm_budget: int = 1 << 30, max_batch: int = 64, inflight: int = 2,
                 fast_threshold: float = 0.15, lowe_ratio: float = 0.85,
                 keep_outputs: bool = False, log_batches: bool = False,
                 frames_per_stream: int = 3, **launch_shape):
        L = _lib()
        self.specs = [StreamSpec(s.width, s.height, s.cap) for s in specs]
        cfg = ServerConfig()
        L.orbfe_server_config_default(C.byref(cfg))
        cfg.detect.fast_threshold = fast_threshold
        cfg.detect.lowe_ratio = lowe_ratio
        for k, v in launch_shape.items():
            setattr(cfg.detect, k, v)
        cfg.hbm_budget = hbm_budget
        cfg.max_batch = max_batch
        cfg.inflight = inflight
        cfg.frames_per_stream = frames_per_stream
        cfg.log_batches = 1 if log_batches else 0
        arr = (StreamSpec * len(self.specs))(*self.specs)
        h = C.c_void_p()
        rc = L.orbfe_server_create(C.byref(cfg), arr, len(self.specs), device, C.byref(h))
        if rc != orbfe.OK:
            raise orbfe.OrbfeError(f"orbfe_server_create failed ({rc})")
        self._h = h
        self.keep_outputs = keep_outputs
        self.outputs: Dict[Tuple[int, int], Tuple[np.ndarray, Optional[np.ndarray]]] = {}
        self._pending_out: Dict[Tuple[int, int], Tuple[torch.Tensor, torch.Tensor]] = {}
        self._hold: Dict[Tuple[int, int], torch.Tensor] = {}  # frames until they complete
        self.completed: List[Completed] = []
        self._done_buf = (Completion * 4096)()

    def close(self):
        if getattr(self, "_h", None):
            _lib().orbfe_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != orbfe.OK:
            msg = _lib().orbfe_server_last_error(self._h)
            raise orbfe.OrbfeError(f"liborbfe server error {rc}: {msg.decode() if msg else ''}")

    def clock(self) -> float:
        return float(_lib().orbfe_server_clock(self._h))

    @property
    def batches(self) -> int:
        b, m = C.c_uint64(), C.c_uint32()
        self._check(_lib().orbfe_server_stats(self._h, C.byref(b), C.byref(m)))
        return int(b.value)

    @property
    def max_batch_seen(self) -> int:
        b, m = C.c_uint64(), C.c_uint32()
        self._check(_lib().orbfe_server_stats(self._h, C.byref(b), C.byref(m)))
        return int(m.value)

    @property
    def batch_log(self) -> List[List[Tuple[int, int, int]]]:
        """Per launched batch: its frames as (stream, seq, submission number)."""
        n = C.c_size_t()
        self._check(_lib().orbfe_server_batch_log(self._h, None, 0, C.byref(n)))
        buf = (BatchEntry * max(n.value, 1))()
        self._check(_lib().orbfe_server_batch_log(self._h, buf, n.value, C.byref(n)))
        out: List[List[Tuple[int, int, int]]] = []
        for e in buf[:n.value]:
            while len(out) <= e.batch:
                out.append([])
            out[e.batch].append((e.stream, e.seq, e.order))
        return out

    def submit(self, stream: int, host_frame: torch.Tensor, t_arrival: Optional[float] = None):
        """Queues a pinned u8 frame of `stream`; returns its seq."""
        sp = self.specs[stream]
        assert host_frame.is_pinned() and host_frame.numel() == sp.width * sp.height
        fo = mo = None
        if self.keep_outputs:
            cap = sp.cap or sp.width * sp.height // 250
            fo = torch.empty(orbfe.feature_array_bytes(LEVELS, cap), dtype=torch.uint8,
                             pin_memory=True)
            mo = torch.empty(LEVELS * cap, dtype=torch.int32, pin_memory=True)
        seq = C.c_uint32()
        t = self.clock() if t_arrival is None else t_arrival
        self._check(_lib().orbfe_server_submit(self._h, stream, host_frame.data_ptr(), t,
                                               None if fo is None else fo.data_ptr(),
                                               None if mo is None else mo.data_ptr(),
                                               C.byref(seq)))
        key = (stream, int(seq.value))
        self._hold[key] = host_frame
        if fo is not None:
            self._pending_out[key] = (fo, mo)
        return key[1]

    def pump(self) -> int:
        """Retire finished batches, launch new ones; returns batches launched."""
        n, launched = C.c_uint32(), C.c_uint32()
        self._check(_lib().orbfe_server_pump(self._h, self._done_buf, len(self._done_buf),
                                             C.byref(n), C.byref(launched)))
        for c in self._done_buf[:n.value]:
            key = (c.stream, c.seq)
            self.completed.append(Completed(c.stream, c.seq, c.t_arrival, c.t_done))
            self._hold.pop(key, None)
            out = self._pending_out.pop(key, None)
            if out is not None:
                self.outputs[key] = (out[0].numpy().copy(),
                                     out[1].numpy().copy() if c.seq > 0 else None)
        return int(launched.value)

    def idle(self) -> bool:
        return bool(_lib().orbfe_server_idle(self._h))


def poisson_arrivals(rates: Sequence[float], duration: float, seed: int):
    """Per-stream Poisson arrival times over [0, duration): list of
    (t, stream, seq) sorted by t."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for s, lam in enumerate(rates):
        t, k = 0.0, 0
        while True:
            t += rng.exponential(1.0 / lam)
            if t >= duration:
                break
            out.append((t, s, k))
            k += 1
    out.sort()
    return out


def run_schedule(server: Server, arrivals, frame_pool, time_scale: float = 1.0):
    """Replays arrivals (t, stream, seq) through server in real time (scaled
    by time_scale; 0 = all waiting at once).  frame_pool[stream] is a list of
    pinned host frames, used cyclically (every stream the same length).
    Without keep_outputs the whole loop runs natively (orbfe_server_run);
    with it, from Python through submit/pump.  Returns (wall seconds,
    completed list)."""
    P = len(frame_pool[0])
    assert all(len(p) == P for p in frame_pool)
    if server.keep_outputs:
        t0 = time.perf_counter()
        i, n = 0, len(arrivals)
        while i < n or not server.idle():
            now = time.perf_counter() - t0
            while i < n and arrivals[i][0] * time_scale <= now:
                _, s, k = arrivals[i]
                server.submit(s, frame_pool[s][k % P])
                i += 1
            if not server.pump() and (i < n or not server.idle()):
                time.sleep(20e-6)
        return time.perf_counter() - t0, server.completed
    arr = (Arrival * max(len(arrivals), 1))(*[Arrival(t, s, k) for t, s, k in arrivals])
    pool = (C.c_void_p * (len(frame_pool) * P))(
        *[f.data_ptr() for frames in frame_pool for f in frames])
    done = (Completion * max(len(arrivals), 1))()
    wall = C.c_double()
    server._check(_lib().orbfe_server_run(server._h, arr, len(arrivals), pool, P, time_scale,
                                          done, C.byref(wall)))
    out = [Completed(c.stream, c.seq, c.t_arrival, c.t_done) for c in done[:len(arrivals)]]
    server.completed.extend(out)
    return float(wall.value), out


def plain_backlog(specs: Sequence[StreamSpec], arrivals, frame_pool, device: int = 0,
                  fast_threshold: float = 0.15, lowe_ratio: float = 0.85) -> float:
    """The plain per-frame path over the same frames: every frame of the
    schedule, in arrival order, uploaded from pinned memory and run through
    its stream's per-frame detect and match launches on one HIP stream, with
    no batching.  Returns the wall seconds to drain the whole backlog."""
    dev = torch.device(f"cuda:{device}")
    ctxs: Dict[Tuple[int, int, int], orbfe.Context] = {}
    state = []
    for sp in specs:
        key = (sp.width, sp.height, sp.cap)
        if key not in ctxs:
            ctxs[key] = orbfe.Context(sp.width, sp.height, device=device,
                                      features_per_level=sp.cap,
                                      fast_threshold=fast_threshold, lowe_ratio=lowe_ratio)
        fb = orbfe.feature_array_bytes(LEVELS, sp.cap)
        state.append((ctxs[key], torch.empty((sp.height, sp.width), dtype=torch.uint8,
                                             device=dev),
                      [torch.zeros(fb, dtype=torch.uint8, device=dev) for _ in range(2)],
                      torch.full((LEVELS * sp.cap,), -1, dtype=torch.int32, device=dev)))
    hs = torch.cuda.current_stream(dev)

    def one(t, s, k):
        c, up, feats, match = state[s]
        pool = frame_pool[s]
        up.copy_(pool[k % len(pool)], non_blocking=True)
        c.detect_async(up, specs[s].width, feats[k & 1], stream=hs)
        if k > 0:
            c.match_pipeline_async(feats[1 - (k & 1)], feats[k & 1], match, stream=hs)

    for a in arrivals[:2 * len(specs)]:  # warm every context
        one(*a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a in arrivals:
        one(*a)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    for c in ctxs.values():
        c.close()
    return wall


def latency_stats(completed: List[Completed]):
    lat = np.array([c.t_done - c.t_arrival for c in completed]) * 1e3
    if len(lat) == 0:
        return {}
    return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "p999_ms": float(np.percentile(lat, 99.9)), "max_ms": float(lat.max()),
            "mean_ms": float(lat.mean()), "frames": int(len(lat))}
