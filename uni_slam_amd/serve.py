"""Continuous-batching front-end server for many camera streams on one GPU.

Frames arrive from many streams at their own times.  The server keeps, per
stream, a ring of slots resident in HBM (upload buffer, feature array,
pyramid and match table; frame k uses slot k mod slots), so a stream's
previous-frame descriptors stay on the device.  Each scheduling tick takes
the frames that have arrived, admits the oldest ones (a few consecutive
frames per stream at most, each stream in its own order) as far as the HBM
budget and the in-flight slots allow, uploads them and launches one batched
detect (all admitted frames, any mix of sizes) and one batched match (each
frame against its stream's previous frame) on one of two HIP streams.  A
batch's uploads overlap the kernels of the batch before it; its kernels wait
(on the device) for the batches holding its streams' previous frames.
Latency = completion - arrival.

Results are bit-identical to the plain per-frame path: the batched kernels
compute each frame exactly as the per-frame launch does.
"""
from __future__ import annotations

import collections
import gc
import heapq
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import orbfe

LEVELS = 6


def frame_scratch_bytes(width: int, height: int, cap: int) -> int:
    """Device scratch one in-flight frame uses in a batch, as laid out by
    liborbfe: survivor bitmap, keypoint list and smoothed image of the
    detect stage; match table, query list and the two sides' cell indices
    (cell starts and the records sorted by cell) of the match stage."""
    dims = orbfe.level_dims(width, height)
    words = sum(h * ((w + 63) // 64) for w, h in dims)
    smooth = sum((w + 4) * (h + 4) for w, h in dims)

    def al(v):
        return (v + 255) // 256 * 256

    def al4(v):
        return (v + 3) // 4 * 4
    idx_words = al4(48 * 24 + 1) + 12 * cap
    detect = al(8 * words) + al(4 * LEVELS * cap) + al(4 * smooth)
    qlist_words = 64 * 32 + LEVELS * cap + 64 * LEVELS + 64 * 64
    match = al(4 * LEVELS * cap) + al(4 * qlist_words) + 2 * al(4 * LEVELS * idx_words)
    return detect + match


@dataclass
class StreamSpec:
    width: int
    height: int
    cap: int


@dataclass
class _Request:
    stream: int
    seq: int            # frame number within its stream
    t_arrival: float
    host: torch.Tensor  # pinned u8 (H, W)
    order: int = 0      # submission number (FIFO admission key)


@dataclass
class _Stream:
    spec: StreamSpec
    feats: List[torch.Tensor]     # per slot
    pyramid: List[torch.Tensor]   # per slot
    match: List[torch.Tensor]     # per slot: the slot's frame against the frame before
    upload: List[torch.Tensor]    # per slot
    scratch: int
    # prebuilt C descriptors of the frame / pair of each slot
    frame_c: List[orbfe.Frame] = field(default_factory=list)
    pair_c: List[orbfe.MatchPair] = field(default_factory=list)
    pending: collections.deque = field(default_factory=collections.deque)
    events: Dict[int, torch.cuda.Event] = field(default_factory=dict)  # seq -> its batch


@dataclass
class Completed:
    stream: int
    seq: int
    t_arrival: float
    t_done: float


def _frame_struct(st: _Stream, p: int) -> orbfe.Frame:
    sp = st.spec
    f = orbfe.Frame()
    f.image = st.upload[p].data_ptr()
    f.pitch = sp.width
    f.width, f.height, f.cap = sp.width, sp.height, sp.cap
    f.pyramid = st.pyramid[p].data_ptr()
    f.features = st.feats[p].data_ptr()
    return f


def _pair_struct(st: _Stream, p: int) -> orbfe.MatchPair:
    """Pair of the frame in slot p against the stream's previous frame."""
    sp = st.spec
    n = len(st.feats)
    m = orbfe.MatchPair()
    m.prev_features = st.feats[(p - 1) % n].data_ptr()
    m.curr_features = st.feats[p].data_ptr()
    m.prev_match = st.match[p].data_ptr()
    m.width, m.height, m.cap = sp.width, sp.height, sp.cap
    return m


class Server:
    """frames_per_stream: how many consecutive frames of one stream a batch
    may hold (a stream with a backlog catches up in fewer, fuller batches).
    Each stream keeps 2 * frames_per_stream + 2 slots (upload, features,
    pyramid, match table) resident; frame k uses slot k mod slots."""

    def __init__(self, specs: Sequence[StreamSpec], device: int = 0,
                 hbm_budget: int = 1 << 30, max_batch: int = 64, inflight: int = 2,
                 fast_threshold: float = 0.15, lowe_ratio: float = 0.85,
                 keep_outputs: bool = False, warmup: bool = True, log_batches: bool = False,
                 frames_per_stream: int = 3, **launch_shape):
        self.dev = torch.device(f"cuda:{device}")
        self.specs = list(specs)
        self.budget = hbm_budget
        self.max_batch = max_batch
        self.per_stream = max(1, frames_per_stream)
        self.slots = 2 * self.per_stream + 2
        self.keep_outputs = keep_outputs
        # with log_batches: per launched batch, its (stream, seq, submission number)
        self.batch_log: Optional[List[List[Tuple[int, int, int]]]] = [] if log_batches else None
        self.outputs: Dict[Tuple[int, int], Tuple[np.ndarray, Optional[np.ndarray]]] = {}
        self._snap: Dict[Tuple[int, int], Tuple[torch.Tensor, Optional[torch.Tensor]]] = {}
        w0, h0 = self.specs[0].width, self.specs[0].height
        self.ctxs = [orbfe.Context(w0, h0, device=device, fast_threshold=fast_threshold,
                                   lowe_ratio=lowe_ratio, **launch_shape)
                     for _ in range(inflight)]
        self.hip_streams = [torch.cuda.Stream(device=self.dev) for _ in range(inflight)]
        self.busy: List[Optional[Tuple[torch.cuda.Event, List[_Request], int]]] = [None] * inflight
        self.streams: List[_Stream] = []
        n = self.slots
        for sp in self.specs:
            fb = orbfe.feature_array_bytes(LEVELS, sp.cap)
            npyr = sum(w * h for w, h in orbfe.level_dims(sp.width, sp.height)[1:])
            st = _Stream(
                spec=sp,
                feats=[torch.zeros(fb, dtype=torch.uint8, device=self.dev) for _ in range(n)],
                pyramid=[torch.zeros(npyr, dtype=torch.float32, device=self.dev)
                         for _ in range(n)],
                match=[torch.full((LEVELS * sp.cap,), -1, dtype=torch.int32, device=self.dev)
                       for _ in range(n)],
                upload=[torch.empty((sp.height, sp.width), dtype=torch.uint8, device=self.dev)
                        for _ in range(n)],
                scratch=frame_scratch_bytes(sp.width, sp.height, sp.cap))
            st.frame_c = [_frame_struct(st, p) for p in range(n)]
            st.pair_c = [_pair_struct(st, p) for p in range(n)]
            self.streams.append(st)
        # streams with pending frames, keyed by their oldest frame's submission
        # number: popping in key order is FIFO admission over all frames, a
        # stream's frames in its own order
        self._ready: List[Tuple[int, int]] = []
        self._order = 0
        self.completed: List[Completed] = []
        self.in_flight_bytes = 0
        self.max_batch_seen = 0
        self.batches = 0
        if warmup:
            self.warmup()

    def warmup(self):
        """Creates every stream's plan in each in-flight context and grows
        the contexts' scratch arenas and pinned staging slots to the largest
        batch admission can form (frames_per_stream frames of every stream),
        so serving never allocates device or pinned memory.  Stream state is
        reset."""
        k = self.per_stream
        frames = [st.frame_c[p] for st in self.streams for p in range(1, k + 1)]
        pairs = [st.pair_c[p] for st in self.streams for p in range(1, k + 1)]
        frames = (orbfe.Frame * len(frames))(*frames)
        pairs = (orbfe.MatchPair * len(pairs))(*pairs)
        for ctx, hs in zip(self.ctxs, self.hip_streams):
            with torch.cuda.stream(hs):
                for _ in range(4):  # each slot of the staging ring
                    ctx.detect_batch_async(frames, stream=hs)
                    ctx.match_batch_async(pairs, stream=hs)
        torch.cuda.synchronize()
        for st in self.streams:
            for f in st.feats:
                f.zero_()
            for m in st.match:
                m.fill_(-1)
        torch.cuda.synchronize()

    # -- requests --------------------------------------------------------------
    def submit(self, stream: int, seq: int, host_frame: torch.Tensor, t_arrival: float):
        st = self.streams[stream]
        r = _Request(stream, seq, t_arrival, host_frame, self._order)
        self._order += 1
        st.pending.append(r)
        if len(st.pending) == 1:
            heapq.heappush(self._ready, (r.order, stream))

    def idle(self) -> bool:
        return not self._ready and all(b is None for b in self.busy)

    # -- scheduling -------------------------------------------------------------
    def poll(self, now_fn=time.perf_counter):
        """Retire finished batches; record their completion time."""
        for i, b in enumerate(self.busy):
            if b is None:
                continue
            ev, reqs, nbytes = b
            if ev.query():
                t = now_fn()
                for r in reqs:
                    self.completed.append(Completed(r.stream, r.seq, r.t_arrival, t))
                    if self.keep_outputs:
                        f, m = self._snap.pop((r.stream, r.seq))
                        self.outputs[(r.stream, r.seq)] = (
                            f.cpu().numpy(), None if m is None else m.cpu().numpy())
                self.in_flight_bytes -= nbytes
                self.busy[i] = None

    def _admit(self) -> Tuple[List[_Request], int]:
        """FIFO admission: the oldest pending frames, at most frames_per_stream
        per stream (a stream's frames stay in order), up to max_batch, within
        the HBM budget (a frame is always admitted when nothing else is in
        flight)."""
        chosen: List[_Request] = []
        nbytes = 0
        ready = self._ready
        taken: Dict[int, int] = {}
        deferred: List[int] = []  # streams at their per-batch limit with frames left
        while ready and len(chosen) < self.max_batch:
            s = ready[0][1]
            st = self.streams[s]
            if (self.in_flight_bytes + nbytes + st.scratch > self.budget
                    and (chosen or self.in_flight_bytes > 0)):
                break
            heapq.heappop(ready)
            chosen.append(st.pending.popleft())
            nbytes += st.scratch
            taken[s] = taken.get(s, 0) + 1
            if st.pending:
                if taken[s] < self.per_stream:
                    heapq.heappush(ready, (st.pending[0].order, s))
                else:
                    deferred.append(s)
        for s in deferred:  # back in line for later batches
            heapq.heappush(ready, (self.streams[s].pending[0].order, s))
        return chosen, nbytes

    def launch(self) -> bool:
        """Admit pending frames into a free slot; returns True if launched.

        Device order on the slot's HIP stream: each frame's upload (after the
        batches that last used its stream slot: frames k - slots and
        k - slots + 1), then a wait for the batches holding the admitted
        frames' previous frames (their features are this batch's match
        input), then one batched detect and one batched match.  The detect
        stage of a batch completes before its match stage, so a stream's
        consecutive frames can share a batch.  The uploads of a batch overlap
        the kernels of the batch before it."""
        if not self._ready:
            return False
        try:
            slot = self.busy.index(None)
        except ValueError:
            return False
        chosen, nbytes = self._admit()
        if not chosen:
            return False
        ctx, hs = self.ctxs[slot], self.hip_streams[slot]
        n = self.slots
        frames, pairs = [], []
        waited = set()
        in_batch = {(r.stream, r.seq) for r in chosen}

        def wait(e):
            if e is not None and id(e) not in waited:
                hs.wait_event(e)
                waited.add(id(e))

        with torch.cuda.stream(hs):
            for r in chosen:
                st = self.streams[r.stream]
                p = r.seq % n
                wait(st.events.get(r.seq - n))
                wait(st.events.get(r.seq - n + 1))
                st.upload[p].copy_(r.host, non_blocking=True)
                frames.append(st.frame_c[p])
                if r.seq > 0:
                    pairs.append(st.pair_c[p])
            for r in chosen:
                if (r.stream, r.seq - 1) not in in_batch:
                    wait(self.streams[r.stream].events.get(r.seq - 1))
            ctx.detect_batch_async((orbfe.Frame * len(frames))(*frames), stream=hs)
            if pairs:
                ctx.match_batch_async((orbfe.MatchPair * len(pairs))(*pairs), stream=hs)
            if self.keep_outputs:  # stream-ordered snapshots of the results
                for r in chosen:
                    st = self.streams[r.stream]
                    p = r.seq % n
                    self._snap[(r.stream, r.seq)] = (
                        st.feats[p].clone(), st.match[p].clone() if r.seq > 0 else None)
            ev = torch.cuda.Event()
            ev.record(hs)
        for r in chosen:
            st = self.streams[r.stream]
            st.events[r.seq] = ev
            st.events.pop(r.seq - n, None)  # no later frame waits for it
        self.busy[slot] = (ev, chosen, nbytes)
        if self.batch_log is not None:
            self.batch_log.append([(r.stream, r.seq, r.order) for r in chosen])
        self.in_flight_bytes += nbytes
        self.max_batch_seen = max(self.max_batch_seen, len(chosen))
        self.batches += 1
        return True


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


def run_schedule(server: Server, arrivals, frame_pool, time_scale: float = 1.0,
                 spin_s: float = 20e-6):
    """Replays arrivals in real time (scaled by time_scale) through server.
    frame_pool[stream] is a list of pinned host frames, used cyclically.
    Returns (wall seconds, completed list).  Python's cyclic garbage
    collector is paused for the run (its pauses would land in the latency
    tail)."""
    gc_was_enabled = gc.isenabled()
    gc.collect()
    gc.disable()
    try:
        return _run(server, arrivals, frame_pool, time_scale, spin_s)
    finally:
        if gc_was_enabled:
            gc.enable()


def _run(server, arrivals, frame_pool, time_scale, spin_s):
    t0 = time.perf_counter()
    i = 0
    n = len(arrivals)
    while i < n or not server.idle():
        now = time.perf_counter() - t0
        while i < n and arrivals[i][0] * time_scale <= now:
            ta, s, k = arrivals[i]
            pool = frame_pool[s]
            server.submit(s, k, pool[k % len(pool)], t0 + ta * time_scale)
            i += 1
        server.poll()
        launched = server.launch()
        if not launched and (i < n or not server.idle()):
            time.sleep(spin_s)
    wall = time.perf_counter() - t0
    return wall, server.completed


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
