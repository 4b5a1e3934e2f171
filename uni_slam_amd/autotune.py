"""Launch-shape autotuner of the batched detect and match stages.

Three launch shapes of liborbfe are free parameters that leave every output
bit unchanged: the rows of a detect tile (8, 16 or 32; LDS per workgroup and
so occupancy, halo overhead and tile count move with it), the lanes that
describe one keypoint (8, 16 or 32; fewer lanes do more tests each, more
lanes repeat the per-keypoint moment sum more often) and the lanes that
match one query (16, 32 or 64; fewer lanes walk longer candidate lists,
more lanes leave more of them idle on short lists).  Which shape is fastest
depends on the mix of frame sizes and caps in a batch, so the tuner times
every shape on the caller's own batch: a prepared batch per shape, captured
in a HIP graph, replayed `reps` times between two events.  Each variant's
outputs must equal the default shape's byte for byte; a difference raises
AutotuneError (it would be a kernel bug, never a tuning outcome).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Sequence, Tuple

from . import orbfe

TILE_ROWS = (8, 16, 32)
DESCRIBE_LANES = (8, 16, 32)
MATCH_LANES = (16, 32, 64)
DEFAULT = (16, 16)
MATCH_DEFAULT = 16
SHAPES: List[Tuple[int, int]] = [(t, d) for t in TILE_ROWS for d in DESCRIBE_LANES]


class AutotuneError(RuntimeError):
    pass


@dataclass
class TuneResult:
    best: Tuple[int, int]                       # (detect_tile_rows, describe_lanes)
    seconds: Dict[Tuple[int, int], float] = field(default_factory=dict)  # per replay

    @property
    def speedup_vs_default(self) -> float:
        return self.seconds[DEFAULT] / self.seconds[self.best]

    def config(self) -> dict:
        """Keyword overrides for orbfe.Context."""
        return {"detect_tile_rows": self.best[0], "describe_lanes": self.best[1]}


def _capture(torch, launch):
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        launch(side)  # warm-up outside capture
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        launch(torch.cuda.current_stream())
    return g


def tune_detect(frames: Sequence[dict], width: int, height: int, device: int = 0,
                reps: int = 10, shapes: Sequence[Tuple[int, int]] = SHAPES,
                **config) -> TuneResult:
    """Times the batched detect stage of `frames` (frame dicts of device
    tensors, as orbfe.Context.batch_create takes them; their feature arrays
    are overwritten) under each launch shape and returns the fastest.
    `width`, `height` and `config` are the context's orbfe_config."""
    import torch

    shapes = list(shapes)
    if DEFAULT not in shapes:
        shapes.insert(0, DEFAULT)
    else:
        shapes.remove(DEFAULT)
        shapes.insert(0, DEFAULT)
    want = None
    res = TuneResult(best=DEFAULT)
    for shape in shapes:
        ctx = orbfe.Context(width, height, device=device, detect_tile_rows=shape[0],
                            describe_lanes=shape[1], **config)
        try:
            b = ctx.batch_create(frames)
            g = _capture(torch, lambda st: b.detect_async(stream=st))
            g.replay()
            torch.cuda.synchronize()
            got = [f["features"].clone() for f in frames]
            if want is None:
                want = got
            else:
                for i, (a, c) in enumerate(zip(want, got)):
                    if not torch.equal(a, c):
                        raise AutotuneError(
                            f"launch shape {shape} changed the features of frame {i}")
            t0 = torch.cuda.Event(enable_timing=True)
            t1 = torch.cuda.Event(enable_timing=True)
            t0.record()
            for _ in range(reps):
                g.replay()
            t1.record()
            torch.cuda.synchronize()
            res.seconds[shape] = t0.elapsed_time(t1) / 1e3 / reps
            del g
            b.close()
        finally:
            ctx.close()
    res.best = min(res.seconds, key=res.seconds.get)
    return res


@dataclass
class MatchTuneResult:
    best: int                                   # match_lanes
    seconds: Dict[int, float] = field(default_factory=dict)  # per replay

    def config(self) -> dict:
        """Keyword overrides for orbfe.Context."""
        return {"match_lanes": self.best}


def tune_match(pairs: Sequence[dict], width: int, height: int, device: int = 0,
               reps: int = 10, lanes: Sequence[int] = MATCH_LANES, **config) -> MatchTuneResult:
    """Times the batched match stage of `pairs` (pair dicts of device
    tensors, as orbfe.Context.match_batch_create takes them; their match
    tables are overwritten) under each lanes-per-query shape and returns the
    fastest.  `width`, `height` and `config` are the context's orbfe_config."""
    import torch

    order = [MATCH_DEFAULT] + [n for n in lanes if n != MATCH_DEFAULT]
    want = None
    res = MatchTuneResult(best=MATCH_DEFAULT)
    for n in order:
        ctx = orbfe.Context(width, height, device=device, match_lanes=n, **config)
        try:
            mb = ctx.match_batch_create(pairs)
            g = _capture(torch, lambda st: mb.match_async(stream=st))
            g.replay()
            torch.cuda.synchronize()
            got = [p["prev_match"].clone() for p in pairs]
            if want is None:
                want = got
            else:
                for i, (a, c) in enumerate(zip(want, got)):
                    if not torch.equal(a, c):
                        raise AutotuneError(f"{n} lanes per query changed the matches of pair {i}")
            t0 = torch.cuda.Event(enable_timing=True)
            t1 = torch.cuda.Event(enable_timing=True)
            t0.record()
            for _ in range(reps):
                g.replay()
            t1.record()
            torch.cuda.synchronize()
            res.seconds[n] = t0.elapsed_time(t1) / 1e3 / reps
            del g
            mb.close()
        finally:
            ctx.close()
    res.best = min(res.seconds, key=res.seconds.get)
    return res
