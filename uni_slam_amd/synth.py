"""Seeded synthetic inputs for the ORB front-end workloads.

The reference reads EuRoC PNGs (data/src/provider_base.cpp:90-128); the
workloads here replace them with a fully specified generator:

* scene: 4-octave value noise in [0, 255] (lattice spacings 64, 32, 16, 8 px,
  amplitudes 1, 1/2, 1/4, 1/8, lattice values uniform in [0, 1), smoothstep
  interpolation, normalised by the amplitude sum), overlaid with filled
  axis-aligned rectangles (sides uniform in [8, 80] px, intensity uniform in
  [0, 255]); 200 rectangles per frame area, scaled with the canvas area;
* a frame is a W x H crop of the scene plus fresh N(0, 3) noise, rounded half
  to even and clamped to [0, 255], as u8;
* a pair is frame A at the canvas origin and frame B shifted by an integer
  (dx, dy) uniform in [-8, 8]^2;
* a sequence shifts each frame by (dx, dy) uniform in [-4, 4]^2 from the
  previous one.

All randomness comes from numpy's PCG64 generator seeded with the given seed,
so a seed fixes every byte.
"""
from __future__ import annotations

import numpy as np

RESOLUTIONS = ((640, 480), (752, 480), (1280, 720), (1920, 1080))


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def _smoothstep(t: np.ndarray) -> np.ndarray:
    return t * t * (3.0 - 2.0 * t)


def value_noise(rng: np.random.Generator, width: int, height: int,
                octaves: int = 4, base_cell: int = 64) -> np.ndarray:
    """Value noise in [0, 1] of shape (height, width), float64."""
    out = np.zeros((height, width), dtype=np.float64)
    amp_sum = 0.0
    for k in range(octaves):
        cell = base_cell >> k
        amp = 0.5 ** k
        gw = width // cell + 2
        gh = height // cell + 2
        lattice = rng.random((gh, gw))
        xs = np.arange(width, dtype=np.float64) / cell
        ys = np.arange(height, dtype=np.float64) / cell
        x0 = np.floor(xs).astype(np.int64)
        y0 = np.floor(ys).astype(np.int64)
        tx = _smoothstep(xs - x0)[None, :]
        ty = _smoothstep(ys - y0)[:, None]
        v00 = lattice[y0[:, None], x0[None, :]]
        v10 = lattice[y0[:, None], x0[None, :] + 1]
        v01 = lattice[y0[:, None] + 1, x0[None, :]]
        v11 = lattice[y0[:, None] + 1, x0[None, :] + 1]
        top = v00 + (v10 - v00) * tx
        bot = v01 + (v11 - v01) * tx
        out += amp * (top + (bot - top) * ty)
        amp_sum += amp
    return out / amp_sum


def scene(seed: int, width: int, height: int, margin: int = 0,
          frame_area: int | None = None) -> np.ndarray:
    """Clean scene canvas of shape (height + 2*margin, width + 2*margin)."""
    rng = _rng(seed)
    cw = width + 2 * margin
    ch = height + 2 * margin
    canvas = value_noise(rng, cw, ch) * 255.0
    area = frame_area if frame_area is not None else width * height
    n_rect = int(round(200 * (cw * ch) / area))
    xs = rng.integers(0, cw, size=n_rect)
    ys = rng.integers(0, ch, size=n_rect)
    ws = rng.integers(8, 81, size=n_rect)
    hs = rng.integers(8, 81, size=n_rect)
    vals = rng.integers(0, 256, size=n_rect)
    for x, y, w, h, v in zip(xs, ys, ws, hs, vals):
        canvas[y:y + h, x:x + w] = float(v)
    return canvas


def render(canvas: np.ndarray, x_off: int, y_off: int, width: int,
           height: int, rng: np.random.Generator,
           sigma: float = 3.0) -> np.ndarray:
    """Crop + N(0, sigma) noise, round half to even, clamp, u8."""
    crop = canvas[y_off:y_off + height, x_off:x_off + width]
    noisy = crop + rng.normal(0.0, sigma, size=crop.shape)
    return np.clip(np.rint(noisy), 0, 255).astype(np.uint8)


def frame_pair(seed: int, width: int = 752, height: int = 480,
               max_shift: int = 8):
    """Frame A and its translated twin B (config-0 generator).

    Returns (A, B, (dx, dy)) with A, B u8 arrays of shape (height, width)."""
    m = max_shift
    canvas = scene(seed, width, height, margin=m)
    rng = _rng(seed ^ 0x5EED_0001)
    dx, dy = (int(v) for v in rng.integers(-m, m + 1, size=2))
    a = render(canvas, m, m, width, height, rng)
    b = render(canvas, m + dx, m + dy, width, height, rng)
    return a, b, (dx, dy)


def sequence(seed: int, width: int = 1280, height: int = 720,
             n_frames: int = 32, max_shift: int = 4):
    """A sequence in which each frame is the previous one shifted by an
    integer (dx, dy) uniform in [-max_shift, max_shift]^2 plus fresh noise.

    Returns (frames u8 array (n, height, width), shifts list)."""
    m = max_shift * max(n_frames - 1, 1)
    canvas = scene(seed, width, height, margin=m, frame_area=width * height)
    rng = _rng(seed ^ 0x5EED_0002)
    x, y = m, m
    frames = np.empty((n_frames, height, width), dtype=np.uint8)
    shifts = []
    for t in range(n_frames):
        if t > 0:
            dx, dy = (int(v) for v in rng.integers(-max_shift, max_shift + 1, size=2))
            x += dx
            y += dy
            shifts.append((dx, dy))
        frames[t] = render(canvas, x, y, width, height, rng)
    return frames, shifts


def mixed_batch(seed: int, n_frames: int = 64, cap_lo: int = 1000,
                cap_hi: int = 4000):
    """Mixed-resolution batch: per frame a resolution uniform over
    RESOLUTIONS, a keypoint cap uniform in [cap_lo, cap_hi] and a translated
    twin.  Returns a list of dicts {a, b, cap, shift}."""
    rng = _rng(seed)
    out = []
    for i in range(n_frames):
        w, h = RESOLUTIONS[int(rng.integers(0, len(RESOLUTIONS)))]
        cap = int(rng.integers(cap_lo, cap_hi + 1))
        a, b, shift = frame_pair(seed * 1000003 + i, w, h)
        out.append({"a": a, "b": b, "cap": cap, "shift": shift})
    return out


def descriptors(seed: int, n_query: int, n_train: int, dup_frac: float = 0.3,
                p_flip: float = 0.05):
    """Matcher stress input: uniform random 256-bit train descriptors; a
    `dup_frac` share of the queries are copies of a random train descriptor
    with each bit flipped with probability `p_flip`, the rest uniform.

    Returns (queries u32 (n_query, 8), train u32 (n_train, 8),
    planted mask, planted train index (-1 where not planted))."""
    rng = _rng(seed)
    train = rng.integers(0, 2**32, size=(n_train, 8), dtype=np.uint64).astype(np.uint32)
    query = rng.integers(0, 2**32, size=(n_query, 8), dtype=np.uint64).astype(np.uint32)
    planted = rng.random(n_query) < dup_frac
    src = rng.integers(0, max(n_train, 1), size=n_query)
    src = np.where(planted, src, -1)
    if n_train > 0 and planted.any():
        flips = rng.random((n_query, 256)) < p_flip
        bits = np.packbits(flips.astype(np.uint8), axis=1, bitorder="little")
        flip_words = bits.view(np.uint32).reshape(n_query, 8)
        idx = np.nonzero(planted)[0]
        query[idx] = train[src[idx]] ^ flip_words[idx]
    return query, train, planted, src
