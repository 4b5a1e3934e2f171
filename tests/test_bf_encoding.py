"""CPU checks of the arithmetic the matrix-core brute-force matcher relies on
(k_bf_mfma): descriptors read as int8 signs, their dot products, the key the
accumulator input turns them into, and the key the split merge decodes.

The operand bytes are built here the way the kernel builds them (16 bits of a
32-bit word per lane half, 4 bytes per int32, bit 7 - j of each byte of the
word shifted left by 4 * half), and the dot products are taken in int32 as the
matrix cores do; the expected values are Hamming distances from XOR and
popcount.  A mistake in the sign convention, the bit selection or the key
scale fails these before any GPU run.
"""
from __future__ import annotations

import numpy as np

KEY_SHIFT = 23          # d << 23 | t: the keys the merge and the C-ABI use
WINDOW = 8192           # train rows per split (13-bit index field)
BIAS = 1 << 20
PAST_END = 1 << 30


def signs16(v: np.ndarray) -> np.ndarray:
    """bm_signs16: (..., ) u32 words (already shifted left by 4 * half) ->
    (..., 16) int8, byte j*4+k = +64 if bit (8k + 7 - j) of v is 0 else -64."""
    v = np.asarray(v, np.uint64) & 0xFFFFFFFF
    out = np.empty(v.shape + (16,), np.int8)
    for j in range(4):
        word = ((v << j) & 0x80808080) | 0x40404040
        for k in range(4):
            out[..., 4 * j + k] = ((word >> (8 * k)) & 0xFF).astype(np.uint8).view(np.int8)
    return out


def operand(desc: np.ndarray, query_side: bool) -> np.ndarray:
    """(n, 8) u32 descriptors -> (n, 8 steps, 2 halves, 16) int8 operand
    bytes: train words as they are, query words inverted."""
    w = desc.astype(np.uint64)
    if query_side:
        w = ~w & 0xFFFFFFFF
    halves = [signs16((w << (4 * h)) & 0xFFFFFFFF) for h in (0, 1)]
    return np.stack(halves, axis=2)


def hamming(q: np.ndarray, t: np.ndarray) -> np.ndarray:
    x = (q[:, None, :] ^ t[None, :, :]).astype(np.uint32)
    return np.unpackbits(x.view(np.uint8), axis=-1).sum(-1).astype(np.int64)


def dots(q: np.ndarray, t: np.ndarray) -> np.ndarray:
    a = operand(t, False).reshape(len(t), -1).astype(np.int32)
    b = operand(q, True).reshape(len(q), -1).astype(np.int32)
    return b @ a.T  # (nq, nt) int32, exact


def global_key(k: int, tb: int, t0: int) -> int:
    """bm_global_key: tile-relative key -> d << 23 | t (0xFFFFFFFF: none)."""
    k = int(k)
    if k >= PAST_END - BIAS - WINDOW:
        return 0xFFFFFFFF
    u = (k + BIAS + (tb - t0)) & 0xFFFFFFFF
    return ((u >> 13) << KEY_SHIFT) | (t0 + (u & (WINDOW - 1)))


def test_every_bit_lands_in_exactly_one_byte():
    # one set bit at a time: exactly one of the 32 operand bytes (both halves)
    # flips sign, and each bit maps to a different byte
    seen = set()
    base = operand(np.zeros((1, 8), np.uint32), False)[0, 0]
    for b in range(32):
        d = np.zeros((1, 8), np.uint32)
        d[0, 0] = np.uint32(1 << b)
        diff = np.argwhere(operand(d, False)[0, 0] != base)
        assert len(diff) == 1
        seen.add(tuple(diff[0]))
    assert len(seen) == 32


def test_dot_product_is_affine_in_hamming_distance():
    rng = np.random.default_rng(11)
    q = rng.integers(0, 2**32, (40, 8), dtype=np.uint64).astype(np.uint32)
    t = rng.integers(0, 2**32, (50, 8), dtype=np.uint64).astype(np.uint32)
    t[:3] = q[:3]                    # d = 0
    t[3] = ~q[3]                     # d = 256
    t[4] = q[4] ^ np.uint32(1)       # d = 1
    assert np.array_equal(dots(q, t), 8192 * hamming(q, t) - BIAS)


def test_keys_order_as_distance_then_index():
    # the accumulator input (row offset) makes key = 8192 d + t - 2^20, which
    # sorts as (d, t) for t < 8192; over a split of tiles the kept keys are
    # lowered by 32 per tile and decoded back to d << 23 | t
    rng = np.random.default_rng(12)
    q = rng.integers(0, 2**32, (6, 8), dtype=np.uint64).astype(np.uint32)
    nt = 300                                       # 9 full tiles + 12 rows
    t = rng.integers(0, 2**32, (nt, 8), dtype=np.uint64).astype(np.uint32)
    t[[7, 40, 299]] = q[0]                         # ties at d = 0 across tiles
    t0 = 4096                                      # the split's first row
    d = hamming(q, t)
    prod = dots(q, t)
    for i in range(len(q)):
        m1 = m2 = 0x7FFFFFFF
        tiles = (nt + 31) // 32
        for k in range(tiles):
            tb = 32 * k
            for o in range(32):
                # rows past the end read the last row again, plus the bias
                key = int(prod[i, min(tb + o, nt - 1)]) + o + (PAST_END if tb + o >= nt else 0)
                m1, m2 = min(m1, key), sorted((m1, m2, key))[1]
            if k + 1 < tiles:
                m1 -= 32
                m2 -= 32
        tb_last = t0 + 32 * (tiles - 1)
        order = sorted(range(nt), key=lambda j: (d[i, j], j))
        want1 = (int(d[i, order[0]]) << KEY_SHIFT) | (t0 + order[0])
        want2 = (int(d[i, order[1]]) << KEY_SHIFT) | (t0 + order[1])
        assert global_key(m1, tb_last, t0) == want1
        assert global_key(m2, tb_last, t0) == want2
    assert global_key(0x7FFFFFFF - 32 * 255, t0, t0) == 0xFFFFFFFF   # nothing kept


def test_one_train_row_leaves_no_second_best():
    q = np.zeros((1, 8), np.uint32)
    t = np.full((1, 8), 0xFFFFFFFF, np.uint32)
    prod = int(dots(q, t)[0, 0])
    keys = [prod] + [prod + o + PAST_END for o in range(1, 32)]
    m1, m2 = sorted(keys)[:2]
    assert global_key(m1, 0, 0) == (256 << KEY_SHIFT)
    assert global_key(m2, 0, 0) == 0xFFFFFFFF
