"""Synthetic product workloads (operands as little-endian uint64 limb arrays).

* size_sweep(k): a of exactly 2^k bits, b of floor(2^k / 3) bits, seed k.
* tree_level_replay(): the independent products of one binary-splitting tree
  level, mimicked: bit lengths lognormal with median 2^18 and sigma 0.6,
  clipped to [2^12, 2^24], every third product a square.

All draws come from one splitmix64 stream per workload: first the Box-Muller
uniforms (four per product: two for a's length, two for b's), then the limbs
of every operand in product order (a, then b unless the product squares).
"""
from __future__ import annotations

import numpy as np

from . import draw_operand, limbs_for_bits, operands, splitmix64


def size_sweep(k: int):
    return operands(k, 1 << k, (1 << k) // 3)


def _uniforms(seed: int, start: int, count: int) -> np.ndarray:
    """count doubles in [0, 1) from 53-bit draws of the stream."""
    x = splitmix64(seed, start, count)
    return (x >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def tree_level_sizes(count: int = 4096, seed: int = 7, median_bits: int = 1 << 18, sigma: float = 0.6,
                     lo_bits: int = 1 << 12, hi_bits: int = 1 << 24):
    u = _uniforms(seed, 0, 4 * count).reshape(count, 4)
    # Box-Muller: z = sqrt(-2 ln(1 - u1)) cos(2 pi u2)
    za = np.sqrt(-2.0 * np.log1p(-u[:, 0])) * np.cos(2.0 * np.pi * u[:, 1])
    zb = np.sqrt(-2.0 * np.log1p(-u[:, 2])) * np.cos(2.0 * np.pi * u[:, 3])
    bits_a = np.clip(np.rint(median_bits * np.exp(sigma * za)), lo_bits, hi_bits).astype(np.int64)
    bits_b = np.clip(np.rint(median_bits * np.exp(sigma * zb)), lo_bits, hi_bits).astype(np.int64)
    square = (np.arange(count) % 3) == 2
    bits_b = np.where(square, bits_a, bits_b)
    return bits_a, bits_b, square


def tree_level_replay(count: int = 4096, seed: int = 7, **kw):
    """List of (a, b, square) with b is a for squares."""
    bits_a, bits_b, square = tree_level_sizes(count, seed, **kw)
    pos = 4 * count
    out = []
    for i in range(count):
        a = draw_operand(seed, pos, int(bits_a[i]))
        pos += a.size
        if square[i]:
            out.append((a, a, True))
        else:
            b = draw_operand(seed, pos, int(bits_b[i]))
            pos += b.size
            out.append((a, b, False))
    return out


def total_limbs(products) -> int:
    return sum(a.size + (0 if sq else b.size) for a, b, sq in products)


__all__ = ["size_sweep", "tree_level_sizes", "tree_level_replay", "total_limbs", "limbs_for_bits"]
