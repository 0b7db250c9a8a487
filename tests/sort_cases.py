"""Shared generators and the expected stable order for sort tests."""

from __future__ import annotations

import numpy as np


def xorshift(n: int, salt: int) -> np.ndarray:
    """The reference's key generator (crates/dynamis-sort/tests/gpu.rs:214-224)."""
    out = np.empty(n, dtype=np.uint32)
    state = (0x9E3779B9 ^ salt) & 0xFFFFFFFF
    for i in range(n):
        state ^= (state << 13) & 0xFFFFFFFF
        state ^= state >> 17
        state ^= (state << 5) & 0xFFFFFFFF
        out[i] = state
    return out


def expected_order(major, minor, major_words: int, minor_words: int) -> np.ndarray:
    """Stable order by (low major_words bytes of major, low minor_words bytes of minor), from
    numpy's stable lexsort; input order breaks ties."""
    major = np.asarray(major, dtype=np.uint64)
    minor = np.asarray(minor, dtype=np.uint64)
    mmask = np.uint64((1 << (8 * major_words)) - 1)
    nmask = np.uint64((1 << (8 * minor_words)) - 1)
    return np.lexsort((minor & nmask, major & mmask)).astype(np.uint32)
