"""Synthetic box scenes and the brute-force pair set for broadphase tests and the benchmark."""

from __future__ import annotations

import numpy as np

from dynamis_amd import abi
from dynamis_amd.scenes import uniform_boxes  # noqa: F401


def brute_force_pairs(boxes: np.ndarray, groups=None, flags=None) -> np.ndarray:
    """Every (i < j) pair whose closed boxes overlap, from distinct owners, with at least one
    mobile and at least one awake collider: the set the grid must reproduce exactly. O(K^2)."""
    k = len(boxes)
    mn = boxes["min"].astype(np.float32)
    mx = boxes["max"].astype(np.float32)
    g = np.arange(k) if groups is None else np.asarray(groups)
    f = np.full(k, abi.BP_MOBILE | abi.BP_AWAKE) if flags is None else np.asarray(flags)
    mobile = (f & abi.BP_MOBILE) != 0
    awake = (f & abi.BP_AWAKE) != 0
    out = []
    for i in range(k - 1):
        j = np.arange(i + 1, k)
        hit = np.all(mn[i] <= mx[j], axis=1) & np.all(mn[j] <= mx[i], axis=1)
        hit &= g[j] != g[i]
        hit &= mobile[j] | mobile[i]
        hit &= awake[j] | awake[i]
        for jj in j[hit]:
            out.append((i, jj))
    if not out:
        return np.zeros((0, 2), np.uint32)
    return np.array(sorted(out), dtype=np.uint32)


def canonical_pairs(pairs: np.ndarray) -> np.ndarray:
    """Rows sorted lexicographically."""
    pairs = np.asarray(pairs, dtype=np.uint32).reshape(-1, 2)
    order = np.lexsort((pairs[:, 1], pairs[:, 0]))
    return pairs[order]


def entry_tuples(keys: np.ndarray, order: np.ndarray, entries: np.ndarray) -> np.ndarray:
    """Sorted (key, collider, cell offset) triples of a sorted entry region: the canonical form
    of the entries whatever order equal keys were emitted in."""
    info = entries["info"][order]
    collider = info & abi.ENTRY_INDEX_MASK
    offset = (info & abi.ENTRY_CELL_MASK) >> abi.ENTRY_CELL_SHIFT
    t = np.stack([keys.astype(np.uint64), collider.astype(np.uint64), offset.astype(np.uint64)], 1)
    return t[np.lexsort((t[:, 2], t[:, 1], t[:, 0]))]
