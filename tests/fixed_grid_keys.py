"""Host restatement of the fixed-cell broadphase's entries (keys of dyn_fixed_grid): every cell a
box touches, x-major, keyed by (cell hash, box index) and stably sorted."""

from __future__ import annotations

import numpy as np

HASH_MASK = 0x07FFFFFF


def cell_hash(x: np.ndarray, y: np.ndarray, z: np.ndarray) -> np.ndarray:
    """grid_index.wgsl:28-33 on int32 cell coordinates, in uint32 arithmetic."""
    with np.errstate(over="ignore"):
        ux = x.astype(np.int64).astype(np.uint32) * np.uint32(0x9E3779B9)
        uy = y.astype(np.int64).astype(np.uint32) * np.uint32(0x85EBCA77)
        uz = z.astype(np.int64).astype(np.uint32) * np.uint32(0xC2B2AE3D)
        h = ux ^ uy ^ uz ^ (ux << np.uint32(7)) ^ (uy >> np.uint32(3)) ^ (uz << np.uint32(11))
    return (h & np.uint32(HASH_MASK)).astype(np.uint32)


def sorted_entries(boxes: np.ndarray, cell: float) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    c = np.float32(cell)
    lo = np.floor(boxes["min"].astype(np.float32) / c).astype(np.int64)
    hi = np.floor(boxes["max"].astype(np.float32) / c).astype(np.int64)
    span = hi - lo + 1
    count = span[:, 0] * span[:, 1] * span[:, 2]
    hashes, owners, ordinals = [], [], []
    for o in range(int(count.max()) if len(count) else 0):
        sel = np.flatnonzero(count > o)
        sy, sz = span[sel, 1], span[sel, 2]
        x = lo[sel, 0] + o // (sy * sz)
        rem = o % (sy * sz)
        y = lo[sel, 1] + rem // sz
        z = lo[sel, 2] + rem % sz
        hashes.append(cell_hash(x, y, z))
        owners.append(sel.astype(np.uint32))
        ordinals.append(np.full(len(sel), o, np.uint32))
    h = np.concatenate(hashes) if hashes else np.zeros(0, np.uint32)
    b = np.concatenate(owners) if owners else np.zeros(0, np.uint32)
    o = np.concatenate(ordinals) if ordinals else np.zeros(0, np.uint32)
    # the device sorts (hash, box) stably over entries emitted box by box in ordinal order
    order = np.lexsort((o, b, h))
    return h[order], b[order], o[order]
