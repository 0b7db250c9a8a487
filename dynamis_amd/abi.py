"""numpy views of the device records and constants of include/dyn_abi.h (the reference's
crates/dynamis-abi/abi/records.wgsl layouts and src/constant.rs / src/counter.rs values)."""

from __future__ import annotations

import numpy as np

AABB = np.dtype([("min", "<f4", 3), ("_pad0", "<f4"), ("max", "<f4", 3), ("_pad1", "<f4")])
GRID_ENTRY = np.dtype([("min", "<f4", 3), ("group", "<u4"), ("max", "<f4", 3), ("info", "<u4")])

assert AABB.itemsize == 32 and GRID_ENTRY.itemsize == 32

# constant.rs:67-87
MAX_CELLS_PER_AXIS = 2
MAX_CELLS_PER_COLLIDER = 8
ENTRY_INDEX_MASK = (1 << 24) - 1
ENTRY_CELL_SHIFT = 24
ENTRY_CELL_MASK = 7 << 24
ENTRY_KIND_SHIFT = 27
ENTRY_KIND_MASK = 3 << 27
ENTRY_MOBILE = 1 << 29
ENTRY_AWAKE = 1 << 30
ENTRY_PRIMARY = 1 << 31
LEVEL_KEY_SHIFT = 27
CELL_HASH_MASK = 0x07FF_FFFF

BP_MOBILE = 1
BP_AWAKE = 2

# counter.rs:75-128 slot order
COUNTER_NAMES = [
    "ENTRIES", "RESTING_ENTRIES", "RESTING_LEVELS", "RESTING_REBUILD", "RESTING_SORT",
    "AWAKE_BASE", "IMMOVABLE_ENTRIES", "IMMOVABLE_EMITTED", "IMMOVABLE_LEVELS", "ENTRY_BASE",
    "PAIRS", "GRID_LEVELS", "CONTACTS", "ARCHIVED", "JOINTS", "EVENTS", "REFUSED_PAIRS",
    "REFUSED_EVENTS", "ENTRY_FAULTS", "RESTING", "REFUSED_RESTING", "SLEPT", "WOKE", "ACTIVE",
    "WOKE_DEFERRED", "RESTING_GATHER", "RESTING_INDEX", "RESTING_PENDING", "ISLANDS",
    "ISLAND_FAULTS", "BLOCKS", "SOLVER_ROWS", "SOLVE_LINEAR_RESIDUAL",
    "SOLVE_ANGULAR_RESIDUAL", "COARSE_ACTIVE", "GRID_SCALE", "GRID_EXTENT", "BREAKS",
    "PARTICLE_REACH", "COARSE_NEIGHBOURS", "LIVE", "LIVE_FAULTS", "ROW_FAULTS", "SOFT_ACTIVE",
    "SOFT_SLEPT", "SOFT_WOKE", "REFUSED_CONTACTS", "STEP", "IMPACTS", "REFUSED_IMPACTS",
    "SOFT_EVENTS", "REFUSED_SOFT_EVENTS",
]
COUNTER = {name: slot for slot, name in enumerate(COUNTER_NAMES)}
COUNTER_DEVICE_COUNT = len(COUNTER_NAMES)
COUNTER_STRIDE_WORDS = 64
COUNTER_BLOCK_WORDS = COUNTER_DEVICE_COUNT * COUNTER_STRIDE_WORDS


def counter(block: np.ndarray, name: str) -> int:
    """Value of a named counter in a downloaded counter block."""
    return int(block[COUNTER[name] * COUNTER_STRIDE_WORDS])


def aabbs(mins: np.ndarray, maxs: np.ndarray) -> np.ndarray:
    out = np.zeros(len(mins), dtype=AABB)
    out["min"] = mins
    out["max"] = maxs
    return out
