"""Host side of the broadphase stage (dyn_broadphase in include/dynamis.h).

Mirrors the reference's `Broadphase::record` (crates/dynamis-broadphase/src/passes.rs:145-160)
together with the entry emission that feeds it (crates/dynamis-rigid/src/emit_entries.rs:127-144)
for one awake region: boxes in, sorted grid entries and candidate pairs out.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import abi
from ._native import call, lib
from .device import DeviceBuffer, Stream


class BroadphaseIO(ctypes.Structure):
    _fields_ = [
        ("d_aabbs", ctypes.c_void_p),
        ("d_groups", ctypes.c_void_p),
        ("d_flags", ctypes.c_void_p),
        ("d_resolution", ctypes.c_void_p),
        ("colliders", ctypes.c_uint32),
        ("d_entry_keys", ctypes.c_void_p),
        ("d_entry_order", ctypes.c_void_p),
        ("d_entries", ctypes.c_void_p),
        ("entry_capacity", ctypes.c_uint32),
        ("d_pair_major", ctypes.c_void_p),
        ("d_pair_minor", ctypes.c_void_p),
        ("pair_capacity", ctypes.c_uint32),
        ("d_counters", ctypes.c_void_p),
        ("sort_pairs", ctypes.c_uint32),
        ("_reserved", ctypes.c_uint32),
        ("d_scratch", ctypes.c_void_p),
        ("scratch_bytes", ctypes.c_size_t),
        ("d_worlds", ctypes.c_void_p),
    ]


@dataclass
class BroadphaseResult:
    entry_keys: np.ndarray      # sorted cell keys
    entry_order: np.ndarray     # entry slot of each sorted key
    entries: np.ndarray         # GRID_ENTRY records by slot
    pairs: np.ndarray           # (P, 2) uint32 (min, max) collider pairs
    counters: np.ndarray        # counter block words


class Broadphase:
    """Device buffers for K colliders; `run()` enqueues one pass, `read()` copies results out."""

    def __init__(self, colliders: int, entry_capacity: int | None = None,
                 pair_capacity: int | None = None, stream: Stream | None = None):
        self.colliders = int(colliders)
        self.entry_capacity = int(entry_capacity if entry_capacity is not None
                                  else abi.MAX_CELLS_PER_COLLIDER * max(self.colliders, 1))
        self.pair_capacity = int(pair_capacity if pair_capacity is not None
                                 else 8 * max(self.colliders, 1))
        self.stream = stream
        k = max(self.colliders, 1)
        self.aabbs = DeviceBuffer(k * abi.AABB.itemsize)
        self.groups = DeviceBuffer(k * 4)
        self.flags = DeviceBuffer(k * 4)
        self.resolution = DeviceBuffer(k * abi.AABB.itemsize)
        self.entry_keys = DeviceBuffer(max(self.entry_capacity, 1) * 4)
        self.entry_order = DeviceBuffer(max(self.entry_capacity, 1) * 4)
        self.entries = DeviceBuffer(max(self.entry_capacity, 1) * abi.GRID_ENTRY.itemsize)
        self.pair_major = DeviceBuffer(max(self.pair_capacity, 1) * 4)
        self.pair_minor = DeviceBuffer(max(self.pair_capacity, 1) * 4)
        self.counters = DeviceBuffer.zeros(abi.COUNTER_BLOCK_WORDS * 4)
        need = lib().dyn_broadphase_scratch_bytes(self.colliders, self.entry_capacity,
                                                  self.pair_capacity)
        self.scratch = DeviceBuffer(need)
        self._has_groups = self._has_flags = self._has_resolution = False

    def upload(self, boxes: np.ndarray, groups: np.ndarray | None = None,
               flags: np.ndarray | None = None, resolution: np.ndarray | None = None) -> None:
        boxes = np.ascontiguousarray(boxes, dtype=abi.AABB)
        if len(boxes) != self.colliders:
            raise ValueError("box count differs from the collider count")
        if self.colliders:
            self.aabbs.upload(boxes)
        self._has_groups = groups is not None
        if groups is not None:
            self.groups.upload(np.asarray(groups, dtype=np.uint32))
        self._has_flags = flags is not None
        if flags is not None:
            self.flags.upload(np.asarray(flags, dtype=np.uint32))
        self._has_resolution = resolution is not None
        if resolution is not None:
            self.resolution.upload(np.ascontiguousarray(resolution, dtype=abi.AABB))

    def io(self, sort_pairs: bool = True) -> BroadphaseIO:
        return BroadphaseIO(
            self.aabbs.ptr, self.groups.ptr if self._has_groups else None,
            self.flags.ptr if self._has_flags else None,
            self.resolution.ptr if self._has_resolution else None, self.colliders,
            self.entry_keys.ptr, self.entry_order.ptr, self.entries.ptr, self.entry_capacity,
            self.pair_major.ptr, self.pair_minor.ptr, self.pair_capacity, self.counters.ptr,
            1 if sort_pairs else 0, 0, self.scratch.ptr, self.scratch.nbytes)

    def run(self, sort_pairs: bool = True) -> None:
        io = self.io(sort_pairs)
        call("dyn_broadphase", ctypes.c_void_p(self.stream.ptr if self.stream else None),
             ctypes.byref(io))

    def read(self) -> BroadphaseResult:
        if self.stream:
            self.stream.synchronize()
        counters = self.counters.download(np.uint32, abi.COUNTER_BLOCK_WORDS)
        entries = min(abi.counter(counters, "ENTRIES"), self.entry_capacity)
        pairs = min(abi.counter(counters, "PAIRS"), self.pair_capacity)
        major = self.pair_major.download(np.uint32, pairs)
        minor = self.pair_minor.download(np.uint32, pairs)
        return BroadphaseResult(
            self.entry_keys.download(np.uint32, entries),
            self.entry_order.download(np.uint32, entries),
            self.entries.download(abi.GRID_ENTRY, entries),
            np.stack([major, minor], axis=1) if pairs else np.zeros((0, 2), np.uint32),
            counters)


@dataclass
class FixedGridResult:
    entry_hash: np.ndarray    # sorted entries: cell hash, box index, cell ordinal
    entry_box: np.ndarray
    entry_ordinal: np.ndarray
    pairs: np.ndarray         # (P, 2) canonical overlapping pairs, first < second
    refused_boxes: int
    refused_pairs: int


class FixedGrid:
    """The literal fixed-cell broadphase (dyn_fixed_grid): a uniform grid of side `cell_size`,
    keys (cell hash, box index), radix sort, cell runs, canonical overlapping pairs."""

    def __init__(self, boxes: int, cell_size: float = 4.0, entry_capacity: int | None = None,
                 pair_capacity: int | None = None, stream: Stream | None = None):
        self.k = int(boxes)
        self.cell_size = float(cell_size)
        self.entry_capacity = int(entry_capacity if entry_capacity is not None else 8 * max(self.k, 1))
        self.pair_capacity = int(pair_capacity if pair_capacity is not None else 4 * max(self.k, 1))
        self.stream = stream or Stream()
        self.boxes = DeviceBuffer(max(self.k, 1) * abi.AABB.itemsize)
        self.entry_hash = DeviceBuffer(max(self.entry_capacity, 1) * 4)
        self.entry_box = DeviceBuffer(max(self.entry_capacity, 1) * 4)
        self.entry_ordinal = DeviceBuffer(max(self.entry_capacity, 1) * 4)
        self.pair_first = DeviceBuffer(max(self.pair_capacity, 1) * 4)
        self.pair_second = DeviceBuffer(max(self.pair_capacity, 1) * 4)
        self.counts = DeviceBuffer.zeros(16)
        need = lib().dyn_fixed_grid_scratch_bytes(self.k, self.entry_capacity, self.pair_capacity)
        self.scratch = DeviceBuffer(need)

    def upload(self, boxes: np.ndarray) -> None:
        self.boxes.upload(np.ascontiguousarray(boxes, dtype=abi.AABB))

    def run(self) -> None:
        call("dyn_fixed_grid", self.stream.handle, self.boxes.ptr, self.k, ctypes.c_float(self.cell_size),
             self.entry_hash.ptr, self.entry_box.ptr, self.entry_ordinal.ptr, self.entry_capacity,
             self.pair_first.ptr, self.pair_second.ptr, self.pair_capacity, self.counts.ptr,
             self.scratch.ptr, self.scratch.nbytes)

    def read(self) -> FixedGridResult:
        self.stream.synchronize()
        counts = self.counts.download(np.uint32, 4)
        e = min(int(counts[0]), self.entry_capacity)
        p = min(int(counts[1]), self.pair_capacity)
        first = self.pair_first.download(np.uint32, p)
        second = self.pair_second.download(np.uint32, p)
        return FixedGridResult(self.entry_hash.download(np.uint32, e), self.entry_box.download(np.uint32, e),
                               self.entry_ordinal.download(np.uint32, e), np.stack([first, second], axis=1),
                               int(counts[2]), int(counts[3]))
