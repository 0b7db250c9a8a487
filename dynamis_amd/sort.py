"""Host mirror of the reference's radix sort (crates/dynamis-sort/src/radix.rs).

`RadixSort.sort(channels, plan, major_words, minor_words)` keeps the reference's call shape
(radix.rs:397) and its `SortChannels` contract (radix.rs:76-85): the count (and optional base)
live in device words, the primaries hold the result. The work runs in libdynamis_amd.so.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from ._native import call, lib
from .device import DeviceBuffer, Stream


def units_for(live: int) -> int:
    """radix.rs:63-67."""
    return int(lib().dyn_sort_units_for(int(live)))


def key_words(elements: int) -> int:
    """radix.rs:71-74."""
    return int(lib().dyn_sort_key_words(int(elements) & 0xFFFFFFFF))


@dataclass
class SortChannels:
    """Device lanes of one sort (radix.rs:76-85). `capacity` is the lane length in words."""

    count: DeviceBuffer
    major: DeviceBuffer
    minor: DeviceBuffer
    payload: DeviceBuffer
    capacity: int
    base: DeviceBuffer | None = None

    @classmethod
    def upload(cls, major, minor, payload, slots: int | None = None) -> "SortChannels":
        major = np.asarray(major, dtype=np.uint32)
        minor = np.asarray(minor, dtype=np.uint32)
        payload = np.asarray(payload, dtype=np.uint32)
        n = major.size
        slots = max(int(slots if slots is not None else n), 1)
        lanes = []
        for data in (major, minor, payload):
            buf = DeviceBuffer(slots * 4)
            if n:
                buf.upload(data)
            lanes.append(buf)
        count = DeviceBuffer.from_array(np.array([n], dtype=np.uint32))
        return cls(count, lanes[0], lanes[1], lanes[2], slots)

    def download(self, n: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        return (self.major.download(np.uint32, n), self.minor.download(np.uint32, n),
                self.payload.download(np.uint32, n))


class RadixSort:
    """Owns the scratch of one sort call site (the reference's RadixSort owns its counts,
    offsets and block totals, radix.rs:337-395)."""

    def __init__(self, stream: Stream | None = None):
        self.stream = stream
        self._scratch: DeviceBuffer | None = None

    def scratch_for(self, capacity: int, passes: int) -> DeviceBuffer:
        need = int(lib().dyn_sort_scratch_bytes(capacity, passes))
        if self._scratch is None or self._scratch.nbytes < need:
            self._scratch = DeviceBuffer(need)
        return self._scratch

    def sort(self, channels: SortChannels, plan: int, major_words: int,
             minor_words: int) -> None:
        passes = major_words + minor_words
        scratch = self.scratch_for(channels.capacity, max(passes, 1))
        call("dyn_sort_pairs",
             ctypes.c_void_p(self.stream.ptr if self.stream else None),
             channels.major.handle, channels.minor.handle, channels.payload.handle,
             channels.capacity, channels.count.handle,
             channels.base.handle if channels.base else None,
             int(plan), int(major_words), int(minor_words), scratch.handle, scratch.nbytes)
        if self.stream:
            self.stream.synchronize()
