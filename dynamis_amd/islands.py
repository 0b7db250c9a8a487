"""Host side of the island stage (dyn_islands in include/dynamis.h).

Mirrors the contact part of the reference's island pass (crates/dynamis-rigid/src/islands.rs:
237-258: island_init, island_link_contacts, the island_flatten rounds, island_count): compacted
contacts and the collider owners in, each simulated row's island root (the island's minimum row)
and the island count out.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import abi
from ._native import call, lib
from .device import DeviceBuffer, Stream


class IslandsIO(ctypes.Structure):
    _fields_ = [
        ("d_contacts", ctypes.c_void_p),
        ("d_contact_count", ctypes.c_void_p),
        ("contact_capacity", ctypes.c_uint32),
        ("colliders", ctypes.c_uint32),
        ("d_owners", ctypes.c_void_p),
        ("bodies", ctypes.c_uint32),
        ("dynamic_count", ctypes.c_uint32),
        ("slop", ctypes.c_float),
        ("d_parents", ctypes.c_void_p),
        ("d_counters", ctypes.c_void_p),
        ("d_scratch", ctypes.c_void_p),
        ("scratch_bytes", ctypes.c_size_t),
    ]


class Islands:
    """`run(contacts)` enqueues the pass over a CONTACT array; `read()` returns (parents, islands,
    island_faults)."""

    def __init__(self, owners: np.ndarray, bodies: int, dynamic_count: int, slop: float,
                 stream: Stream | None = None):
        self.owners_host = np.ascontiguousarray(owners, dtype=np.uint32)
        self.bodies, self.dynamic_count, self.slop = int(bodies), int(dynamic_count), float(slop)
        self.stream = stream
        self.owners = DeviceBuffer.from_array(self.owners_host) if len(self.owners_host) else DeviceBuffer(4)
        self.parents = DeviceBuffer(max(self.dynamic_count, 1) * 4)
        self.counters = DeviceBuffer.zeros(abi.COUNTER_BLOCK_WORDS * 4)
        self.scratch = DeviceBuffer(lib().dyn_islands_scratch_bytes(self.bodies))
        self.count = DeviceBuffer(4)
        self.contacts = None

    def run(self, contacts: np.ndarray) -> None:
        contacts = np.ascontiguousarray(contacts, dtype=abi.CONTACT)
        self.contacts = DeviceBuffer.from_array(contacts) if len(contacts) else DeviceBuffer(abi.CONTACT.itemsize)
        self.count.upload(np.array([len(contacts)], np.uint32))
        io = IslandsIO(self.contacts.ptr, self.count.ptr, len(contacts), len(self.owners_host), self.owners.ptr,
                       self.bodies, self.dynamic_count, self.slop, self.parents.ptr, self.counters.ptr,
                       self.scratch.ptr, self.scratch.nbytes)
        call("dyn_islands", ctypes.c_void_p(self.stream.ptr if self.stream else None), ctypes.byref(io))

    def read(self) -> tuple[np.ndarray, int, int]:
        if self.stream:
            self.stream.synchronize()
        else:
            call("dyn_stream_synchronize", None)
        block = self.counters.download(np.uint32, abi.COUNTER_BLOCK_WORDS)
        return (self.parents.download(np.uint32, self.dynamic_count), abi.counter(block, "ISLANDS"),
                abi.counter(block, "ISLAND_FAULTS"))
