"""Host side of the solver stage (dyn_solve in include/dynamis.h).

Mirrors the reference's solver passes (crates/dynamis-rigid/src/solver.rs:89-101 SolverReset and
:304-334 SolveSubsteps): body records and compacted, warm-started contacts in; the bodies after
the substeps and the contacts with their new accumulated impulses out.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import abi
from ._native import call, lib
from .device import DeviceBuffer, Stream


class SolveIO(ctypes.Structure):
    _fields_ = [
        ("params", ctypes.c_void_p),
        ("d_states", ctypes.c_void_p),
        ("d_descs", ctypes.c_void_p),
        ("bodies", ctypes.c_uint32),
        ("dynamic_count", ctypes.c_uint32),
        ("colliders", ctypes.c_uint32),
        ("contact_capacity", ctypes.c_uint32),
        ("d_owners", ctypes.c_void_p),
        ("d_contacts", ctypes.c_void_p),
        ("d_contact_count", ctypes.c_void_p),
        ("d_counters", ctypes.c_void_p),
        ("d_scratch", ctypes.c_void_p),
        ("scratch_bytes", ctypes.c_size_t),
    ]


class Solver:
    """Device copies of a scene's descriptors and owners; `run(states, contacts, dt)` enqueues the
    solver on those states and contacts and `read()` returns (states, contacts, counters)."""

    def __init__(self, scene, contact_capacity: int, stream: Stream | None = None):
        self.scene = scene
        self.capacity = int(contact_capacity)
        self.stream = stream
        self.params = np.ascontiguousarray(scene.params, dtype=abi.STEP_PARAMS).copy()
        self.bodies = len(scene.states)
        self.states = DeviceBuffer(max(self.bodies, 1) * abi.BODY_STATE.itemsize)
        self.descs = DeviceBuffer.from_array(np.ascontiguousarray(scene.descs, dtype=abi.BODY_DESC))
        self.owners = DeviceBuffer.from_array(np.ascontiguousarray(scene.owners, dtype=np.uint32))
        self.contacts = DeviceBuffer(max(self.capacity, 1) * abi.CONTACT.itemsize)
        self.count_word = DeviceBuffer(4)
        self.counters = DeviceBuffer.zeros(abi.COUNTER_BLOCK_WORDS * 4)
        self.scratch = DeviceBuffer(lib().dyn_solve_scratch_bytes(self.bodies, self.capacity))
        self.count = 0

    def run(self, states: np.ndarray, contacts: np.ndarray, dt: float) -> None:
        contacts = np.ascontiguousarray(contacts, dtype=abi.CONTACT)
        if len(contacts) > self.capacity:
            raise ValueError("more contacts than the contact capacity")
        self.states.upload(np.ascontiguousarray(states, dtype=abi.BODY_STATE))
        if len(contacts):
            self.contacts.upload(contacts)
        self.count = len(contacts)
        self.count_word.upload(np.array([self.count], np.uint32))
        self.again(dt)

    def again(self, dt: float) -> None:
        """Enqueues the solver once more on the device's states and contacts as the last run left
        them (no upload)."""
        io = SolveIO(self.params.ctypes.data, self.states.ptr, self.descs.ptr, self.bodies,
                     int(self.scene.dynamic_count), len(self.scene.colliders), self.capacity, self.owners.ptr,
                     self.contacts.ptr, self.count_word.ptr, self.counters.ptr, self.scratch.ptr,
                     self.scratch.nbytes)
        call("dyn_solve", ctypes.c_void_p(self.stream.ptr if self.stream else None), ctypes.byref(io),
             ctypes.c_float(dt))

    def read(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        if self.stream:
            self.stream.synchronize()
        else:
            call("dyn_stream_synchronize", None)
        return (self.states.download(abi.BODY_STATE, self.bodies), self.contacts.download(abi.CONTACT, self.count),
                self.counters.download(np.uint32, abi.COUNTER_BLOCK_WORDS))
