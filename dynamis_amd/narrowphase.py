"""Host side of the narrowphase stage (dyn_narrowphase in include/dynamis.h).

Mirrors the reference's narrowphase pass (crates/dynamis-rigid/src/narrowphase.rs:110-134, the
`work` program of crates/dynamis-rigid/shaders/narrowphase.wgsl:518-685 dispatched at :126):
body and collider records and candidate pairs in, one raw contact slot and one valid word per
pair out. The caller compacts, as the reference's compact_* programs do after the pass.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import abi
from ._native import call, lib
from .device import DeviceBuffer, Stream


SUPPORT_CELLS = 6 * 8 * 8  # cube-map cells per hull source (support_cells.h)


class NarrowphaseIO(ctypes.Structure):
    _fields_ = [
        ("params", ctypes.c_void_p),
        ("d_states", ctypes.c_void_p),
        ("d_descs", ctypes.c_void_p),
        ("bodies", ctypes.c_uint32),
        ("colliders", ctypes.c_uint32),
        ("d_colliders", ctypes.c_void_p),
        ("d_owners", ctypes.c_void_p),
        ("d_sources", ctypes.c_void_p),
        ("d_vertices", ctypes.c_void_p),
        ("source_count", ctypes.c_uint32),
        ("pair_capacity", ctypes.c_uint32),
        ("d_pair_major", ctypes.c_void_p),
        ("d_pair_minor", ctypes.c_void_p),
        ("d_counters", ctypes.c_void_p),
        ("d_contacts", ctypes.c_void_p),
        ("d_valid", ctypes.c_void_p),
        ("d_scratch", ctypes.c_void_p),
        ("scratch_bytes", ctypes.c_size_t),
        ("d_support_cells", ctypes.c_void_p),
        ("d_support_candidates", ctypes.c_void_p),
    ]


def support_cells(sources: np.ndarray, vertices: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """The hull sources' support cells (dyn_hull_support_cells): 2 words per source and cell,
    and the candidate vertex indices."""
    cells = np.zeros((len(sources), SUPPORT_CELLS, 2), np.uint32)
    total = ctypes.c_uint32()
    call("dyn_hull_support_cells", sources.ctypes.data, len(sources), vertices.ctypes.data, len(vertices),
         cells.ctypes.data, None, 0, ctypes.byref(total))
    cands = np.zeros(max(total.value, 1), np.uint16)
    call("dyn_hull_support_cells", sources.ctypes.data, len(sources), vertices.ctypes.data, len(vertices),
         cells.ctypes.data, cands.ctypes.data, total.value, ctypes.byref(total))
    return cells, cands


def set_epa_small_vertices(vertices: int) -> None:
    """Test hook: the vertex capacity (4-16, default 16) of the EPA kernel that gives eight lanes
    to a polytope; larger polytopes are expanded again by the wave-per-polytope kernel. Results
    do not depend on it."""
    call("dyn_set_epa_small_vertices", int(vertices))


class Narrowphase:
    """Device copies of a scene's records; `run(pairs)` enqueues the pass over the (P, 2)
    collider pairs and `read()` returns (contacts_raw, contact_valid) for them."""

    def __init__(self, scene, pair_capacity: int, stream: Stream | None = None, support_tables: bool = True):
        self.scene = scene
        self.pair_capacity = int(pair_capacity)
        self.stream = stream
        self.params = np.ascontiguousarray(scene.params, dtype=abi.STEP_PARAMS).copy()
        self.states = DeviceBuffer.from_array(np.ascontiguousarray(scene.states, dtype=abi.BODY_STATE))
        self.descs = DeviceBuffer.from_array(np.ascontiguousarray(scene.descs, dtype=abi.BODY_DESC))
        self.colliders = DeviceBuffer.from_array(np.ascontiguousarray(scene.colliders, dtype=abi.COLLIDER))
        self.owners = DeviceBuffer.from_array(np.ascontiguousarray(scene.owners, dtype=np.uint32))
        sources = scene.sources if scene.sources is not None else np.zeros(0, abi.SHAPE_SOURCE)
        self.source_count = len(sources)
        self.sources = self.vertices = self.cells = self.cands = None
        if self.source_count:
            src = np.ascontiguousarray(sources, dtype=abi.SHAPE_SOURCE)
            verts = np.ascontiguousarray(scene.vertices, dtype=np.float32)
            self.sources = DeviceBuffer.from_array(src)
            self.vertices = DeviceBuffer.from_array(verts)
            if support_tables:
                cells, cands = support_cells(src, verts)
                self.cells = DeviceBuffer.from_array(cells)
                self.cands = DeviceBuffer.from_array(cands)
        cap = max(self.pair_capacity, 1)
        self.pair_major = DeviceBuffer(cap * 4)
        self.pair_minor = DeviceBuffer(cap * 4)
        self.counters = DeviceBuffer.zeros(abi.COUNTER_BLOCK_WORDS * 4)
        self.contacts = DeviceBuffer(cap * abi.CONTACT.itemsize)
        self.valid = DeviceBuffer(cap * 4)
        self.scratch = DeviceBuffer(lib().dyn_narrowphase_scratch_bytes(self.pair_capacity))
        self.count = 0

    def set_states(self, states: np.ndarray) -> None:
        self.states.upload(np.ascontiguousarray(states, dtype=abi.BODY_STATE))

    def run(self, pairs: np.ndarray) -> None:
        pairs = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
        if len(pairs) > self.pair_capacity:
            raise ValueError("more pairs than the pair capacity")
        self.count = len(pairs)
        if self.count:
            self.pair_major.upload(np.ascontiguousarray(pairs[:, 0]))
            self.pair_minor.upload(np.ascontiguousarray(pairs[:, 1]))
        block = np.zeros(abi.COUNTER_BLOCK_WORDS, np.uint32)
        block[abi.COUNTER["PAIRS"] * abi.COUNTER_STRIDE_WORDS] = self.count
        self.counters.upload(block)
        io = NarrowphaseIO(
            self.params.ctypes.data, self.states.ptr, self.descs.ptr, len(self.scene.states),
            len(self.scene.colliders), self.colliders.ptr, self.owners.ptr,
            self.sources.ptr if self.sources else None, self.vertices.ptr if self.vertices else None,
            self.source_count, self.pair_capacity, self.pair_major.ptr, self.pair_minor.ptr,
            self.counters.ptr, self.contacts.ptr, self.valid.ptr, self.scratch.ptr, self.scratch.nbytes,
            self.cells.ptr if self.cells else None, self.cands.ptr if self.cands else None)
        call("dyn_narrowphase", ctypes.c_void_p(self.stream.ptr if self.stream else None), ctypes.byref(io))

    def read(self) -> tuple[np.ndarray, np.ndarray]:
        if self.stream:
            self.stream.synchronize()
        else:
            call("dyn_stream_synchronize", None)
        return (self.contacts.download(abi.CONTACT, self.count), self.valid.download(np.uint32, self.count))
