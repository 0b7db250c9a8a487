"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so) for tests and the CPU baseline.

The oracle is the checker: tests compare the HIP library against it. It is built from
oracle/*.c with the Makefile there when the shared object is missing.
"""

from __future__ import annotations

import ctypes
import pathlib
import subprocess

import numpy as np

from dynamis_amd import abi
from dynamis_amd.world import BodyCommands

ROOT = pathlib.Path(__file__).resolve().parents[1]
ORACLE_DIR = ROOT / "oracle"
ORACLE_SO = ORACLE_DIR / "_build" / "liboracle.so"
ORACLE_OMP_SO = ORACLE_DIR / "_build" / "liboracle_omp.so"  # OpenMP build: same results, all cores

_u32 = ctypes.c_uint32
_p = ctypes.c_void_p


def build() -> pathlib.Path:
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)
    return ORACLE_SO


class Oracle:
    def __init__(self, handle: ctypes.CDLL):
        self.h = handle
        h = handle
        h.oracle_threads.restype = _u32
        h.oracle_threads.argtypes = []
        h.oracle_units_for.restype = _u32
        h.oracle_units_for.argtypes = [_u32]
        h.oracle_key_words.restype = _u32
        h.oracle_key_words.argtypes = [_u32]
        h.oracle_sort_pairs.restype = ctypes.c_int
        h.oracle_sort_pairs.argtypes = [_p, _p, _p, _u32, _u32, _u32, _u32, _u32]
        h.oracle_cell_hash.restype = _u32
        h.oracle_cell_hash.argtypes = [ctypes.c_int32] * 3
        h.oracle_cell_key.restype = _u32
        h.oracle_cell_key.argtypes = [_u32] + [ctypes.c_int32] * 3
        h.oracle_grid_cell_size.restype = ctypes.c_float
        h.oracle_grid_cell_size.argtypes = [_u32, _u32]
        h.oracle_shape_levels.restype = _u32
        h.oracle_shape_levels.argtypes = [_p, ctypes.c_float]
        h.oracle_broadphase.restype = _u32
        h.oracle_broadphase.argtypes = [_p, _p, _p, _p, _u32, _p, _p, _p, _u32, _p, _p, _u32,
                                        _p, _u32]
        h.oracle_world_create.restype = _p
        h.oracle_world_create.argtypes = [_p, _p, _p, _u32, _u32, _p, _p, _u32, _u32, _u32]
        h.oracle_world_destroy.restype = None
        h.oracle_world_destroy.argtypes = [_p]
        h.oracle_world_step.restype = None
        h.oracle_world_step.argtypes = [_p, ctypes.c_float]
        h.oracle_world_capacity.restype = None
        h.oracle_world_capacity.argtypes = [_p, _p]
        h.oracle_world_states.restype = None
        h.oracle_world_states.argtypes = [_p, _p]
        h.oracle_world_contacts.restype = _u32
        h.oracle_world_contacts.argtypes = [_p, _p, _u32]
        h.oracle_world_edit.restype = ctypes.c_int
        h.oracle_world_edit.argtypes = [_p, _p, _p, _u32]
        h.oracle_world_resting.restype = _u32
        h.oracle_world_resting.argtypes = [_p, _p, _u32]
        h.oracle_world_counters.restype = None
        h.oracle_world_counters.argtypes = [_p, _p]
        h.oracle_world_pairs.restype = _u32
        h.oracle_world_pairs.argtypes = [_p, _p, _p, _u32]
        h.oracle_world_narrowphase.restype = ctypes.c_int
        h.oracle_world_narrowphase.argtypes = [_p, _p, _p, _u32, _p, _p]
        h.oracle_world_solve.restype = ctypes.c_int
        h.oracle_world_solve.argtypes = [_p, _p, _u32, ctypes.c_float]
        h.oracle_world_tape.restype = _u32
        h.oracle_world_tape.argtypes = [_p, _p, _p, _p, _u32]
        h.oracle_world_shapes.restype = ctypes.c_int
        h.oracle_world_shapes.argtypes = [_p, _p, _u32, _p, _u32]
        h.oracle_convex_query.restype = ctypes.c_int
        h.oracle_convex_query.argtypes = [_p, _u32, _p, _u32, _p, _p, _p, _p, _p]
        h.oracle_inertia_inverse.restype = None
        h.oracle_inertia_inverse.argtypes = [_p, _p]
        h.oracle_relay_impulses.restype = None
        h.oracle_relay_impulses.argtypes = [_p, _p, _p, _p]

    def threads(self) -> int:
        return int(self.h.oracle_threads())

    def units_for(self, live: int) -> int:
        return int(self.h.oracle_units_for(live))

    def key_words(self, elements: int) -> int:
        return int(self.h.oracle_key_words(elements & 0xFFFFFFFF))

    def sort_pairs(self, major, minor, payload, major_words, minor_words, count=None, base=0,
                   capacity=None):
        major = np.array(major, dtype=np.uint32)
        minor = np.array(minor, dtype=np.uint32)
        payload = np.array(payload, dtype=np.uint32)
        cap = len(major) if capacity is None else capacity
        cnt = len(major) if count is None else count
        rc = self.h.oracle_sort_pairs(major.ctypes.data, minor.ctypes.data, payload.ctypes.data,
                                      cap, cnt, base, major_words, minor_words)
        if rc != 0:
            raise ValueError("oracle_sort_pairs refused the request")
        return major, minor, payload

    def cell_hash(self, x, y, z) -> int:
        return int(self.h.oracle_cell_hash(x, y, z))

    def cell_key(self, level, x, y, z) -> int:
        return int(self.h.oracle_cell_key(level, x, y, z))

    def broadphase(self, boxes, groups=None, flags=None, resolution=None, entry_capacity=None,
                   pair_capacity=None, sort_pairs=True):
        boxes = np.ascontiguousarray(boxes, dtype=abi.AABB)
        k = len(boxes)
        ecap = 8 * max(k, 1) if entry_capacity is None else entry_capacity
        pcap = 8 * max(k, 1) if pair_capacity is None else pair_capacity
        keys = np.zeros(max(ecap, 1), np.uint32)
        order = np.zeros(max(ecap, 1), np.uint32)
        entries = np.zeros(max(ecap, 1), abi.GRID_ENTRY)
        major = np.zeros(max(pcap, 1), np.uint32)
        minor = np.zeros(max(pcap, 1), np.uint32)
        counters = np.zeros(abi.COUNTER_BLOCK_WORDS, np.uint32)
        g = None if groups is None else np.ascontiguousarray(groups, dtype=np.uint32)
        f = None if flags is None else np.ascontiguousarray(flags, dtype=np.uint32)
        r = None if resolution is None else np.ascontiguousarray(resolution, dtype=abi.AABB)
        live = self.h.oracle_broadphase(
            boxes.ctypes.data, None if g is None else g.ctypes.data,
            None if f is None else f.ctypes.data, None if r is None else r.ctypes.data, k,
            keys.ctypes.data, order.ctypes.data, entries.ctypes.data, ecap, major.ctypes.data,
            minor.ctypes.data, pcap, counters.ctypes.data, 1 if sort_pairs else 0)
        pairs = min(abi.counter(counters, "PAIRS"), pcap)
        return {
            "entry_keys": keys[:live].copy(),
            "entry_order": order[:live].copy(),
            "entries": entries[:min(abi.counter(counters, "ENTRIES"), ecap)].copy(),
            "pairs": np.stack([major[:pairs], minor[:pairs]], axis=1),
            "counters": counters,
        }

    def inertia_inverse(self, inertia):
        tensor = np.ascontiguousarray(inertia, dtype=np.float32)
        out = np.zeros(6, np.float32)
        self.h.oracle_inertia_inverse(tensor.ctypes.data, out.ctypes.data)
        return out

    def convex_query(self, scene, row_a, row_b):
        """GJK closest points and convex_hit of the colliders of two rows of a scene (their
        body states as encoded): a dict of distance, point_a, point_b, normal, penetrating,
        hit_distance, hit_normal."""
        out = np.zeros(15, np.float32)
        ca = scene.colliders[[int(np.nonzero(scene.owners == row_a)[0][0])]]
        cb = scene.colliders[[int(np.nonzero(scene.owners == row_b)[0][0])]]
        sa = scene.states[[row_a]]
        sb = scene.states[[row_b]]
        src = scene.sources if scene.sources is not None and len(scene.sources) else np.zeros(1, abi.SHAPE_SOURCE)
        verts = scene.vertices if scene.vertices is not None and len(scene.vertices) else np.zeros((1, 4), np.float32)
        self.h.oracle_convex_query(src.ctypes.data, len(src), verts.ctypes.data, len(verts),
                                   ca.ctypes.data, sa.ctypes.data, cb.ctypes.data, sb.ctypes.data,
                                   out.ctypes.data)
        return {"distance": float(out[0]), "point_a": out[1:4].copy(), "point_b": out[4:7].copy(),
                "normal": out[7:10].copy(), "penetrating": bool(out[10]),
                "hit_distance": float(out[11]), "hit_normal": out[12:15].copy()}

    def relay(self, current, held):
        """identity.wgsl's relay alone: (current with held's impulses relayed, the held point
        relayed into each point or -1)."""
        cur = np.ascontiguousarray(current, dtype=abi.CONTACT).reshape(1)
        hld = np.ascontiguousarray(held, dtype=abi.CONTACT).reshape(1)
        out = np.zeros(1, abi.CONTACT)
        picks = np.zeros(4, np.uint32)
        self.h.oracle_relay_impulses(cur.ctypes.data, hld.ctypes.data, out.ctypes.data, picks.ctypes.data)
        return out[0], np.where(picks == 0xFFFFFFFF, -1, picks.astype(np.int64))

    def world(self, scene, contact_capacity=None, pair_capacity=None) -> "OracleWorld":
        return OracleWorld(self, scene, contact_capacity, pair_capacity)


class OracleWorld(BodyCommands):
    """The oracle's serial restatement of World::step over the same device records."""

    def __init__(self, oracle: Oracle, scene, contact_capacity=None, pair_capacity=None):
        from dynamis_amd.world import default_capacities
        self.o = oracle
        n = len(scene.states)
        self.n = n
        capacity, pairs = default_capacities(n, contact_capacity, pair_capacity)
        self.scene = scene
        self.w = oracle.h.oracle_world_create(
            scene.params.ctypes.data, scene.states.ctypes.data, scene.descs.ctypes.data, n,
            scene.dynamic_count, scene.colliders.ctypes.data, scene.owners.ctypes.data, n,
            capacity, pairs)
        if not self.w:
            raise MemoryError("oracle_world_create failed")
        if scene.sources is not None and len(scene.sources):
            rc = oracle.h.oracle_world_shapes(self.w, scene.sources.ctypes.data, len(scene.sources),
                                              scene.vertices.ctypes.data, len(scene.vertices))
            if rc != 0:
                raise ValueError("oracle_world_shapes refused the shape sources")

    def step(self, dt: float) -> None:
        self.o.h.oracle_world_step(self.w, ctypes.c_float(dt))

    def capacities(self) -> tuple[int, int, int]:
        """(pair capacity, contact capacity, growths so far), as World.capacities."""
        out = np.zeros(3, np.uint32)
        self.o.h.oracle_world_capacity(self.w, out.ctypes.data)
        return int(out[0]), int(out[1]), int(out[2])

    @property
    def capacity(self) -> int:
        return self.capacities()[1]

    @property
    def pair_capacity(self) -> int:
        return self.capacities()[0]

    def states(self) -> np.ndarray:
        out = np.zeros(self.n, dtype=abi.BODY_STATE)
        self.o.h.oracle_world_states(self.w, out.ctypes.data)
        return out

    def measured(self) -> dict:
        block = np.zeros(abi.COUNTER_BLOCK_WORDS, dtype=np.uint32)
        self.o.h.oracle_world_counters(self.w, block.ctypes.data)
        return {name: abi.counter(block, name) for name in abi.COUNTER_NAMES}

    def inspect_contacts(self) -> np.ndarray:
        out = np.zeros(self.capacity, dtype=abi.CONTACT)
        count = self.o.h.oracle_world_contacts(self.w, out.ctypes.data, self.capacity)
        return out[:min(count, self.capacity)].copy()

    def _queue_edits(self, ids: np.ndarray, edits: np.ndarray) -> None:
        if self.o.h.oracle_world_edit(self.w, ids.ctypes.data, edits.ctypes.data, len(ids)) != 0:
            raise ValueError("oracle_world_edit refused the edits")

    def resting_contacts(self) -> np.ndarray:
        out = np.zeros(self.capacity, dtype=abi.CONTACT)
        count = self.o.h.oracle_world_resting(self.w, out.ctypes.data, self.capacity)
        return out[:min(count, self.capacity)].copy()

    def pairs(self) -> np.ndarray:
        major = np.zeros(self.pair_capacity, np.uint32)
        minor = np.zeros(self.pair_capacity, np.uint32)
        count = self.o.h.oracle_world_pairs(self.w, major.ctypes.data, minor.ctypes.data,
                                            self.pair_capacity)
        count = min(count, self.pair_capacity)
        return np.stack([major[:count], minor[:count]], axis=1)

    def narrowphase(self, pairs: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        """The narrowphase pass alone over (P, 2) collider pairs and the current records:
        (contacts_raw, contact_valid), one per pair."""
        pairs = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
        major, minor = np.ascontiguousarray(pairs[:, 0]), np.ascontiguousarray(pairs[:, 1])
        raw = np.zeros(len(pairs), dtype=abi.CONTACT)
        valid = np.zeros(len(pairs), np.uint32)
        rc = self.o.h.oracle_world_narrowphase(self.w, major.ctypes.data, minor.ctypes.data, len(pairs),
                                               raw.ctypes.data, valid.ctypes.data)
        if rc != 0:
            raise ValueError("oracle_world_narrowphase refused the pairs")
        return raw, valid

    def solve(self, contacts: np.ndarray, dt: float) -> np.ndarray:
        """The solver alone over the contacts and the current records; returns the contacts with
        their new accumulated impulses (the states move in place: read them with states())."""
        contacts = np.ascontiguousarray(contacts, dtype=abi.CONTACT).copy()
        rc = self.o.h.oracle_world_solve(self.w, contacts.ctypes.data, len(contacts), ctypes.c_float(dt))
        if rc != 0:
            raise ValueError("oracle_world_solve refused the contacts")
        return contacts

    def tape(self) -> tuple[np.ndarray, np.ndarray]:
        """The last step's states after its input clearing and its contacts after the warm-start
        relay and contact_begin, as the solver received them."""
        start = np.zeros(self.n, abi.BODY_STATE)
        pre = np.zeros(self.capacity, abi.CONTACT)
        src = np.zeros((self.capacity, 4), np.uint32)
        count = self.o.h.oracle_world_tape(self.w, start.ctypes.data, pre.ctypes.data, src.ctypes.data,
                                           self.capacity)
        return start, pre[:min(count, self.capacity)].copy()

    def close(self) -> None:
        if self.w:
            self.o.h.oracle_world_destroy(self.w)
            self.w = None

    def __del__(self):
        self.close()


_oracles: dict[bool, Oracle] = {}


def load(parallel: bool = False) -> Oracle:
    """The serial oracle, or with `parallel` its OpenMP build (the all-core CPU baseline)."""
    if parallel not in _oracles:
        path = ORACLE_OMP_SO if parallel else ORACLE_SO
        if not path.exists():
            build()
        _oracles[parallel] = Oracle(ctypes.CDLL(str(path)))
    return _oracles[parallel]
