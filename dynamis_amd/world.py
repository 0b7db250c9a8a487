"""Host side of the rigid-body step (dyn_world_* in include/dynamis.h).

Mirrors the reference's `World` for the rigid path (crates/dynamis-world/src/lib.rs:86 new,
src/body.rs:108 spawn, :326 read_state, src/step.rs:34 step, src/observation.rs:484 wait,
src/readback.rs:95 inspect_contacts, :215 measured): bodies with one collider each are described
on the host, encoded into the device records by the library (dyn_encode_body), and stepped on the
GPU. Rows follow the reference's partition (crates/dynamis-world/src/body.rs:292-321): simulated
bodies first, then static and kinematic ones; a handle is the spawn index.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import abi
from ._native import call, lib


@dataclass
class HullSource:
    """A hull shape source: vertices, outward triangles and its solid geometry."""

    vertices: np.ndarray      # (V, 3) float32
    triangles: np.ndarray     # (T, 3) uint32
    volume: float
    centroid: np.ndarray      # (3,) float32
    second: np.ndarray        # (6,) float32 unit second moment about the centroid


def build_hull(points) -> HullSource:
    """Quickhull over a point set and the hull's solid geometry (World::add_hull_from_mesh,
    crates/dynamis-world/src/shape.rs:50-57; crates/dynamis-hull/src/{hull,solid}.rs)."""
    pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    cap = max(8, 2 * len(pts) + 8)
    verts = np.zeros((cap, 3), np.float32)
    tris = np.zeros((cap, 3), np.uint32)
    nv, nt = ctypes.c_uint32(), ctypes.c_uint32()
    call("dyn_hull_build", pts.ctypes.data, len(pts), verts.ctypes.data, ctypes.byref(nv),
         tris.ctypes.data, ctypes.byref(nt), cap)
    verts, tris = verts[:nv.value].copy(), tris[:nt.value].copy()
    volume = np.zeros(1, np.float32)
    centroid = np.zeros(3, np.float32)
    second = np.zeros(6, np.float32)
    call("dyn_hull_solid", verts.ctypes.data, len(verts), tris.ctypes.data, len(tris),
         volume.ctypes.data, centroid.ctypes.data, second.ctypes.data)
    return HullSource(vertices=verts, triangles=tris, volume=float(volume[0]), centroid=centroid,
                      second=second)


def shape_records(hulls: list[HullSource]) -> tuple[np.ndarray, np.ndarray]:
    """ShapeSource records and float4 vertex rows of a list of hulls (source id = index)."""
    sources = np.zeros(len(hulls), dtype=abi.SHAPE_SOURCE)
    rows = []
    offset = 0
    for i, h in enumerate(hulls):
        sources[i]["kind"] = abi.SHAPE_HULL
        sources[i]["vertex_offset"] = offset
        sources[i]["vertex_count"] = len(h.vertices)
        sources[i]["local_min"] = h.vertices.min(axis=0)
        sources[i]["local_max"] = h.vertices.max(axis=0)
        r = np.zeros((len(h.vertices), 4), np.float32)
        r[:, :3] = h.vertices
        rows.append(r)
        offset += len(h.vertices)
    vertices = np.concatenate(rows) if rows else np.zeros((0, 4), np.float32)
    return sources, np.ascontiguousarray(vertices)


def default_config() -> np.ndarray:
    """PhysicsConfig::default (crates/dynamis-world/src/config.rs:76-100) as a CONFIG record."""
    config = np.zeros(1, dtype=abi.CONFIG)
    lib().dyn_config_default(config.ctypes.data)
    return config


@dataclass
class BodyDesc:
    """One body with one collider at its origin (crates/dynamis-model/src/body.rs:89-108 and
    collider.rs:52-68). `density` >= 0 derives the mass from the shape's volume
    (BodyDesc::density); a negative density stands for None, and `mass` is then the body's
    mass (BodyDesc::new's 1.0 by default); a zero mass makes a static body."""

    shape: int
    radius: float = 0.0
    half_height: float = 0.0
    half_extents: tuple = (0.0, 0.0, 0.0)
    position: tuple = (0.0, 0.0, 0.0)
    orientation: tuple = (0.0, 0.0, 0.0, 1.0)
    velocity: tuple = (0.0, 0.0, 0.0)
    angular_velocity: tuple = (0.0, 0.0, 0.0)
    mass: float = 1.0
    density: float = -1.0
    friction: float = 0.5
    restitution: float = 0.0
    rolling_friction: float = 0.0
    spin_friction: float = 0.0
    gravity_scale: float = 1.0
    linear_damping: float = -1.0
    angular_damping: float = -1.0
    kinematic: bool = False
    collision_group: int = 1
    collision_mask: int = abi.NO_COLLISION_FILTER
    source: int = 0

    # constructors named after crates/dynamis-model/src/shape.rs
    @staticmethod
    def sphere(radius: float, **kw) -> "BodyDesc":
        return BodyDesc(shape=abi.SHAPE_SPHERE, radius=radius, **kw)

    @staticmethod
    def cuboid(half_extents, **kw) -> "BodyDesc":
        return BodyDesc(shape=abi.SHAPE_CUBOID, half_extents=tuple(half_extents), **kw)

    @staticmethod
    def capsule(radius: float, half_height: float, **kw) -> "BodyDesc":
        return BodyDesc(shape=abi.SHAPE_CAPSULE, radius=radius, half_height=half_height, **kw)

    @staticmethod
    def cylinder(radius: float, half_height: float, **kw) -> "BodyDesc":
        return BodyDesc(shape=abi.SHAPE_CYLINDER, radius=radius, half_height=half_height, **kw)

    @staticmethod
    def hull(source: int, **kw) -> "BodyDesc":
        """A body whose collider is hull shape source `source` (an index into the hulls
        passed to `encode` / `World.add_hull`)."""
        return BodyDesc(shape=abi.SHAPE_HULL, source=source, **kw)

    @staticmethod
    def plane(**kw) -> "BodyDesc":
        kw.setdefault("mass", 0.0)
        return BodyDesc(shape=abi.SHAPE_PLANE, **kw)

    def fixed(self) -> "BodyDesc":
        """A static body: no mass, never moves (BodyKind::of with zero mass)."""
        self.density = -1.0
        self.mass = 0.0
        return self

    def spec(self, hulls: list[HullSource] | None = None) -> np.ndarray:
        s = np.zeros(1, dtype=abi.BODY_SPEC)
        for name in abi.BODY_SPEC.names:
            if name.startswith("solid_"):
                continue
            value = getattr(self, name)
            s[name] = int(value) if name == "kinematic" else value
        if self.shape == abi.SHAPE_HULL:
            if hulls is None or not 0 <= self.source < len(hulls):
                raise ValueError(f"hull body names shape source {self.source}, which is not given")
            h = hulls[self.source]
            s["solid_volume"] = h.volume
            s["solid_centroid"] = h.centroid
            s["solid_second"] = h.second
        return s


@dataclass
class Scene:
    """Device records of a set of bodies, in row order, ready for dyn_world_create."""

    params: np.ndarray
    states: np.ndarray
    descs: np.ndarray
    colliders: np.ndarray
    owners: np.ndarray
    dynamic_count: int
    rows: np.ndarray                 # handle -> row
    handles: np.ndarray = field(default=None)  # row -> handle
    sources: np.ndarray = field(default=None)  # SHAPE_SOURCE records
    vertices: np.ndarray = field(default=None)  # (V, 4) float32 rows


def spawn_rows(simulates: list[bool]) -> list[int]:
    """Row of each spawned body after the reference's spawn bookkeeping for a fresh world:
    a body is appended at the end of the rows (pool.rs:118-127 with compact removal), and a
    simulated body is swapped with the first non-simulated row (body.rs:292-321
    migrate_partition). Returns row -> handle."""
    alive: list[int] = []
    dynamic = 0
    for handle, sim in enumerate(simulates):
        slot = len(alive)
        alive.append(handle)
        if sim:
            if slot != dynamic:
                alive[slot], alive[dynamic] = alive[dynamic], alive[slot]
            dynamic += 1
    return alive


def encode(bodies: list[BodyDesc], config: np.ndarray | None = None,
           hulls: list[HullSource] | None = None) -> Scene:
    """Encodes bodies spawned in this order into device records with the library's host
    encoder. Body handle ids are the spawn indices; collider slots follow spawn order (the
    collider arena, collider.rs:180-214); rows follow the simulated-first partition."""
    specs = (np.concatenate([b.spec(hulls) for b in bodies]) if bodies
             else np.zeros(0, dtype=abi.BODY_SPEC))
    return encode_specs(specs, config, hulls)


def encode_specs(specs: np.ndarray, config: np.ndarray | None = None,
                 hulls: list[HullSource] | None = None) -> Scene:
    """`encode` of an array of BODY_SPEC records (spawn order), one library call for all."""
    config = default_config() if config is None else config
    specs = np.ascontiguousarray(specs, dtype=abi.BODY_SPEC)
    n = len(specs)
    enc_states = np.zeros(n, dtype=abi.BODY_STATE)
    enc_descs = np.zeros(n, dtype=abi.BODY_DESC)
    colliders = np.zeros(n, dtype=abi.COLLIDER)
    if n:
        call("dyn_encode_bodies", specs.ctypes.data, n, config.ctypes.data, 0, enc_states.ctypes.data,
             enc_descs.ctypes.data, colliders.ctypes.data)
    simulates = (enc_descs["inverse_mass"] > 0.0) & ((enc_descs["flags"] & abi.BODY_KINEMATIC) == 0)
    order = spawn_rows(simulates.tolist())
    rows = np.zeros(n, dtype=np.uint32)
    rows[np.asarray(order, dtype=np.int64)] = np.arange(n, dtype=np.uint32)
    states = enc_states[order].copy() if n else enc_states
    descs = enc_descs[order].copy() if n else enc_descs
    dynamic_count = int(simulates.sum())
    params = np.zeros(1, dtype=abi.STEP_PARAMS)
    call("dyn_step_params_build", config.ctypes.data, n, dynamic_count, n, params.ctypes.data)
    sources, vertices = shape_records(hulls or [])
    return Scene(params=params, states=states, descs=descs, colliders=colliders,
                 owners=rows.copy(), dynamic_count=dynamic_count, rows=rows,
                 handles=np.asarray(order, dtype=np.uint32), sources=sources, vertices=vertices)


def default_capacities(bodies: int, contacts: int | None = None,
                       pairs: int | None = None) -> tuple[int, int]:
    """Per-step stream capacities: 8 contacts and 32 candidate pairs per body unless given."""
    c = int(contacts if contacts is not None else max(64, 8 * bodies))
    p = int(pairs if pairs is not None else max(256, 32 * bodies))
    return c, p


def body_edit(kind: int, mask: int = 0, position=None, orientation=None, velocity=None,
              angular_velocity=None, force=None, torque=None) -> np.ndarray:
    """One BODY_EDIT record (records.wgsl:193-199): the fields a kind reads (body_edits.wgsl)."""
    e = np.zeros(1, dtype=abi.BODY_EDIT)
    e["kind"] = kind
    e["mask"] = mask
    for name, value in (("position", position), ("orientation", orientation), ("velocity", velocity),
                        ("angular_velocity", angular_velocity), ("force", force), ("torque", torque)):
        if value is not None:
            e["state"][name] = value
    return e


class BodyCommands:
    """The body commands of the reference's World (crates/dynamis-world/src/command.rs and
    body.rs: set_velocity, set_position, apply_force, apply_impulse, sleep, wake, ...), each
    queued as a body edit for the next step. `_queue_edits(ids, edits)` queues the records;
    `_body_id(handle)` maps a handle to its body id."""

    def _body_id(self, handle: int) -> int:
        return int(handle)

    def edit(self, handle: int, record: np.ndarray) -> None:
        ids = np.array([self._body_id(handle)], dtype=np.uint32)
        self._queue_edits(ids, np.ascontiguousarray(record, dtype=abi.BODY_EDIT).reshape(1))

    def set_position(self, handle, position) -> None:
        self.edit(handle, body_edit(abi.EDIT_PATCH, abi.PATCH_POSITION, position=position))

    def set_orientation(self, handle, orientation) -> None:
        self.edit(handle, body_edit(abi.EDIT_PATCH, abi.PATCH_ORIENTATION, orientation=orientation))

    def set_velocity(self, handle, velocity) -> None:
        self.edit(handle, body_edit(abi.EDIT_PATCH, abi.PATCH_VELOCITY, velocity=velocity))

    def set_angular_velocity(self, handle, angular_velocity) -> None:
        self.edit(handle, body_edit(abi.EDIT_PATCH, abi.PATCH_ANGULAR_VELOCITY, angular_velocity=angular_velocity))

    def apply_force(self, handle, force) -> None:
        self.edit(handle, body_edit(abi.EDIT_FORCE, force=force))

    def apply_force_at_point(self, handle, force, point) -> None:
        self.edit(handle, body_edit(abi.EDIT_FORCE_AT_POINT, force=force, position=point))

    def apply_torque(self, handle, torque) -> None:
        self.edit(handle, body_edit(abi.EDIT_TORQUE, torque=torque))

    def apply_impulse(self, handle, impulse) -> None:
        self.edit(handle, body_edit(abi.EDIT_IMPULSE, velocity=impulse))

    def apply_impulse_at_point(self, handle, impulse, point) -> None:
        self.edit(handle, body_edit(abi.EDIT_IMPULSE_AT_POINT, velocity=impulse, position=point))

    def apply_angular_impulse(self, handle, impulse) -> None:
        self.edit(handle, body_edit(abi.EDIT_ANGULAR_IMPULSE, angular_velocity=impulse))

    def sleep(self, handle) -> None:
        self.edit(handle, body_edit(abi.EDIT_SLEEP))

    def wake(self, handle) -> None:
        self.edit(handle, body_edit(abi.EDIT_WAKE))


class World(BodyCommands):
    """A device world. Bodies are spawned before the first step; `step(dt)` enqueues one step
    on the world's stream, reads wait for it."""

    def __init__(self, config: np.ndarray | None = None, contact_capacity: int | None = None,
                 pair_capacity: int | None = None):
        self.config = default_config() if config is None else config.copy()
        self.contact_capacity = contact_capacity
        self._pair_request = pair_capacity
        self._bodies: list[BodyDesc] = []
        self._hulls: list[HullSource] = []
        self._handle = ctypes.c_void_p()
        self.scene: Scene | None = None

    # -- bodies ------------------------------------------------------------------------------
    def spawn(self, desc: BodyDesc) -> int:
        if self.scene is not None:
            raise RuntimeError("bodies are spawned before the first step of a world")
        self._bodies.append(desc)
        return len(self._bodies) - 1

    def count(self) -> int:
        return len(self._bodies)

    def add_hull(self, points) -> int:
        """Builds a hull shape source from points and returns its id (World::add_hull_from_mesh)."""
        if self.scene is not None:
            raise RuntimeError("shape sources are added before the first step of a world")
        self._hulls.append(build_hull(points))
        return len(self._hulls) - 1

    def _ensure(self) -> None:
        if self.scene is not None:
            return
        if not self._bodies:
            raise RuntimeError("a world needs at least one body")
        self.scene = encode(self._bodies, self.config, self._hulls)
        self.create(self.scene, self.contact_capacity, self._pair_request)

    def create(self, scene: Scene, contact_capacity: int | None = None,
               pair_capacity: int | None = None, stream: int | None = None) -> None:
        """Creates the device world from already-encoded records, on its own stream or on the
        given one (worlds sharing a stream step one after another)."""
        self.scene = scene
        n = len(scene.states)
        capacity, pairs = default_capacities(n, contact_capacity, pair_capacity)
        call("dyn_world_create", ctypes.byref(self._handle), stream, scene.params.ctypes.data,
             scene.states.ctypes.data, scene.descs.ctypes.data, n, scene.dynamic_count,
             scene.colliders.ctypes.data, scene.owners.ctypes.data, n, capacity, pairs)
        if scene.sources is not None and len(scene.sources):
            call("dyn_world_upload_shapes", self._handle, scene.sources.ctypes.data,
                 len(scene.sources), scene.vertices.ctypes.data, len(scene.vertices))

    # -- stepping ----------------------------------------------------------------------------
    def step(self, dt: float) -> None:
        self._ensure()
        call("dyn_world_step", self._handle, ctypes.c_float(dt))

    def wait(self) -> None:
        if self.scene is not None:
            call("dyn_world_wait", self._handle)

    # -- capacity planning (World::apply_plan, crates/dynamis-world/src/backend/mod.rs:194-230) --
    def capacities(self) -> tuple[int, int, int]:
        """(pair capacity, contact capacity, growths so far): a step's demand grows the streams
        two steps later (dyn_world_capacity)."""
        pairs, contacts, grows = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        call("dyn_world_capacity", self._handle, ctypes.byref(pairs), ctypes.byref(contacts), ctypes.byref(grows))
        return int(pairs.value), int(contacts.value), int(grows.value)

    @property
    def capacity(self) -> int:
        """Contacts (and resting manifolds) a step can hold now."""
        if not self._handle.value:
            return default_capacities(len(self._bodies), self.contact_capacity, None)[0]
        return self.capacities()[1]

    @property
    def pair_capacity(self) -> int:
        """Candidate pairs a step can hold now."""
        if not self._handle.value:
            return default_capacities(len(self._bodies), None, self._pair_request)[1]
        return self.capacities()[0]

    def set_solver_threads(self, threads: int) -> None:
        """Workgroup size (64, 128, 256) of the solver substeps' kernels; results do not change."""
        self._ensure()
        call("dyn_world_set_solver_threads", self._handle, int(threads))

    def tune_solver(self, steps: int = 3) -> tuple[int, list[float]]:
        """Times the solver's workgroup sizes on this world's next steps (from a snapshot, then
        restored) and keeps the fastest: (size, ms per step of 64, 128, 256)."""
        self._ensure()
        threads = ctypes.c_uint32()
        ms = (ctypes.c_float * 3)()
        call("dyn_world_tune_solver", self._handle, int(steps), ctypes.byref(threads), ms)
        return int(threads.value), [float(v) for v in ms]

    def reserve(self, pair_capacity: int = 0, contact_capacity: int = 0) -> None:
        """Grows the pair and contact streams to at least the given capacities now."""
        self._ensure()
        call("dyn_world_reserve", self._handle, int(pair_capacity), int(contact_capacity))

    def refusals(self) -> list[dict]:
        """What the last step could not hold (World::refusals, crates/dynamis-world/src/
        readback.rs:226): every refusing counter above zero with its label and shortfall class
        (crates/dynamis-abi/src/counter.rs:75-128)."""
        m = self.measured()
        return [{"counter": name, "label": label, "shortfall": shortfall, "count": m[name]}
                for name, (label, shortfall) in abi.REFUSING_COUNTERS.items() if m[name] > 0]

    def convex_work(self) -> dict[str, int]:
        """Pairs the last step gave to GJK (those with a hull, those GJK resumed after its first
        iterations), to EPA (those whose polytope went to the wave kernel) and to the convex
        manifold kernel."""
        self._ensure()
        out = (ctypes.c_uint32 * 6)()
        call("dyn_world_read_work", self._handle, out)
        return {"gjk": int(out[0]), "manifold": int(out[1]), "epa": int(out[2]), "gjk_hull": int(out[3]),
                "gjk_resumed": int(out[4]), "epa_wave": int(out[5])}

    def _queue_edits(self, ids: np.ndarray, edits: np.ndarray) -> None:
        self._ensure()
        call("dyn_world_edit", self._handle, ids.ctypes.data, edits.ctypes.data, len(ids))

    def set_graphs(self, enabled: bool) -> None:
        """Replays each step as a captured HIP graph (the default) or launches it eagerly."""
        self._ensure()
        call("dyn_world_set_graphs", self._handle, 1 if enabled else 0)

    # -- reads -------------------------------------------------------------------------------
    def states(self) -> np.ndarray:
        """Body states in row order."""
        self._ensure()
        out = np.zeros(len(self.scene.states), dtype=abi.BODY_STATE)
        call("dyn_world_read_states", self._handle, out.ctypes.data)
        return out

    def read_state(self, handle: int) -> np.ndarray:
        return self.states()[int(self.scene.rows[handle])]

    def measured(self) -> dict[str, int]:
        """Counters of the last step by name (crates/dynamis-abi/src/counter.rs slots)."""
        self._ensure()
        block = np.zeros(abi.COUNTER_BLOCK_WORDS, dtype=np.uint32)
        call("dyn_world_read_counters", self._handle, block.ctypes.data)
        return {name: abi.counter(block, name) for name in abi.COUNTER_NAMES}

    def solve_residual(self) -> tuple[float, float]:
        """Largest linear and angular velocity change of the last solve round of the first
        substep (fixed-point words of crates/dynamis-world/src/readback.rs:219-229)."""
        m = self.measured()
        scale = float(abi.SOLVER_VELOCITY_SCALE)
        return (np.int32(np.uint32(m["SOLVE_LINEAR_RESIDUAL"])) / scale,
                np.int32(np.uint32(m["SOLVE_ANGULAR_RESIDUAL"])) / scale)

    def inspect_contacts(self) -> np.ndarray:
        """Contacts of the last step in canonical pair order (CONTACT records)."""
        self._ensure()
        out = np.zeros(self.capacity, dtype=abi.CONTACT)
        count = ctypes.c_uint32()
        call("dyn_world_read_contacts", self._handle, out.ctypes.data, self.capacity,
             ctypes.byref(count))
        return out[:count.value].copy()

    def resting_contacts(self) -> np.ndarray:
        """Manifolds frozen with their sleeping pairs (the resting archive)."""
        self._ensure()
        out = np.zeros(self.capacity, dtype=abi.CONTACT)
        count = ctypes.c_uint32()
        call("dyn_world_read_resting", self._handle, out.ctypes.data, self.capacity, ctypes.byref(count))
        return out[:min(count.value, self.capacity)].copy()

    def pairs(self) -> np.ndarray:
        """Sorted candidate pairs of the last step as (P, 2) collider indices."""
        self._ensure()
        major = np.zeros(self.pair_capacity, dtype=np.uint32)
        minor = np.zeros(self.pair_capacity, dtype=np.uint32)
        count = ctypes.c_uint32()
        call("dyn_world_read_pairs", self._handle, major.ctypes.data, minor.ctypes.data,
             self.pair_capacity, ctypes.byref(count))
        return np.stack([major[:count.value], minor[:count.value]], axis=1)

    def stream(self) -> int:
        self._ensure()
        return lib().dyn_world_stream(self._handle)

    # -- checkpoints (World::snapshot / restore, crates/dynamis-world/src/snapshot.rs:64-123) ---
    def snapshot(self):
        """The durable state copied into a new device buffer (on the world's stream)."""
        from .device import DeviceBuffer
        self._ensure()
        nbytes = ctypes.c_size_t()
        call("dyn_world_snapshot_bytes", self._handle, ctypes.byref(nbytes))
        buf = DeviceBuffer(nbytes.value)
        call("dyn_world_snapshot", self._handle, ctypes.c_void_p(buf.ptr))
        return buf

    def restore(self, buf) -> None:
        """Puts a snapshot of this world back (on the world's stream)."""
        self._ensure()
        call("dyn_world_restore", self._handle, ctypes.c_void_p(buf.ptr))

    def close(self) -> None:
        if self._handle.value:
            call("dyn_world_destroy", self._handle)
            self._handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self) -> "World":
        return self

    def __exit__(self, *exc) -> None:
        self.close()


@dataclass
class BatchLayout:
    """Where each world of a batch lives in the merged records."""

    dynamic_base: np.ndarray   # first merged row of each world's simulated rows
    static_base: np.ndarray    # first merged row of each world's other rows
    dynamic_count: np.ndarray
    body_count: np.ndarray
    collider_base: np.ndarray  # first merged collider slot of each world
    collider_count: np.ndarray
    id_base: np.ndarray        # offset added to each world's body ids
    collider_worlds: np.ndarray  # world of each merged collider

    def rows_of(self, world: int) -> np.ndarray:
        """Merged row of each of the world's own rows, in its row order."""
        d, n = int(self.dynamic_count[world]), int(self.body_count[world])
        dyn = self.dynamic_base[world] + np.arange(d, dtype=np.int64)
        rest = self.static_base[world] + np.arange(n - d, dtype=np.int64)
        return np.concatenate([dyn, rest])


def merge_scenes(scenes: list[Scene], config: np.ndarray | None = None) -> tuple[Scene, BatchLayout]:
    """The records of a batch of independent worlds (dyn_batch_create): all worlds' simulated
    rows first, world by world, then their other rows; colliders world by world in each world's
    arena order; body ids offset so they are unique; hull sources concatenated. Every world
    keeps the relative order of its rows, colliders and ids, so it steps as it does alone."""
    config = default_config() if config is None else config
    w = len(scenes)
    d = np.array([s.dynamic_count for s in scenes], dtype=np.int64)
    n = np.array([len(s.states) for s in scenes], dtype=np.int64)
    k = np.array([len(s.colliders) for s in scenes], dtype=np.int64)
    caps = np.array([int(s.states["body_id"].max()) + 1 if len(s.states) else 0 for s in scenes], dtype=np.int64)
    dyn_base = np.concatenate([[0], np.cumsum(d)[:-1]]) if w else d
    total_dynamic = int(d.sum())
    static_base = total_dynamic + (np.concatenate([[0], np.cumsum(n - d)[:-1]]) if w else d)
    collider_base = np.concatenate([[0], np.cumsum(k)[:-1]]) if w else k
    id_base = np.concatenate([[0], np.cumsum(caps)[:-1]]) if w else caps
    total = int(n.sum())
    states = np.zeros(total, dtype=abi.BODY_STATE)
    descs = np.zeros(total, dtype=abi.BODY_DESC)
    colliders = np.zeros(int(k.sum()), dtype=abi.COLLIDER)
    owners = np.zeros(int(k.sum()), dtype=np.uint32)
    collider_worlds = np.zeros(int(k.sum()), dtype=np.uint32)
    source_parts, vertex_parts = [], []
    source_base = vertex_base = 0
    layout = BatchLayout(dyn_base, static_base, d, n, collider_base, k, id_base, collider_worlds)
    for i, s in enumerate(scenes):
        rows = layout.rows_of(i)
        states[rows] = s.states
        states["body_id"][rows] += np.uint32(id_base[i])
        descs[rows] = s.descs
        c0, c1 = int(collider_base[i]), int(collider_base[i] + k[i])
        colliders[c0:c1] = s.colliders
        owners[c0:c1] = rows[s.owners.astype(np.int64)]
        collider_worlds[c0:c1] = i
        if s.sources is not None and len(s.sources):
            src = s.sources.copy()
            src["vertex_offset"] += np.uint32(vertex_base)
            source_parts.append(src)
            vertex_parts.append(s.vertices)
            hull = c0 + np.flatnonzero(colliders["kind"][c0:c1] == abi.SHAPE_HULL)
            field_view = colliders["source"]
            field_view[hull] += np.uint32(source_base)
            source_base += len(src)
            vertex_base += len(s.vertices)
    params = np.zeros(1, dtype=abi.STEP_PARAMS)
    call("dyn_step_params_build", config.ctypes.data, total, total_dynamic, int(k.sum()), params.ctypes.data)
    sources = np.concatenate(source_parts) if source_parts else None
    vertices = np.concatenate(vertex_parts) if vertex_parts else None
    merged = Scene(params=params, states=states, descs=descs, colliders=colliders, owners=owners,
                   dynamic_count=total_dynamic, rows=np.arange(total, dtype=np.uint32),
                   handles=np.arange(total, dtype=np.uint32), sources=sources, vertices=vertices)
    return merged, layout


class Batch(World):
    """Independent worlds stepped together as one device world (dyn_batch_create)."""

    def __init__(self, scenes: list[Scene], config: np.ndarray | None = None,
                 contact_capacity: int | None = None, pair_capacity: int | None = None):
        super().__init__(config, contact_capacity, pair_capacity)
        self.worlds = len(scenes)
        merged, self.layout = merge_scenes(scenes, self.config)
        self.scene = merged
        n = len(merged.states)
        capacity, pairs = default_capacities(n, contact_capacity, pair_capacity)
        call("dyn_batch_create", ctypes.byref(self._handle), None, merged.params.ctypes.data,
             merged.states.ctypes.data, merged.descs.ctypes.data, n, merged.dynamic_count,
             merged.colliders.ctypes.data, merged.owners.ctypes.data,
             self.layout.collider_worlds.ctypes.data, len(merged.colliders), self.worlds, capacity, pairs)
        if merged.sources is not None and len(merged.sources):
            call("dyn_world_upload_shapes", self._handle, merged.sources.ctypes.data,
                 len(merged.sources), merged.vertices.ctypes.data, len(merged.vertices))

    def world_states(self, world: int, states: np.ndarray | None = None) -> np.ndarray:
        """One world's body states in its own row order, with its own body ids."""
        states = self.states() if states is None else states
        out = states[self.layout.rows_of(world)].copy()
        out["body_id"] -= np.uint32(self.layout.id_base[world])
        return out

    def world_contacts(self, world: int, contacts: np.ndarray | None = None) -> np.ndarray:
        """One world's contacts in step order, with its own collider slots and body ids."""
        contacts = self.inspect_contacts() if contacts is None else contacts
        c0 = int(self.layout.collider_base[world])
        mine = (contacts["a"] >= c0) & (contacts["a"] < c0 + int(self.layout.collider_count[world]))
        out = contacts[mine].copy()
        out["a"] -= np.uint32(c0)
        out["b"] -= np.uint32(c0)
        out["first_body_id"] -= np.uint32(self.layout.id_base[world])
        out["second_body_id"] -= np.uint32(self.layout.id_base[world])
        return out
