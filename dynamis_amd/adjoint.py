"""The differentiable rollout: record layouts of the frozen-contact tapes and the device adjoint.

A rollout of `steps` steps runs on a World or Batch (dyn_adjoint_forward), keeping a checkpoint of
the durable state every `every` steps. The backward pass (dyn_adjoint_backward) walks the segments
last to first: it restores a segment's checkpoint, re-steps it while recording each step's tape
(the step-start states and the contacts as the solver received them), then runs, per step and per
world, the replay and reverse of dynamis_amd/csrc/adjoint.h in one workgroup. Its outputs are the
adjoints of the initial states, of each body's mass (at fixed shape: inertia scales with it) and
of each collider's friction coefficient.

The record builders below restate on the host what the device recorder writes, for tapes taken
from other sources (the CPU checks build them from the oracle's steps).
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import abi

STATE_FLOATS = 16
BODY_FLOATS = 20
CONTACT_FLOATS = 76
ACC_FLOATS = 12
NO_SOURCE = 0xFFFFFFFF

BODY_SIMULATES, BODY_MOVES, BODY_INERT = 1, 2, 4
FLAG_POINTS, FLAG_INERT_F, FLAG_INERT_S, FLAG_RESOLVES = 7, 8, 16, 32

PARAMS = np.dtype([
    ("dt", "<f4"), ("h", "<f4"), ("gravity", "<f4", 3), ("slop", "<f4"), ("relaxation", "<f4"),
    ("restitution_threshold", "<f4"), ("max_velocity", "<f4"), ("max_angular_velocity", "<f4"),
    ("substeps", "<u4"), ("iterations", "<u4"), ("position_iterations", "<u4"),
    ("friction_combine", "<u4"),
])
RANGE = np.dtype([("row_base", "<u4"), ("row_count", "<u4"), ("collider_base", "<u4"),
                  ("collider_count", "<u4")])


def adj_params(step_params: np.ndarray, dt: float) -> np.ndarray:
    """The differentiable step's parameters from a STEP_PARAMS record and the step's dt."""
    p = step_params.reshape(-1)[0]
    out = np.zeros(1, dtype=PARAMS)
    out["dt"] = np.float32(dt)
    out["h"] = np.float32(dt) / np.float32(p["substeps"])
    out["gravity"] = p["gravity"][:3]
    for name in ("slop", "relaxation", "restitution_threshold", "max_velocity", "max_angular_velocity",
                 "friction_combine"):
        out[name] = p[name]
    out["substeps"] = p["substeps"]
    out["iterations"] = p["solve_iterations"]
    out["position_iterations"] = p["position_iterations"]
    return out


def simulates(descs: np.ndarray) -> np.ndarray:
    return ((descs["flags"] & abi.BODY_KINEMATIC) == 0) & (descs["inverse_mass"] > 0)


def moves(descs: np.ndarray) -> np.ndarray:
    return ((descs["flags"] & abi.BODY_KINEMATIC) != 0) | (descs["inverse_mass"] > 0)


def body_records(states: np.ndarray, descs: np.ndarray) -> np.ndarray:
    """Per row: inverse mass, gravity scale, dampings, body inverse inertia and inertia, centre of
    mass, flags (simulates, moves, inert in the solver)."""
    n = len(descs)
    out = np.zeros((n, BODY_FLOATS), np.float32)
    out[:, 0] = descs["inverse_mass"]
    out[:, 1] = descs["gravity_scale"]
    out[:, 2] = descs["linear_damping"]
    out[:, 3] = descs["angular_damping"]
    out[:, 4:10] = descs["inverse_inertia"]
    out[:, 10:16] = descs["inertia"]
    out[:, 16:19] = descs["com"]
    sim, mov = simulates(descs), moves(descs)
    inert = ~sim | (states["sleeping"] != 0)
    flags = (sim * BODY_SIMULATES) | (mov * BODY_MOVES) | (inert * BODY_INERT)
    out[:, 19] = flags.astype(np.uint32).view(np.float32)
    return out


def state_records(states: np.ndarray) -> np.ndarray:
    out = np.zeros((len(states), STATE_FLOATS), np.float32)
    out[:, 0:3] = states["position"]
    out[:, 4:8] = states["orientation"]
    out[:, 8:11] = states["velocity"]
    out[:, 12:15] = states["angular_velocity"]
    return out


def friction_partials(fa: np.ndarray, fb: np.ndarray, mode: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """material_combine (core.wgsl) of two coefficients and its partial derivatives."""
    fa = fa.astype(np.float32)
    fb = fb.astype(np.float32)
    if mode == 1:
        pick_a = fa < fb
        return np.where(pick_a, fa, fb), pick_a.astype(np.float32), (~pick_a).astype(np.float32)
    if mode == 2:
        pick_a = fa > fb
        return np.where(pick_a, fa, fb), pick_a.astype(np.float32), (~pick_a).astype(np.float32)
    if mode == 3:
        half = np.full_like(fa, 0.5)
        return (fa + fb) * np.float32(0.5), half, half
    c = np.sqrt(fa * fb)
    safe = np.where(c > 0, c, 1.0).astype(np.float32)
    return c, np.where(c > 0, fb / (2 * safe), 0).astype(np.float32), np.where(c > 0, fa / (2 * safe), 0).astype(np.float32)


def contact_records(contacts: np.ndarray, relay_src: np.ndarray, owners: np.ndarray, colliders: np.ndarray,
                    states: np.ndarray, descs: np.ndarray, friction_combine: int) -> np.ndarray:
    """Contacts as the solver received them (after the warm-start relay) into tape records."""
    m = len(contacts)
    out = np.zeros((m, CONTACT_FLOATS), np.float32)
    u = out.view(np.uint32)
    if m == 0:
        return out
    fb = owners[contacts["a"]]
    sb = owners[contacts["b"]]
    inert = ~simulates(descs) | (states["sleeping"] != 0)
    pc = contacts["point_count"].astype(np.uint32)
    resolves = (pc > 0) & (contacts["sensor"] == 0)
    flags = pc | (inert[fb] * FLAG_INERT_F) | (inert[sb] * FLAG_INERT_S) | (resolves * FLAG_RESOLVES)
    u[:, 0] = fb
    u[:, 1] = sb
    u[:, 2] = flags
    u[:, 3] = contacts["a"]
    out[:, 4:7] = contacts["normal"]
    out[:, 7] = contacts["friction"]
    _, da, db = friction_partials(colliders["friction"][contacts["a"]], colliders["friction"][contacts["b"]],
                                  friction_combine)
    out[:, 8] = da
    out[:, 9] = db
    out[:, 10] = contacts["restitution"]
    u[:, 11] = contacts["b"]
    pts = contacts["points"]
    for p in range(4):
        base = 12 + 16 * p
        out[:, base:base + 3] = pts["local_a"][:, p]
        out[:, base + 3] = pts["depth"][:, p]
        out[:, base + 4:base + 7] = pts["local_b"][:, p]
        out[:, base + 8:base + 11] = pts["position"][:, p]
        u[:, base + 11] = relay_src[:, p]
        out[:, base + 12] = pts["accumulated_normal"][:, p]
        out[:, base + 13] = pts["accumulated_tangent_1"][:, p]
        out[:, base + 14] = pts["accumulated_tangent_2"][:, p]
    return out


def unsupported(contacts: np.ndarray) -> str | None:
    """Why a step's contacts fall outside the differentiable model, or None."""
    resolves = (contacts["point_count"] > 0) & (contacts["sensor"] == 0)
    if np.any(resolves & (contacts["relaxation"] > 0)):
        return "soft contacts (relaxation > 0)"
    if np.any(resolves & ((contacts["rolling_friction"] != 0) | (contacts["spin_friction"] != 0))):
        return "rolling or spin friction"
    return None


def world_ranges(world) -> np.ndarray:
    """The adjoint's per-world ranges of a World (one world) or a Batch (its layout)."""
    layout = getattr(world, "layout", None)
    if layout is None:
        scene = world.scene
        return np.array([[0, scene.dynamic_count, 0, len(scene.colliders)]], np.uint32)
    return np.stack([layout.dynamic_base, layout.dynamic_count, layout.collider_base, layout.collider_count],
                    axis=1).astype(np.uint32)


class _Device:
    """A device buffer of the library's allocator (no framework needed)."""

    def __init__(self, nbytes: int):
        from ._native import call
        self.ptr = ctypes.c_void_p()
        self.nbytes = max(int(nbytes), 4)
        call("dyn_malloc", ctypes.byref(self.ptr), self.nbytes)

    def upload(self, array: np.ndarray) -> None:
        from ._native import call
        a = np.ascontiguousarray(array)
        call("dyn_memcpy_htod", self.ptr, a.ctypes.data, a.nbytes)

    def download(self, array: np.ndarray) -> np.ndarray:
        from ._native import call
        call("dyn_memcpy_dtoh", array.ctypes.data, self.ptr, array.nbytes)
        return array

    def free(self) -> None:
        from ._native import call
        if self.ptr:
            call("dyn_free", self.ptr)
            self.ptr = ctypes.c_void_p()


class Adjoint:
    """A differentiable rollout of `steps` steps of a World or Batch (dyn_adjoint_*), with a
    checkpoint every `every` steps. forward(dt) steps the world; backward(grad) takes the adjoint
    of the final states (rows x 16 floats, the state layout above) and returns the adjoints of the
    initial states, of each row's mass and of each collider's friction coefficient."""

    def __init__(self, world, steps: int, every: int = 8, ranges: np.ndarray | None = None):
        from ._native import call
        self.world = world
        self.steps = int(steps)
        self.every = int(every)
        self.ranges = world_ranges(world) if ranges is None else np.ascontiguousarray(ranges, np.uint32)
        self.n = len(world.scene.states)
        self.k = len(world.scene.colliders)
        self._handle = ctypes.c_void_p()
        call("dyn_adjoint_create", ctypes.byref(self._handle), world._handle, self.steps, self.every,
             self.ranges.ctypes.data, len(self.ranges))
        self._grad = _Device(4 * self.n * STATE_FLOATS)
        self._grad0 = _Device(4 * self.n * STATE_FLOATS)
        self._mass = _Device(4 * self.n)
        self._friction = _Device(4 * max(self.k, 1))

    def forward(self, dt: float) -> None:
        from ._native import call
        call("dyn_adjoint_forward", self._handle, ctypes.c_float(dt))

    def backward(self, grad_final: np.ndarray) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        from ._native import call
        g = np.ascontiguousarray(grad_final, np.float32).reshape(self.n, STATE_FLOATS)
        self._grad.upload(g)
        call("dyn_adjoint_backward", self._handle, self._grad.ptr, self._grad0.ptr, self._mass.ptr,
             self._friction.ptr)
        call("dyn_world_wait", self.world._handle)
        g0 = self._grad0.download(np.zeros((self.n, STATE_FLOATS), np.float32))
        gm = self._mass.download(np.zeros(self.n, np.float32))
        gf = self._friction.download(np.zeros(max(self.k, 1), np.float32))[:self.k]
        return g0, gm, gf

    def tape(self, slot: int) -> tuple[np.ndarray, np.ndarray]:
        """The recorded start states and contact records of a step of the last recomputed segment."""
        from ._native import call
        cap = self.world.capacity
        states = np.zeros((self.n, STATE_FLOATS), np.float32)
        contacts = np.zeros((cap, CONTACT_FLOATS), np.float32)
        count = ctypes.c_uint32()
        call("dyn_adjoint_read_tape", self._handle, slot, states.ctypes.data, contacts.ctypes.data, cap,
             ctypes.byref(count))
        return states, contacts[:min(count.value, cap)].copy()

    def bodies(self) -> np.ndarray:
        from ._native import call
        out = np.zeros((self.n, BODY_FLOATS), np.float32)
        call("dyn_adjoint_read_bodies", self._handle, out.ctypes.data)
        return out

    def close(self) -> None:
        from ._native import call
        if self._handle:
            call("dyn_adjoint_destroy", self._handle)
            self._handle = ctypes.c_void_p()
        for b in (self._grad, self._grad0, self._mass, self._friction):
            b.free()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
