"""Host side of the differentiable-step checks: the frozen-contact rollout in FP64 and its reverse
(oracle/_build/libadjhost.so, built from oracle/adjoint_host.cpp), and tapes recorded by the
CPU oracle's steps.

Central differences of the FP64 rollout's loss are the gradient oracle; the reverse run on the host
checks the adjoint's mathematics on the CPU, and the device adjoint is checked against the same
differences on the device's own tapes.
"""

from __future__ import annotations

import ctypes
import pathlib
import subprocess

import numpy as np

from dynamis_amd import abi
from dynamis_amd import adjoint as adj

from . import oracle_binding

ROOT = pathlib.Path(__file__).resolve().parents[1]
HOST_SO = ROOT / "oracle" / "_build" / "libadjhost.so"

_p = ctypes.c_void_p
_u32 = ctypes.c_uint32


class Host:
    def __init__(self):
        if not HOST_SO.exists():
            subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
        h = ctypes.CDLL(str(HOST_SO))
        h.adj_host_rollout.restype = ctypes.c_int
        h.adj_host_rollout.argtypes = [_p, _p, _u32, _u32, _u32, _u32, _p, _u32, _p, _p, _p, _p, _p, _p, _p,
                                       ctypes.c_int]
        h.adj_host_gradient.restype = ctypes.c_int
        h.adj_host_gradient.argtypes = [_p, _p, _u32, _u32, _u32, _u32, _p, _u32, _p, _p, _p, _p, _p, _p, _p,
                                        _p, ctypes.c_int, ctypes.c_int]
        h.adj_host_sizes.restype = _u32
        h.adj_host_sizes.argtypes = [_p]
        sizes = np.zeros(4, np.uint32)
        h.adj_host_sizes(sizes.ctypes.data)
        assert sizes[0] == adj.PARAMS.itemsize and sizes[1] == adj.RANGE.itemsize
        assert sizes[2] == adj.CONTACT_FLOATS and sizes[3] == adj.BODY_FLOATS
        self.h = h


_host: Host | None = None


def host() -> Host:
    global _host
    if _host is None:
        _host = Host()
    return _host


class Tape:
    """A recorded rollout: parameters, world ranges, body records and, per step, the contact
    records and the start states."""

    def __init__(self, params, ranges, bodies, contacts, starts, colliders, capacity):
        self.params = params
        self.ranges = ranges
        self.bodies = np.ascontiguousarray(bodies, np.float32)
        self.contacts = [np.ascontiguousarray(c, np.float32) for c in contacts]
        self.starts = [np.ascontiguousarray(s, np.float32) for s in starts]
        self.n = len(bodies)
        self.k = colliders
        self.capacity = capacity
        self.steps = len(contacts)
        # the host functions index contact records up to the capacity
        self._padded = []
        for c in self.contacts:
            pad = np.zeros((max(capacity, 1), adj.CONTACT_FLOATS), np.float32)
            pad[:len(c)] = c
            self._padded.append(pad)
        self._cptr = (ctypes.c_void_p * self.steps)(*[c.ctypes.data for c in self._padded])
        self._fptr = (ctypes.c_void_p * self.steps)(*[s.ctypes.data for s in self.starts])
        self._counts = np.array([len(c) for c in self.contacts], np.uint32)

    def x0(self) -> np.ndarray:
        return self.starts[0].astype(np.float64)

    def rollout(self, x0=None, s=None, mu=None, anchored=False) -> np.ndarray:
        x0 = self.x0() if x0 is None else np.ascontiguousarray(x0, np.float64)
        out = np.zeros_like(x0)
        sp = None if s is None else np.ascontiguousarray(s, np.float64)
        mp = None if mu is None else np.ascontiguousarray(mu, np.float64)
        host().h.adj_host_rollout(self.params.ctypes.data, self.ranges.ctypes.data, len(self.ranges), self.n,
                                  self.k, self.capacity, self.bodies.ctypes.data, self.steps, self._cptr,
                                  self._counts.ctypes.data, self._fptr, x0.ctypes.data,
                                  None if sp is None else sp.ctypes.data, None if mp is None else mp.ctypes.data,
                                  out.ctypes.data, int(anchored))
        return out

    def gradient(self, xbar: np.ndarray, single: bool = False, x0=None, anchored=False):
        x0 = self.x0() if x0 is None else np.ascontiguousarray(x0, np.float64)
        xbar = np.ascontiguousarray(xbar, np.float64)
        gx = np.zeros_like(x0)
        gs = np.zeros(self.n, np.float64)
        gm = np.zeros(max(self.k, 1), np.float64)
        host().h.adj_host_gradient(self.params.ctypes.data, self.ranges.ctypes.data, len(self.ranges), self.n,
                                   self.k, self.capacity, self.bodies.ctypes.data, self.steps, self._cptr,
                                   self._counts.ctypes.data, self._fptr, x0.ctypes.data, xbar.ctypes.data,
                                   gx.ctypes.data, gs.ctypes.data, gm.ctypes.data, int(single), int(anchored))
        return gx, gs, gm[:self.k]


def oracle_tape(scene, steps: int, dt: float) -> tuple[Tape, list[np.ndarray]]:
    """Steps the CPU oracle and records each step's tape. Returns the tape and the oracle's states
    after every step."""
    o = oracle_binding.load()
    h = o.h
    h.oracle_world_tape.restype = _u32
    h.oracle_world_tape.argtypes = [_p, _p, _p, _p, _u32]
    world = o.world(scene)
    n, cap = world.n, world.capacity
    params = adj.adj_params(scene.params, dt)
    ranges = np.zeros(1, adj.RANGE)
    ranges[0] = (0, scene.dynamic_count, 0, len(scene.colliders))
    contacts, starts, after = [], [], []
    bodies = None
    mode = int(scene.params["friction_combine"][0])
    for _ in range(steps):
        world.step(dt)
        start = np.zeros(n, abi.BODY_STATE)
        pre = np.zeros(cap, abi.CONTACT)
        src = np.zeros((cap, 4), np.uint32)
        count = h.oracle_world_tape(world.w, start.ctypes.data, pre.ctypes.data, src.ctypes.data, cap)
        if bodies is None:
            bodies = adj.body_records(start, scene.descs)
        why = adj.unsupported(pre[:count])
        assert why is None, why
        contacts.append(adj.contact_records(pre[:count], src[:count], scene.owners, scene.colliders, start,
                                            scene.descs, mode))
        starts.append(adj.state_records(start))
        after.append(world.states())
    world.close()
    return Tape(params, ranges, bodies, contacts, starts, len(scene.colliders), cap), after


def differences(f, h: float, unit: float = 1.0) -> tuple[float, float, float]:
    """Central, forward and backward differences of f (the loss at the parameter moved by e) at
    step h, divided by `unit` (a parameter's scale, for derivatives taken in a scaled variable)."""
    lp, l0, lm = f(h), f(0.0), f(-h)
    return (lp - lm) / (2 * h) / unit, (lp - l0) / h / unit, (l0 - lm) / h / unit


def classify(dev: float, diffs: tuple[float, float, float], scale: float, tol: float = 1e-3) -> str:
    """'smooth' when the central difference agrees with `dev` within tol (relative, floored at 1e-2 of
    the largest derivative, `scale`); 'kink' when the parameter sits at a kink of the piecewise-smooth
    rollout within h: the forward and backward differences differ by more than tol, and `dev` (the
    slope of the branch the recording selects) agrees with one of them within tol; else 'off'."""
    central, forward, backward = diffs
    floor = 1e-2 * scale

    def rel(a, b):
        return abs(a - b) / max(abs(b), floor)

    if rel(dev, central) <= tol:
        return "smooth"
    if rel(forward, backward) > tol and min(rel(dev, forward), rel(dev, backward)) <= tol:
        return "kink"
    return "off"
