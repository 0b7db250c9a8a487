"""Inputs of the solver stage tests: the states and warm-started contacts a step hands its
solver, taken from the oracle's step."""

from __future__ import annotations

import numpy as np

from dynamis_amd import abi
from tests.narrow_cases import with_states
from tests.worlds import DT, config, cube_drop, encode

NO_SLEEP = dict(sleep_velocity=0.0, sleep_angular_velocity=0.0)


def begin_step(states: np.ndarray, dynamic_count: int) -> np.ndarray:
    """begin_step.wgsl: awake simulated rows keep their pose as the previous pose."""
    s = states.copy()
    rows = np.arange(dynamic_count)
    awake = rows[s["sleeping"][:dynamic_count] == 0]
    s["prev_position"][awake] = s["position"][awake]
    s["prev_orientation"][awake] = s["orientation"][awake]
    return s


def solver_inputs(oracle, scene, frames: int):
    """Steps the oracle `frames` times and returns (scene with the last step's solver-entry states,
    that step's solver contacts, the step's end states, the step's end contacts)."""
    ref = oracle.world(scene)
    for _ in range(frames):
        ref.step(DT)
    start, pre = ref.tape()
    end, end_contacts = ref.states(), ref.inspect_contacts()
    ref.close()
    return with_states(scene, begin_step(start, scene.dynamic_count)), pre, end, end_contacts


def drop_scene(count=200, seed=3, footprint=8.0, **fields):
    return encode(cube_drop(count, seed=seed, footprint=footprint, high=12.0), config(**{**NO_SLEEP, **fields}))


def state_fields_equal(got: np.ndarray, want: np.ndarray, skip=("sleep_timer",)) -> list[str]:
    return [f for f in abi.BODY_STATE.names if f not in skip and got[f].tobytes() != want[f].tobytes()]
