"""dyn_solve on the device against the oracle's solver alone (pinned to the oracle's whole step
by the CPU tests): from the states and warm-started contacts a step hands its solver, the device
ends bit-identical to the oracle in every state field and every byte of every contact. Cases: a
200-cube drop and a 60-body mixed pile (spheres, boxes, capsules, hulls, cylinders) without
sleeping, a 300-cube drop with sleeping bodies among the rows, no contacts, and a contact naming
a collider outside the records (refused, nothing solved)."""

from __future__ import annotations

import numpy as np
import pytest

from dynamis_amd._native import DynamisError
from dynamis_amd.solver import Solver
from tests.narrow_cases import DT
from tests.solve_cases import NO_SLEEP, drop_scene, solver_inputs
from tests.worlds import config, cube_drop, ellipsoid_hulls, encode, mixed_pile

pytestmark = pytest.mark.gpu


def check_parity(oracle, scene, frames, least_contacts):
    inputs, contacts, _, _ = solver_inputs(oracle, scene, frames)
    assert len(contacts) >= least_contacts
    ref = oracle.world(inputs)
    want_contacts = ref.solve(contacts, DT)
    want_states = ref.states()
    ref.close()
    solver = Solver(inputs, max(len(contacts), 1))
    solver.run(inputs.states, contacts, DT)
    states, got_contacts, counters = solver.read()
    assert states.tobytes() == want_states.tobytes()
    assert got_contacts.tobytes() == want_contacts.tobytes()
    return inputs, states


def test_cube_drop(device, oracle):
    check_parity(oracle, drop_scene(), 45, 100)


def test_mixed_pile(device, oracle):
    hulls = ellipsoid_hulls(4, seed=3)
    scene = encode(mixed_pile(60, seed=2, kinds=5, hull_sources=4, spacing=1.6), config(**NO_SLEEP), hulls)
    check_parity(oracle, scene, 45, 20)


def test_sleeping_rows_stay(device, oracle):
    scene = encode(cube_drop(300, seed=1, footprint=10.0, high=4.0), config())  # default sleeping
    inputs, after = check_parity(oracle, scene, 200, 500)
    asleep = inputs.states["sleeping"][:scene.dynamic_count] != 0
    assert asleep.any()
    assert after[:scene.dynamic_count][asleep].tobytes() == inputs.states[:scene.dynamic_count][asleep].tobytes()


def test_no_contacts_and_foreign_contacts(device, oracle):
    scene = drop_scene(20)
    inputs, contacts, _, _ = solver_inputs(oracle, scene, 30)
    solver = Solver(inputs, max(len(contacts), 1))
    solver.run(inputs.states, contacts[:0], DT)
    ref = oracle.world(inputs)
    ref.solve(contacts[:0], DT)
    states, _, _ = solver.read()
    assert states.tobytes() == ref.states().tobytes()
    ref.close()
    bad = contacts.copy()
    bad[0]["b"] = len(scene.colliders) + 3
    with pytest.raises(DynamisError, match="outside the records"):
        solver.run(inputs.states, bad, DT)
