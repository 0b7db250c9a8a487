"""The oracle's solver alone (oracle_world_solve, the checker of dyn_solve) against the oracle's
whole step: given the states and contacts the step handed its solver, it ends where the step
ended (every state field the solver writes, every byte of every contact)."""

from __future__ import annotations

import numpy as np
import pytest

from tests.narrow_cases import DT, pile_scene
from tests.solve_cases import NO_SLEEP, drop_scene, solver_inputs, state_fields_equal
from tests.worlds import config, encode, ellipsoid_hulls, mixed_pile


@pytest.mark.parametrize("kind", ["cube_drop", "mixed_pile"])
def test_stage_reproduces_the_step(oracle, kind):
    if kind == "cube_drop":
        scene = drop_scene()
    else:
        hulls = ellipsoid_hulls(4, seed=3)
        scene = encode(mixed_pile(60, seed=2, kinds=5, hull_sources=4, spacing=1.6), config(**NO_SLEEP), hulls)
    inputs, contacts, end, end_contacts = solver_inputs(oracle, scene, 45)
    assert len(contacts) > 20 and np.abs(contacts["points"]["accumulated_normal"]).sum() > 0
    stage = oracle.world(inputs)
    solved = stage.solve(contacts, DT)
    assert state_fields_equal(stage.states(), end) == []
    assert solved.tobytes() == end_contacts.tobytes()
    stage.close()


def test_stage_refuses_foreign_contacts(oracle):
    scene = drop_scene(20)
    inputs, contacts, _, _ = solver_inputs(oracle, scene, 30)
    bad = contacts.copy()
    bad[0]["a"] = len(scene.colliders)
    stage = oracle.world(inputs)
    with pytest.raises(ValueError):
        stage.solve(bad, DT)
    stage.close()
