"""dyn_islands on the device against scipy's connected components of the same contacts (each
simulated row's root = its island's minimum row, exactly) and against the island count of the
device's own step. Cases: hand-built contact sets (chains, a static bridge, a sensor, a contact
outside the slop, colliders outside the records, no contacts), and the contacts of a settling
3000-body mixed pile."""

from __future__ import annotations

import numpy as np
import pytest

from dynamis_amd.islands import Islands
from dynamis_amd.world import World
from tests.island_cases import expected_roots
from tests.narrow_cases import DT, pile_scene
from tests.test_island_cases import contact_set
from tests.worlds import config, cube_drop, encode

pytestmark = pytest.mark.gpu


def device_islands(contacts, owners, bodies, dynamic, slop):
    isl = Islands(owners, bodies, dynamic, slop)
    isl.run(contacts)
    return isl.read()


def test_hand_built_sets(device):
    owners = np.arange(7, dtype=np.uint32)
    cases = [contact_set([(1, 2), (2, 0), (4, 6), (6, 5)]), contact_set([(3, 4)], depth=-0.02),
             contact_set([(3, 4), (4, 5)], sensor=(0,)), contact_set([]),
             contact_set([(0, 9), (12, 1), (2, 3)])]
    for i, contacts in enumerate(cases):
        parents, islands, faults = device_islands(contacts, owners, 7, 6, 0.01)
        valid = contacts[(contacts["a"] < 7) & (contacts["b"] < 7)]
        roots, count = expected_roots(valid, owners, 6, 0.01)
        assert parents.tolist() == roots.tolist(), i
        assert islands == count and faults == 0, i


@pytest.mark.parametrize("kind", ["cube_drop", "mixed_pile"])
def test_step_contacts(device, kind):
    if kind == "cube_drop":
        scene = encode(cube_drop(1000, seed=1, footprint=12.0), config(sleep_velocity=0.0, sleep_angular_velocity=0.0))
    else:
        scene = pile_scene(3000, seed=5, hull_sources=64, spacing=1.5)
    world = World()
    world.create(scene)
    slop = float(scene.params["slop"][0])
    checked = 0
    for frame in range(120):
        world.step(DT)
        if frame % 30 == 29:
            contacts = world.inspect_contacts()
            parents, islands, faults = device_islands(contacts, scene.owners, len(scene.states),
                                                      scene.dynamic_count, slop)
            roots, count = expected_roots(contacts, scene.owners, scene.dynamic_count, slop)
            assert np.array_equal(parents, roots), (kind, frame)
            assert islands == count and faults == 0, (kind, frame)
            if kind == "cube_drop":  # no sleeping, so no resting links: the step's own count
                assert islands == world.measured()["ISLANDS"], frame
            checked += count < scene.dynamic_count - 10
    world.close()
    assert checked > 0
