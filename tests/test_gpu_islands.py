"""dyn_islands on the device against scipy's connected components of the same contacts (each
simulated row's root = its island's minimum row, exactly) and against the island count of the
device's own step. Cases: hand-built contact sets (chains, a static bridge, a sensor, a contact
outside the slop, colliders outside the records, no contacts), long chains whose seeded forest is
as deep as the chain (in row order, reversed and shuffled, above and below the one-workgroup
flatten's 8192 rows, with a few crossing links), and the contacts of a settling 3000-body mixed
pile."""

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


def chain_contacts(rows, order, seed):
    rng = np.random.default_rng(seed)
    line = np.arange(rows)
    if order == "reversed":
        line = line[::-1]
    elif order == "shuffled":
        line = rng.permutation(rows)
    links = np.stack([line[:-1], line[1:]], axis=1)
    # two chains instead of one (a cut at the middle), rejoined by a far link except in reverse
    # order; a few extra links across the first half of the chain
    links = np.delete(links, rows // 2, axis=0)
    extra = np.stack([rng.integers(0, rows // 2, 6), rng.integers(0, rows // 2, 6)], axis=1)
    rejoin = [[line[3], line[rows - 3]]] if order != "reversed" else np.zeros((0, 2), np.int64)
    links = np.concatenate([links, extra, rejoin])
    links = links[rng.permutation(len(links))]
    c = np.zeros(len(links), dtype=contact_set([]).dtype)
    c["a"], c["b"], c["point_count"] = links[:, 0], links[:, 1], 1
    return c


@pytest.mark.parametrize("rows", [5000, 20000])
@pytest.mark.parametrize("order", ["rows", "reversed", "shuffled"])
def test_long_chains(device, rows, order):
    contacts = chain_contacts(rows, order, seed=rows)
    owners = np.arange(rows + 1, dtype=np.uint32)  # the last row static
    parents, islands, faults = device_islands(contacts, owners, rows + 1, rows, 0.01)
    roots, count = expected_roots(contacts, owners, rows, 0.01)
    assert np.array_equal(parents, roots)
    assert islands == count and faults == 0


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
