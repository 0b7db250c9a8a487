"""The scipy island expectation (the checker of dyn_islands) on hand-built contact sets, and
against the island count of the oracle's step."""

from __future__ import annotations

import numpy as np

from dynamis_amd import abi
from tests.island_cases import expected_roots
from tests.worlds import DT, config, cube_drop, encode


def contact_set(links, depth=0.0, sensor=()):
    c = np.zeros(len(links), dtype=abi.CONTACT)
    for i, (a, b) in enumerate(links):
        c[i]["a"], c[i]["b"], c[i]["point_count"] = a, b, 1
        c[i]["points"]["depth"][0] = depth
        c[i]["sensor"] = 1 if i in sensor else 0
    return c


def test_hand_built_islands():
    # rows 0-5 simulated, row 6 static; collider i belongs to row i
    owners = np.arange(7, dtype=np.uint32)
    roots, count = expected_roots(contact_set([(1, 2), (2, 0), (4, 6), (6, 5)]), owners, 6, 0.01)
    # 0-1-2 joined; a static row links nothing: 3, 4, 5 stay alone
    assert roots.tolist() == [0, 0, 0, 3, 4, 5] and count == 4
    roots, count = expected_roots(contact_set([(3, 4)], depth=-0.02), owners, 6, 0.01)
    assert count == 6  # not touching within the slop
    roots, count = expected_roots(contact_set([(3, 4), (4, 5)], sensor=(0,)), owners, 6, 0.01)
    assert roots.tolist() == [0, 1, 2, 3, 4, 4] and count == 5


def test_matches_the_oracle_island_count(oracle):
    """Without sleeping there is no resting archive, so the step's islands are the components of
    its touching contacts."""
    scene = encode(cube_drop(120, seed=4, footprint=6.0), config(sleep_velocity=0.0, sleep_angular_velocity=0.0))
    ref = oracle.world(scene)
    checked = 0
    for frame in range(90):
        ref.step(DT)
        if frame % 15 == 14:
            contacts = ref.inspect_contacts()
            _, count = expected_roots(contacts, scene.owners, scene.dynamic_count, float(scene.params["slop"][0]))
            assert count == ref.measured()["ISLANDS"], frame
            checked += count < scene.dynamic_count
    assert checked > 0  # some frames had islands of more than one body
    ref.close()
