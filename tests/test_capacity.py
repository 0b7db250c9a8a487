"""Capacity planning (World::apply_plan, crates/dynamis-world/src/backend/mod.rs:194-230): a step
whose demand outgrew a stream refuses what does not fit, counts it in the refusing counters
(REFUSED_PAIRS, REFUSED_CONTACTS), and the stream grows to max(demand, 2 x capacity)
(crates/dynamis-domain/src/capacity.rs:15-23) two steps later, after which nothing is refused.
On the oracle here; the device's plan is checked against the oracle's in tests/test_gpu_capacity.py.
"""

from __future__ import annotations

import numpy as np

from dynamis_amd.scenes import cube_drop
from dynamis_amd.world import default_config, encode

DT = np.float32(1.0 / 60.0)
LAG = 2


def grown(capacity: int, required: int) -> int:
    """dynamis-domain capacity.rs:15-23 with the floor met."""
    return capacity if required <= capacity else max(required, 2 * capacity)


def trace(world, steps: int):
    """Per step: the capacities it ran with (a step plans before it runs), its demand and its
    refusals."""
    rows = []
    for _ in range(steps):
        world.step(DT)
        m = world.measured()
        rows.append((world.capacities(), m["PAIRS"], max(m["CONTACTS"], m["RESTING"] + m["REFUSED_RESTING"]),
                     m["REFUSED_PAIRS"], m["REFUSED_CONTACTS"]))
    return rows


def check_plan(rows):
    """Each step ran with the capacities the demand of the step two back asked for."""
    pair_cap, contact_cap = rows[0][0][0], rows[0][0][1]
    for t, ((pcap, ccap, _), pairs, contacts, refused_pairs, refused_contacts) in enumerate(rows):
        if t >= LAG:
            _, dp, dc, _, _ = rows[t - LAG]
            pair_cap, contact_cap = grown(pair_cap, dp), grown(contact_cap, dc)
        assert (pcap, ccap) == (pair_cap, contact_cap), t
        assert refused_pairs == max(0, pairs - pcap), t
        assert refused_contacts <= max(0, contacts - ccap), t


def test_streams_grow_two_steps_after_their_demand(oracle):
    scene = encode(cube_drop(200, seed=1), default_config())
    world = oracle.world(scene, 64, 256)
    rows = trace(world, 150)
    check_plan(rows)
    grows = world.capacities()[2]
    assert grows >= 2
    refused = [t for t, r in enumerate(rows) if r[3] or r[4]]
    assert refused, "the starting capacities are below the drop's demand"
    # once grown past the peak demand nothing more is refused
    peak = max(r[1] for r in rows), max(r[2] for r in rows)
    settled = [t for t, r in enumerate(rows) if r[0][0] >= peak[0] and r[0][1] >= peak[1]]
    assert settled and all(not (rows[t][3] or rows[t][4]) for t in settled)


def test_planned_capacity_leaves_the_trajectory_alone(oracle):
    """With room from the start, nothing grows and nothing is refused."""
    scene = encode(cube_drop(120, seed=2), default_config())
    world = oracle.world(scene)
    rows = trace(world, 80)
    assert world.capacities()[2] == 0
    assert all(not (r[3] or r[4]) for r in rows)
