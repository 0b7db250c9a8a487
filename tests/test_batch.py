"""Batches of independent worlds (dyn_batch_create): the merged record layout on the CPU, and
on the device every world of a batch stepping bit-identically to the same world alone."""

from __future__ import annotations

import numpy as np
import pytest

from dynamis_amd import abi
from dynamis_amd.world import Batch, World, merge_scenes
from tests.worlds import DT, BodyDesc, config, cube_drop, ellipsoid_hulls, encode, mixed_pile, tower


def batch_scenes():
    hulls = ellipsoid_hulls(4, seed=11)
    return [
        encode(tower(6), config()),
        encode(cube_drop(40, seed=3, footprint=4.0, low=1.0, high=6.0), config()),
        encode(mixed_pile(30, seed=4, spacing=1.5, kinds=5, hull_sources=4), config(), hulls),
        encode([BodyDesc.plane(), BodyDesc.sphere(0.5, position=(0.0, 1.0, 0.0))], config()),
        # the second world again, in the same place: the two must not see each other
        encode(cube_drop(40, seed=3, footprint=4.0, low=1.0, high=6.0), config()),
    ]


def test_merge_layout_keeps_every_world_in_order():
    scenes = batch_scenes()
    merged, layout = merge_scenes(scenes)
    total = sum(len(s.states) for s in scenes)
    assert len(merged.states) == total
    assert merged.dynamic_count == sum(s.dynamic_count for s in scenes)
    # simulated rows first, world by world
    simulated = (merged.descs["inverse_mass"] > 0) & ((merged.descs["flags"] & abi.BODY_KINEMATIC) == 0)
    assert simulated[:merged.dynamic_count].all() and not simulated[merged.dynamic_count:].any()
    seen = np.zeros(total, dtype=int)
    for i, s in enumerate(scenes):
        rows = layout.rows_of(i)
        seen[rows] += 1
        got = merged.states[rows]
        assert np.array_equal(got["body_id"] - layout.id_base[i], s.states["body_id"])
        assert np.array_equal(got["position"], s.states["position"])
        assert merged.descs[rows].tobytes() == s.descs.tobytes()
        c0, c1 = int(layout.collider_base[i]), int(layout.collider_base[i] + layout.collider_count[i])
        assert np.array_equal(merged.owners[c0:c1], rows[s.owners])
        assert (layout.collider_worlds[c0:c1] == i).all()
        hull = s.colliders["kind"] == abi.SHAPE_HULL
        if hull.any():
            got_src = merged.sources[merged.colliders["source"][c0:c1][hull]]
            want_src = s.sources[s.colliders["source"][hull]]
            assert np.array_equal(got_src["vertex_count"], want_src["vertex_count"])
    assert (seen == 1).all()
    assert len(np.unique(merged.states["body_id"])) == total


@pytest.mark.gpu
def test_batched_worlds_step_as_they_do_alone(device):
    scenes = batch_scenes()
    print("[batch] 5 worlds, 60 frames", flush=True)
    batch = Batch(scenes)
    alone = []
    for s in scenes:
        w = World()
        w.create(s)
        alone.append(w)
    for frame in range(60):
        batch.step(DT)
        for w in alone:
            w.step(DT)
        if frame % 15 == 14:
            states, contacts = batch.states(), batch.inspect_contacts()
            for i, w in enumerate(alone):
                assert batch.world_states(i, states).tobytes() == w.states().tobytes(), (frame, i)
                assert batch.world_contacts(i, contacts).tobytes() == w.inspect_contacts().tobytes(), (frame, i)
            total = {k: sum(w.measured()[k] for w in alone) for k in ("PAIRS", "CONTACTS", "ISLANDS", "BLOCKS")}
            got = batch.measured()
            assert {k: got[k] for k in total} == total, frame
    for w in alone:
        w.close()
    batch.close()
