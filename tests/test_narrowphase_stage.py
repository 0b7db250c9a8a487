"""The oracle's narrowphase pass alone (oracle_world_narrowphase, the checker of dyn_narrowphase)
against the oracle's whole step: over a step's candidate pairs and the states the step started
from, it writes the contacts that step compacted, in every field the pass writes."""

from __future__ import annotations

import numpy as np
import pytest

from tests.narrow_cases import DT, convex_path_contacts, narrow_fields_equal, pile_scene, with_states


def test_stage_reproduces_the_steps_contacts(oracle):
    scene = pile_scene(60)
    ref = oracle.world(scene)
    for _ in range(40):
        ref.step(DT)
    start = ref.states()
    ref.step(DT)
    pairs, contacts = ref.pairs(), ref.inspect_contacts()
    stage = oracle.world(with_states(scene, start))
    raw, valid = stage.narrowphase(pairs)
    assert set(np.unique(valid)) <= {0, 1}
    got = raw[valid == 1]  # compaction keeps pair order (compact_scatter.wgsl)
    assert len(got) == len(contacts) > 20
    assert convex_path_contacts(scene, contacts) > 0
    assert narrow_fields_equal(got, contacts) == []
    stage.close()
    ref.close()


def test_stage_refuses_pairs_outside_the_records(oracle):
    scene = pile_scene(8)
    w = oracle.world(scene)
    with pytest.raises(ValueError):
        w.narrowphase(np.array([[0, len(scene.colliders)]], np.uint32))
    raw, valid = w.narrowphase(np.zeros((0, 2), np.uint32))
    assert len(raw) == len(valid) == 0
    w.close()
