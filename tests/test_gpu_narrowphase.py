"""dyn_narrowphase on the device against the oracle's narrowphase pass: the same records and
candidate pairs give bit-identical contact_valid words and contacts_raw slots (every field of
every contact). Cases: a 60-body mixed pile (spheres, boxes, capsules, hulls, cylinders) a step
into its contacts; a 3000-body mixed pile with 64 hull sources stepped on the device, over that
step's broadphase pairs; an empty pair list; pairs naming colliders outside the records."""

from __future__ import annotations

import numpy as np
import pytest

from dynamis_amd.narrowphase import Narrowphase, set_epa_small_vertices
from dynamis_amd.world import World
from tests.narrow_cases import DT, convex_path_contacts, pile_scene, with_states

pytestmark = pytest.mark.gpu


def device_stage(scene, pairs):
    np_ = Narrowphase(scene, max(len(pairs), 1))
    np_.run(pairs)
    return np_.read()


def assert_stage_parity(oracle, scene, pairs, least_contacts):
    ref = oracle.world(scene, pair_capacity=max(len(pairs), 1))
    want_raw, want_valid = ref.narrowphase(pairs)
    ref.close()
    raw, valid = device_stage(scene, pairs)
    assert np.array_equal(valid, want_valid)
    hit = want_valid == 1
    assert hit.sum() >= least_contacts
    assert convex_path_contacts(scene, want_raw[hit]) > 0
    assert raw[hit].tobytes() == want_raw[hit].tobytes()


def test_small_pile_bit_identical(device, oracle):
    scene = pile_scene(60)
    ref = oracle.world(scene)
    for _ in range(40):
        ref.step(DT)
    start = ref.states()
    ref.step(DT)
    pairs = ref.pairs()
    ref.close()
    assert_stage_parity(oracle, with_states(scene, start), pairs, 20)


def test_large_pile_bit_identical(device, oracle):
    scene = pile_scene(3000, seed=5, hull_sources=64, spacing=1.5)
    world = World()
    world.create(scene)
    for _ in range(60):
        world.step(DT)
    states = world.states()
    pairs = world.pairs()
    world.close()
    assert len(pairs) > 5000
    assert_stage_parity(oracle, with_states(scene, states), pairs, 1000)


def test_empty_and_foreign_pairs(device):
    scene = pile_scene(8)
    raw, valid = device_stage(scene, np.zeros((0, 2), np.uint32))
    assert len(raw) == len(valid) == 0
    k = len(scene.colliders)
    foreign = np.array([[0, k], [k + 5, 1], [1, 2]], np.uint32)
    _, valid = device_stage(scene, foreign)
    assert valid[0] == 0 and valid[1] == 0


def hull_pile(count: int, sources: int = 32, spacing: float = 1.4):
    """Ellipsoid hulls alone on a lattice above a plane."""
    from tests.worlds import BodyDesc, config, ellipsoid_hulls, encode
    side = int(np.ceil(np.sqrt(count / 4.0)))
    bodies = [BodyDesc.plane()]
    for i in range(count):
        layer, cell = divmod(i, side * side)
        gx, gz = divmod(cell, side)
        bodies.append(BodyDesc.hull(i % sources, position=((gx - side / 2) * spacing, 1.5 + layer * spacing,
                                                          (gz - side / 2) * spacing), density=1000.0))
    return encode(bodies, config(), ellipsoid_hulls(sources, seed=11))


def test_support_cells_change_no_bit(device):
    """The hull support over its support cells (dyn_hull_support_cells) against the linear scan
    over every vertex, on the pass over a settled 3000-body pile with 64 hull sources and on the
    hull-heavy pairs of a pile of hulls alone: every byte of every contact equal."""
    for scene, steps in ((pile_scene(3000, seed=5, hull_sources=64, spacing=1.5), 90), (hull_pile(1500), 60)):
        world = World()
        world.create(scene)
        for _ in range(steps):
            world.step(DT)
        states = world.states()
        pairs = world.pairs()
        world.close()
        staged = with_states(scene, states)
        tables = Narrowphase(staged, max(len(pairs), 1))
        linear = Narrowphase(staged, max(len(pairs), 1), support_tables=False)
        assert tables.cells is not None and linear.cells is None
        tables.run(pairs)
        linear.run(pairs)
        raw_t, valid_t = tables.read()
        raw_l, valid_l = linear.read()
        assert valid_t.tobytes() == valid_l.tobytes()
        assert raw_t[valid_t == 1].tobytes() == raw_l[valid_l == 1].tobytes()
        assert (valid_t == 1).sum() > 500


def test_epa_fallback_changes_no_bit(device, oracle):
    """EPA's two kernels (eight lanes per polytope up to 16 vertices, else a wave per polytope
    from the start): with the eight-lane kernel's capacity at 4 vertices every polytope that
    expands goes to the wave kernel, and the pass still equals the oracle's, byte for byte, on a
    3000-body mixed pile and on a pile of hulls alone."""
    for scene, steps in ((pile_scene(3000, seed=5, hull_sources=64, spacing=1.5), 60), (hull_pile(1500), 60)):
        world = World()
        world.create(scene)
        for _ in range(steps):
            world.step(DT)
        states = world.states()
        pairs = world.pairs()
        world.close()
        staged = with_states(scene, states)
        try:
            set_epa_small_vertices(4)
            assert_stage_parity(oracle, staged, pairs, 500)
        finally:
            set_epa_small_vertices(16)
        assert_stage_parity(oracle, staged, pairs, 500)
