"""Body commands between steps (set_velocity, set_position, forces, impulses, sleep, wake) on both
step engines: the reference's command tests restated, closed forms of each edit kind, and (with
-m gpu) the device equal to the oracle bit for bit over a rollout with commands interleaved.

Sources: crates/dynamis-world/tests/suite/commands.rs:51-69 (ordered commands fold in order),
dynamics.rs:523-555 (wake of an unconnected body; a driven contact island stays awake),
contact_identity.rs:57-92 (sliding contacts hold their features); the edit semantics of
crates/dynamis-rigid/shaders/body_edits.wgsl.
"""

from __future__ import annotations

import numpy as np
import pytest

from tests.test_step_identity import asleep, manifolds, settle, settle_until, wide_floor
from tests.worlds import DT, BodyDesc, encode, gravity_config, static_config


def test_ordered_commands_on_one_slot_fold_in_order(engine):
    # commands.rs:51-69: the edits of one body apply in order; the force acts through the step
    scene = encode([BodyDesc.sphere(0.5, mass=2.0)], static_config())
    world = engine(scene)
    world.set_velocity(0, (1.0, 0.0, 0.0))
    world.apply_force(0, (4.0, 0.0, 0.0))
    world.apply_impulse(0, (2.0, 0.0, 0.0))
    world.set_velocity(0, (3.0, 0.0, 0.0))
    world.step(DT)
    state = world.states()[0]
    expected = 3.0 + 4.0 / 2.0 * float(DT)
    assert abs(float(state["velocity"][0]) - expected) < 1e-5, state["velocity"]
    assert float(state["position"][0]) > 0.0


def test_edit_kinds_match_closed_forms(engine):
    """Each kind on a free body without gravity or damping, one step of dt with 2 substeps."""
    mass, r = 3.0, 0.5
    bodies = [BodyDesc.sphere(r, mass=mass, position=(10.0 * i, 0.0, 0.0)) for i in range(6)]
    scene = encode(bodies, static_config())
    inv_i = float(scene.descs["inverse_inertia"][0][0])  # a sphere's inertia is isotropic
    world = engine(scene)
    dt = float(DT)
    world.apply_impulse(0, (0.0, 6.0, 0.0))                                    # v = J / m
    world.apply_force(1, (0.0, 0.0, 9.0))                                      # v = F / m dt
    world.apply_angular_impulse(2, (0.0, 0.0, 1.0))                            # w = I^-1 L
    world.apply_impulse_at_point(3, (0.0, 3.0, 0.0), (30.0 + r, 0.0, 0.0))     # v = J / m, w = I^-1 (r x J)
    world.set_position(4, (40.0, 2.0, 0.0))                                    # teleport
    world.set_angular_velocity(5, (0.0, 2.0, 0.0))
    world.step(DT)
    s = world.states()
    rows = [int(scene.rows[h]) for h in range(6)]
    np.testing.assert_allclose(s[rows[0]]["velocity"], (0.0, 2.0, 0.0), atol=1e-6)
    np.testing.assert_allclose(s[rows[1]]["velocity"], (0.0, 0.0, 3.0 * dt), atol=1e-6)
    np.testing.assert_allclose(s[rows[2]]["angular_velocity"], (0.0, 0.0, inv_i), rtol=1e-6)
    np.testing.assert_allclose(s[rows[3]]["velocity"], (0.0, 1.0, 0.0), atol=1e-6)
    np.testing.assert_allclose(s[rows[3]]["angular_velocity"], (0.0, 0.0, r * 3.0 * inv_i), rtol=1e-6)
    np.testing.assert_allclose(s[rows[4]]["position"], (40.0, 2.0, 0.0), atol=1e-6)
    np.testing.assert_allclose(s[rows[5]]["angular_velocity"], (0.0, 2.0, 0.0), atol=1e-6)


def test_sleep_freezes_and_wake_releases(engine):
    scene = encode([BodyDesc.sphere(0.5, velocity=(1.0, 0.0, 0.0))], static_config())
    world = engine(scene)
    world.step(DT)
    world.sleep(0)
    world.step(DT)
    s = world.states()[0]
    assert s["sleeping"] == 1 and not np.any(s["velocity"])
    assert world.measured()["SLEPT"] == 1
    x = float(s["position"][0])
    world.step(DT)
    assert float(world.states()[0]["position"][0]) == x
    world.wake(0)
    world.set_velocity(0, (0.0, 0.0, 2.0))
    world.step(DT)
    s = world.states()[0]
    assert s["sleeping"] == 0 and world.measured()["WOKE"] == 1
    assert float(s["position"][2]) > 0.0


def test_unknown_body_is_refused(engine):
    world = engine(encode([BodyDesc.sphere(0.5)], static_config()))
    with pytest.raises(Exception):
        world.set_velocity(7, (1.0, 0.0, 0.0))


def test_distant_idle_bodies_sleep_independently(engine):
    # dynamics.rs:523-534
    scene = encode([BodyDesc.sphere(0.3), BodyDesc.sphere(0.3, position=(50.0, 0.0, 0.0))], static_config())
    world = engine(scene)
    settle_until(world, 50, asleep)
    s = world.states()
    assert s[scene.rows[0]]["sleeping"] and s[scene.rows[1]]["sleeping"]
    world.wake(0)
    settle(world, 3)
    s = world.states()
    assert not s[scene.rows[0]]["sleeping"] and s[scene.rows[1]]["sleeping"]


def test_driven_contact_island_stays_awake(engine):
    # dynamics.rs:536-555
    scene = encode([BodyDesc.sphere(5.0, position=(0.0, -1.0, 0.0), mass=0.0),
                    BodyDesc.sphere(0.5, position=(0.0, 4.5, 0.0)),
                    BodyDesc.sphere(0.5, position=(0.0, 5.3, 0.0))], gravity_config())
    world = engine(scene)
    lower, upper = int(scene.rows[1]), int(scene.rows[2])
    settle_until(world, 50, asleep)
    s = world.states()
    assert s[lower]["sleeping"] and s[upper]["sleeping"]
    for _ in range(20):
        world.apply_force(2, (3.0, 0.0, 0.0))
        world.step(DT)
    s = world.states()
    assert not s[lower]["sleeping"] and not s[upper]["sleeping"], "driven contact island must stay awake"
    assert float(s[upper]["position"][0]) > 0.05


def features_of(world, a_body, b_body):
    feats = []
    for c in manifolds(world):
        if {int(c["first_body_id"]), int(c["second_body_id"])} == {a_body, b_body}:
            feats += [int(f) for f in c["points"]["feature"][: int(c["point_count"])]]
    return sorted(feats)


def test_sliding_contacts_hold_their_features(engine):
    # contact_identity.rs:57-92: a frictionless box slid across a frictionless floor keeps the
    # features of its resting manifold every frame
    bodies = [wide_floor(0.0), BodyDesc.cuboid((0.5, 0.5, 0.5), friction=0.0, position=(0.0, 0.5, 0.0))]
    scene = encode(bodies, gravity_config())
    world = engine(scene)
    cube = int(scene.rows[1])
    settle(world, 30)
    resting = features_of(world, 0, 1)
    assert resting, "a resting box must hold contact points"
    world.set_velocity(1, (3.0, 0.0, 0.0))
    travel = 0.0
    previous = float(world.states()[cube]["position"][0])
    for frame in range(40):
        world.step(DT)
        position = float(world.states()[cube]["position"][0])
        travel += position - previous
        previous = position
        assert features_of(world, 0, 1) == resting, frame
    assert travel > 1.5, travel


@pytest.mark.gpu
def test_commands_keep_device_and_oracle_bit_identical(device, oracle):
    """A box stack and a few spheres stepped 90 frames with every edit kind queued along the way
    (several per body per step at times): the device equals the oracle bit for bit."""
    from dynamis_amd.world import World
    bodies = [wide_floor(0.6)]
    bodies += [BodyDesc.cuboid((0.4, 0.4, 0.4), position=(0.0, 0.45 + 0.85 * i, 0.0)) for i in range(4)]
    bodies += [BodyDesc.sphere(0.3, position=(1.5 + i, 1.0, 0.5 * i)) for i in range(4)]
    scene = encode(bodies, gravity_config())
    ref = oracle.world(scene)
    dev = World()
    dev.create(scene)
    for frame in range(90):
        for w in (ref, dev):
            r = np.random.default_rng(frame)
            if frame % 7 == 0:
                w.apply_impulse(int(r.integers(1, 9)), tuple(r.normal(size=3) * 2.0))
            if frame % 11 == 3:
                h = int(r.integers(1, 9))
                w.apply_force_at_point(h, tuple(r.normal(size=3) * 30.0), tuple(r.normal(size=3)))
                w.apply_torque(h, tuple(r.normal(size=3)))
            if frame % 13 == 5:
                w.set_velocity(int(r.integers(5, 9)), tuple(r.normal(size=3)))
                w.apply_angular_impulse(int(r.integers(1, 5)), tuple(r.normal(size=3) * 0.5))
            if frame == 40:
                w.sleep(6)
                w.set_position(7, (3.0, 2.0, -1.0))
                w.set_orientation(7, (0.0, 0.38268343, 0.0, 0.9238795))
            if frame == 60:
                w.wake(6)
                w.apply_impulse_at_point(6, (1.0, 2.0, 0.0), (2.5, 1.0, 1.0))
        ref.step(DT)
        dev.step(DT)
        if frame % 15 == 14:
            assert dev.states().tobytes() == ref.states().tobytes(), frame
            assert dev.inspect_contacts().tobytes() == ref.inspect_contacts().tobytes(), frame
            md, mr = dev.measured(), ref.measured()
            assert {k: md[k] for k in ("SLEPT", "WOKE", "CONTACTS")} == {k: mr[k] for k in ("SLEPT", "WOKE", "CONTACTS")}
    dev.close()
