"""The reference's own world tests, restated, run on both step engines: closed forms (free
fall, isotropic spin), invariants (angular momentum, kinetic energy), resting and stacking
outcomes, solver row counts and determinism. On the oracle (oracle/step.c) they pin the CPU
restatement; on the device (the HIP library through dyn_world_*) they check the product's
behaviour directly, beside the parity tests that compare the two.

Sources: crates/dynamis-world/tests/suite/dynamics.rs, rotation.rs, solver.rs, determinism.rs,
plane.rs.
"""

import math

import numpy as np
import pytest

from tests.worlds import (DT, BodyDesc, angle_between, body_momentum, body_spin, config,
                          cube_drop, encode, floor_box, gravity_config, kinetic_energy, qrot,
                          static_config, symplectic_fall, tower)


def run(engine, bodies, cfg, frames, capacity=None):
    scene = encode(bodies, cfg)
    world = engine(scene, capacity)
    for _ in range(frames):
        world.step(DT)
    return scene, world


def test_free_fall_matches_closed_form(engine):
    # dynamics.rs:11-30
    scene, world = run(engine, [BodyDesc.sphere(0.1, position=(0, 10, 0), mass=1.0)],
                       gravity_config(), 30)
    state = world.states()[0]
    expected_y = 10.0 + symplectic_fall(9.81, 30, 0.0, 2)
    assert abs(state["position"][1] - expected_y) < 1e-3
    assert abs(state["velocity"][1] - (-9.81 * float(DT) * 30)) < 1e-3


def test_free_fall_with_initial_speed_matches_closed_form(engine):
    # the closed form with an initial speed, three substeps
    scene, world = run(engine, [BodyDesc.sphere(0.2, position=(0, 50, 0), velocity=(0, 3, 0))],
                       gravity_config(substeps=3), 20)
    state = world.states()[0]
    assert abs(state["position"][1] - (50.0 + symplectic_fall(9.81, 20, 3.0, 3))) < 1e-3
    assert abs(state["velocity"][1] - (3.0 - 9.81 * float(DT) * 20)) < 1e-3


def test_isotropic_spin_is_exactly_untouched(engine):
    # rotation.rs:164-193
    spin = np.array([1.0, 1.0, 1.0], np.float32)
    scene = encode([BodyDesc.sphere(0.5, angular_velocity=tuple(spin))], static_config())
    world = engine(scene)
    axis = spin / np.linalg.norm(spin)
    for frame in range(1, 601):
        world.step(DT)
        if frame % 30:
            continue
        state = world.states()[0]
        assert np.array_equal(state["angular_velocity"], spin), frame
        half = 0.5 * float(np.linalg.norm(spin)) * float(DT) * frame
        analytic = np.concatenate([axis * math.sin(half), [math.cos(half)]])
        assert angle_between(qrot(state["orientation"], [0, 0, 1]), qrot(analytic, [0, 0, 1])) < 1e-2


def spin_scene(spin):
    return encode([BodyDesc.cuboid((0.5, 1.0, 2.0), mass=1.0, angular_velocity=tuple(spin))],
                  static_config())


@pytest.mark.parametrize("spin,energy", [((math.sqrt(0.5), math.sqrt(0.5), 0.0), False),
                                         ((0.05, 1.0, 0.0), True)])
def test_free_rotation_conserves_angular_momentum(engine, spin, energy):
    # rotation.rs:84-137
    scene = spin_scene(spin)
    inertia = scene.descs[0]["inertia"]
    world = engine(scene)
    initial = world.states()[0]
    momentum = body_momentum(initial, inertia)
    e0 = kinetic_energy(initial, inertia)
    turned = 0.0
    for frame in range(1, 601):
        world.step(DT)
        if frame % 30:
            continue
        state = world.states()[0]
        measured = body_momentum(state, inertia)
        drift = abs(np.linalg.norm(measured) - np.linalg.norm(momentum)) / np.linalg.norm(momentum)
        assert drift < 1e-3, frame
        assert angle_between(measured, momentum) < 3e-2, frame
        if energy:
            assert abs(kinetic_energy(state, inertia) - e0) / e0 < 1e-3, frame
        turned = max(turned, angle_between(body_spin(state), body_spin(initial)))
    assert turned > (math.pi / 2 if energy else 1.0)


def test_principal_axis_spin_stays_on_its_axis(engine):
    # rotation.rs:140-161
    scene = spin_scene((0.0, 0.0, 1.0))
    world = engine(scene)
    for frame in range(1, 601):
        world.step(DT)
        if frame % 30:
            continue
        state = world.states()[0]
        assert np.array_equal(state["angular_velocity"], np.float32([0, 0, 1])), frame
        assert abs(state["orientation"][0]) < 1e-6 and abs(state["orientation"][1]) < 1e-6


def test_ball_rests_on_ground_and_sleeps(engine):
    # dynamics.rs:63-79
    scene, world = run(engine, [BodyDesc.sphere(1.0, mass=0.0), BodyDesc.sphere(0.5, position=(0, 3, 0))],
                       gravity_config(), 90)
    row = scene.rows[1]
    state = world.states()[row]
    assert abs(state["position"][1] - 1.5) < 0.05
    assert state["sleeping"] == 1
    rest = state["position"][1]
    for _ in range(30):
        world.step(DT)
    assert abs(world.states()[row]["position"][1] - rest) < 1e-3


def test_overlapping_bodies_separate(engine):
    # dynamics.rs:82-89
    scene, world = run(engine, [BodyDesc.sphere(0.5, position=(0, -0.4, 0)),
                                BodyDesc.sphere(0.5, position=(0, 0.4, 0))], static_config(), 16)
    st = world.states()
    assert st[scene.rows[1]]["position"][1] - st[scene.rows[0]]["position"][1] >= 0.99


def test_restitution_bounces_fast_and_calms_slow_impact(engine):
    # dynamics.rs:92-122
    scene, world = run(engine, [BodyDesc.sphere(1.0, mass=0.0),
                                BodyDesc.sphere(0.5, position=(0, 1.5, 0), velocity=(0, -2, 0),
                                                restitution=0.8)], static_config(), 1)
    bounce = world.states()[scene.rows[1]]
    assert bounce["velocity"][1] > 1.2 and bounce["position"][1] > 1.4
    scene, world = run(engine, [BodyDesc.sphere(1.0, mass=0.0),
                                BodyDesc.sphere(0.5, position=(0, 1.5, 0), velocity=(0, -0.5, 0),
                                                restitution=0.9)], static_config(), 1)
    v = world.states()[scene.rows[1]]["velocity"][1]
    assert -0.05 < v < 0.01


def test_frictionless_ball_keeps_sliding(engine):
    # dynamics.rs:125-141 (first half: friction 0)
    scene, world = run(engine, [BodyDesc.sphere(100.0, mass=0.0),
                                BodyDesc.sphere(0.5, position=(0, 100.5, 0), velocity=(2, 0, 0),
                                                friction=0.0)], gravity_config(), 30)
    s = world.states()[scene.rows[1]]
    assert s["velocity"][0] > 1.9
    assert abs(s["angular_velocity"][2]) < 0.01


def test_gripping_ball_rolls_without_slipping(engine):
    # dynamics.rs:143-161 (second half: friction 0.8 from the start)
    scene = encode([BodyDesc.sphere(100.0, mass=0.0),
                    BodyDesc.sphere(0.5, position=(0, 100.5, 0), velocity=(2, 0, 0), friction=0.8)],
                   gravity_config())
    world = engine(scene)
    row = scene.rows[1]
    for _ in range(600):
        world.step(DT)
        s = world.states()[row]
        if s["angular_velocity"][2] < -1.5 and abs(s["velocity"][0] + s["angular_velocity"][2] * 0.5) < 0.05:
            break
    assert s["angular_velocity"][2] < -1.5
    assert abs(s["velocity"][0] + s["angular_velocity"][2] * 0.5) < 0.05


def test_stacked_spheres_do_not_collapse(engine):
    # dynamics.rs:164-174
    scene, world = run(engine, [BodyDesc.sphere(1.0, mass=0.0),
                                BodyDesc.sphere(0.5, position=(0, 1.5, 0)),
                                BodyDesc.sphere(0.5, position=(0, 2.5, 0))], gravity_config(), 120)
    st = world.states()
    lower, upper = st[scene.rows[1]]["position"][1], st[scene.rows[2]]["position"][1]
    assert 1.4 < lower < 1.52 and 2.4 < upper < 2.6


def test_resting_tower_holds_its_spacing(engine):
    # dynamics.rs:177-213
    scene, world = run(engine, tower(16), config(), 240)
    st = world.states()
    for level in range(16):
        p = st[scene.rows[level + 1]]["position"]
        assert abs(p[1] - (0.5 + level)) <= 0.02 + level * 0.005, (level, p)
        assert abs(p[0]) < 0.2 and abs(p[2]) < 0.2


def test_dropped_box_settles_within_contact_margin(engine):
    # dynamics.rs:215-240
    scene, world = run(engine, [floor_box(), BodyDesc.cuboid((0.5, 0.5, 0.5), position=(0, 2.5, 0),
                                                             friction=0.7)], config(), 240)
    overlap = 0.5 - world.states()[scene.rows[1]]["position"][1]
    assert overlap <= 0.02 + 0.005


def test_box_rests_on_plane_with_its_face(engine):
    # plane.rs:175-207: a cuboid meets a plane with its supporting face (four points)
    scene, world = run(engine, [BodyDesc.plane(), BodyDesc.cuboid((0.5, 0.5, 0.5), position=(0, 0.49, 0))],
                       config(), 1)
    contacts = world.inspect_contacts()
    assert len(contacts) == 1
    assert contacts[0]["point_count"] == 4
    assert np.allclose(np.abs(contacts[0]["normal"]), [0, 1, 0])


def test_solver_claims_no_row_without_contacts(engine):
    # solver.rs:21-35
    bodies = [BodyDesc.cuboid((0.5, 0.5, 0.5), position=(i * 3.0, 60.0, 0.0)) for i in range(32)]
    _, world = run(engine, bodies, gravity_config(), 2)
    m = world.measured()
    assert m["LIVE"] >= 32 and m["SOLVER_ROWS"] == 0


def test_solver_claims_rows_of_its_blocks_alone(engine):
    # solver.rs:37-71 (first part)
    bodies = [BodyDesc.cuboid((20.0, 0.5, 20.0), mass=0.0, position=(0, -0.5, 0))]
    bodies += [BodyDesc.cuboid((0.5, 0.5, 0.5), position=(i * 3.0, 60.0, 0.0)) for i in range(32)]
    bodies.append(BodyDesc.cuboid((0.5, 0.5, 0.5), position=(0, 0.5, 0), friction=0.9))
    _, world = run(engine, bodies, gravity_config(solve_iterations=4, position_iterations=1), 1)
    m = world.measured()
    assert m["LIVE"] >= 33 and m["SOLVER_ROWS"] == 2


def test_solver_claims_each_shared_row_once(engine):
    # solver.rs:73-100
    bodies = [BodyDesc.cuboid((20.0, 0.5, 20.0), mass=0.0, position=(0, -0.5, 0))]
    bodies += [BodyDesc.cuboid((0.5, 0.5, 0.5), position=(i * 1.05 - 1.6, 0.5, 0.0), friction=0.9)
               for i in range(4)]
    scene, world = run(engine, bodies, gravity_config(), 1)
    assert world.measured()["SOLVER_ROWS"] == 5
    for _ in range(240):
        world.step(DT)
    assert world.measured()["SOLVER_ROWS"] == 0
    world.step(DT)
    assert world.measured()["SOLVER_ROWS"] == 0


def test_step_is_deterministic(engine):
    # determinism.rs:61: two worlds from the same records step bit-identically
    bodies = cube_drop(60, seed=3, footprint=6.0, low=1.0, high=8.0)
    outs = []
    for _ in range(2):
        _, world = run(engine, bodies, config(), 40)
        outs.append(world.states().tobytes())
    assert outs[0] == outs[1]


def test_cube_drop_comes_to_rest_above_the_plane(engine):
    # the 1,000-cube drop at a tenth of its size: every cube ends above the plane within the
    # contact margin and slop, and the pile's kinetic energy has drained
    bodies = cube_drop(100, seed=1, footprint=10.0, low=1.0, high=10.0)
    scene, world = run(engine, bodies, config(solve_iterations=4), 300)
    st = world.states()[:scene.dynamic_count]
    # the lowest point of each cube: centre minus the largest vertical extent of its corners
    lowest = []
    corners = np.array([[sx, sy, sz] for sx in (-0.5, 0.5) for sy in (-0.5, 0.5) for sz in (-0.5, 0.5)])
    for s in st:
        ys = [qrot(s["orientation"], c)[1] for c in corners]
        lowest.append(s["position"][1] + min(ys))
    assert min(lowest) > -(0.02 + 0.005) - 0.01
    speed = np.linalg.norm(st["velocity"], axis=1)
    assert float(np.max(speed)) < 0.5
    m = world.measured()
    assert m["ENTRY_FAULTS"] == 0 and m["REFUSED_PAIRS"] == 0 and m["REFUSED_CONTACTS"] == 0
