"""The reference's world tests of contact identity, warm starting and resting crowds, restated and
run on both step engines (the oracle, and with -m gpu the device).

Sources: crates/dynamis-world/tests/suite/dynamics.rs:557-637 (a resting contact carries its
warm-start impulse; a settled stack drifts nothing), contact_identity.rs:115-200 (manifold points
hold distinct features; a sparse-iteration stack holds through contact identity),
broadphase.rs:60-79 (a wide static floor carries a dense crowd). The helpers restate
tests/suite/common.rs (settle, settle_until polling every 4 frames, asleep = ACTIVE counter 0) and
World::inspect_contacts (crates/dynamis-world/src/readback.rs:95-139: the active contacts, then
the resting archive's manifolds of pairs not already listed).
"""

from __future__ import annotations

import numpy as np

from tests.worlds import DT, BodyDesc, encode, gravity_config

SETTLE_POLL = 4


def settle(world, frames):
    for _ in range(frames):
        world.step(DT)


def settle_until(world, limit, settled):
    for frame in range(1, limit + 1):
        world.step(DT)
        if frame % SETTLE_POLL:
            continue
        if settled(world):
            return frame
    assert settled(world), f"world must settle within {limit} frames"
    return limit


def asleep(world) -> bool:
    return world.measured()["ACTIVE"] == 0


def manifolds(world):
    """World::inspect_contacts: active contacts, then resting ones of pairs not yet seen."""
    out, seen = [], set()
    for c in list(world.inspect_contacts()) + list(world.resting_contacts()):
        key = (int(c["a"]), int(c["b"]))
        if key in seen:
            continue
        seen.add(key)
        out.append(c)
    return out


def wide_floor(friction):
    # contact_identity.rs:49-55
    return BodyDesc.cuboid((16.0, 0.5, 16.0), friction=friction, position=(0.0, -0.5, 0.0), mass=0.0)


def test_resting_contact_carries_warm_start_impulse(engine):
    # dynamics.rs:557-595
    scene = encode([BodyDesc.sphere(5.0, position=(0.0, -1.0, 0.0), mass=0.0),
                    BodyDesc.sphere(0.5, position=(0.0, 4.5, 0.0))], gravity_config())
    world = engine(scene)
    settle_until(world, 60, asleep)
    resting = manifolds(world)
    assert len(resting) == 1, "resting ball must hold exactly one contact"
    impulse = float(resting[0]["points"][0]["accumulated_normal"])
    assert impulse > 0.01, impulse
    n = int(resting[0]["point_count"])
    assert np.all(np.isfinite(resting[0]["points"]["accumulated_normal"][:n]))
    assert np.all(resting[0]["points"]["accumulated_normal"][:n] >= 0.0)
    pair = (int(resting[0]["a"]), int(resting[0]["b"]))
    world.step(DT)
    relayed = manifolds(world)
    assert len(relayed) == 1, "the archived contact must be matched again"
    assert (int(relayed[0]["a"]), int(relayed[0]["b"])) == pair
    assert float(relayed[0]["points"][0]["accumulated_normal"]) > 0.01


def test_settled_stack_drifts_nothing_across_frames(engine):
    # dynamics.rs:597-637
    bodies = [BodyDesc.sphere(5.0, position=(0.0, -1.0, 0.0), mass=0.0)]
    bodies += [BodyDesc.sphere(0.4, position=(0.0, 4.7 + i * 0.81, 0.0), restitution=0.0) for i in range(3)]
    scene = encode(bodies, gravity_config())
    world = engine(scene)
    rows = [int(scene.rows[h]) for h in range(1, 4)]
    previous = [world.states()[r]["position"].copy() for r in rows]

    def still(w):
        nonlocal previous
        current = [w.states()[r]["position"].copy() for r in rows]
        done = all(np.linalg.norm(a.astype(np.float64) - b) < 1e-5 for a, b in zip(current, previous))
        previous = current
        return done

    settle_until(world, 180, still)
    baseline = [float(world.states()[r]["position"][1]) for r in rows]
    settle(world, 30)
    for i, r in enumerate(rows):
        drift = abs(float(world.states()[r]["position"][1]) - baseline[i])
        assert drift < 1e-3, f"settled stack body {i} drifted {drift}"


def test_manifold_points_hold_distinct_features(engine):
    # contact_identity.rs:115-162: a box, a hull of the unit cube, a cylinder, a capsule and a
    # falling sphere on a floor; every manifold names each of its points by a distinct feature
    from dynamis_amd.world import build_hull
    cube = np.array([[x, y, z] for x in (-0.5, 0.5) for y in (-0.5, 0.5) for z in (-0.5, 0.5)], np.float32)
    hulls = [build_hull(cube)]
    bodies = [wide_floor(0.8),
              BodyDesc.cuboid((0.5, 0.5, 0.5), position=(0.0, 0.5, 0.0), friction=0.8),
              BodyDesc.hull(0, position=(1.6, 0.5, 0.0), friction=0.8),
              BodyDesc.cylinder(0.5, 0.5, position=(-1.6, 0.75, 0.0), friction=0.8),
              BodyDesc.capsule(0.4, 0.4, position=(0.0, 0.75, 1.6), friction=0.8),
              BodyDesc.sphere(0.5, position=(3.2, 2.5, 0.0), friction=0.8)]
    world = engine(encode(bodies, gravity_config(), hulls))
    checked = 0
    for frame in range(120):
        world.step(DT)
        if frame % 10:
            continue
        for c in manifolds(world):
            held = [int(f) for f in c["points"]["feature"][: int(c["point_count"])]]
            assert len(set(held)) == len(held), (frame, held)
            checked += 1
    assert checked >= 40


def test_sparse_iteration_stack_holds_through_contact_identity(engine):
    # contact_identity.rs:164-200
    cfg = gravity_config(solve_iterations=2, position_iterations=4)
    bodies = [wide_floor(0.7)]
    bodies += [BodyDesc.cuboid((0.5, 0.5, 0.5), position=(0.0, 0.5 + level * 1.005, 0.0), friction=0.7)
               for level in range(6)]
    scene = encode(bodies, cfg)
    world = engine(scene)
    settle(world, 300)
    states = world.states()
    top = states[int(scene.rows[6])]
    assert abs(float(top["position"][1]) - 5.5) < 0.15, float(top["position"][1])
    for h in range(1, 7):
        p = states[int(scene.rows[h])]["position"]
        assert abs(float(p[0])) < 0.1 and abs(float(p[2])) < 0.1, p


def test_wide_static_floor_carries_a_dense_crowd(engine):
    # broadphase.rs:55-79: 64 spheres of radius 0.4 dropped from 1.2 m onto a 60 m floor rest at
    # y = 0.4 (within 0.25) once all of them are in the resting archive
    crowd = 64
    bodies = [BodyDesc.cuboid((30.0, 0.5, 30.0), mass=0.0, position=(0.0, -0.5, 0.0))]
    for i in range(crowd):
        bodies.append(BodyDesc.sphere(0.4, position=((i % 8) * 2.0 - 7.0, 1.2, (i // 8) * 3.0 - 10.0)))
    scene = encode(bodies, gravity_config())
    world = engine(scene)
    settle_until(world, 240, lambda w: w.measured()["RESTING"] >= crowd)
    states = world.states()
    for s in states[: scene.dynamic_count]:
        assert abs(float(s["position"][1]) - 0.4) < 0.25, float(s["position"][1])
