"""Parity of the device step (dyn_world_*) with the oracle's restatement of World::step.

Tolerances follow the stated FP32 bar for the per-step pipeline: from an identical start state
one step's body positions and orientations agree within 1e-5 relative (absolute below 1 m /
unit length), velocities and contact normal impulses within 1e-4 relative; candidate pairs,
contact identities, point counts and the discrete counters agree exactly. Long runs are
compared by penetration depth, kinetic energy and resting-contact statistics, because contact
dynamics is chaotic and a one-ulp difference in a transcendental grows over hundreds of steps.
"""

from __future__ import annotations

import numpy as np
import pytest

from dynamis_amd.world import World
from tests.worlds import (DT, BodyDesc, config, cube_drop, ellipsoid_hulls, encode, gravity_config,
                          mixed_pile, qrot, tower)

pytestmark = pytest.mark.gpu

POS_TOL = 1e-5
VEL_TOL = 1e-4
IMPULSE_TOL = 1e-4
DISCRETE = ["ENTRIES", "PAIRS", "GRID_LEVELS", "CONTACTS", "ARCHIVED", "RESTING", "SLEPT",
            "WOKE", "ISLANDS", "BLOCKS", "SOLVER_ROWS", "LIVE", "ENTRY_FAULTS", "REFUSED_PAIRS",
            "REFUSED_CONTACTS", "REFUSED_RESTING", "ISLAND_FAULTS", "ROW_FAULTS", "LIVE_FAULTS"]


def device_world(scene, capacity=None) -> World:
    world = World()
    world.create(scene, capacity)
    return world


def rel_close(got, want, tol, floor=1.0):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    scale = np.maximum(np.abs(want), floor)
    return np.abs(got - want) <= tol * scale


def assert_states_close(dev, ref, pos_tol=POS_TOL, vel_tol=VEL_TOL, label=""):
    for field, tol in (("position", pos_tol), ("orientation", pos_tol), ("velocity", vel_tol),
                       ("angular_velocity", vel_tol)):
        ok = rel_close(dev[field], ref[field], tol)
        if not ok.all():
            bad = np.argwhere(~ok)[0][0]
            raise AssertionError(f"{label} {field} row {bad}: device {dev[field][bad]} "
                                 f"oracle {ref[field][bad]}")
    assert np.array_equal(dev["sleeping"], ref["sleeping"]), label
    assert np.array_equal(dev["body_id"], ref["body_id"])


def assert_contacts_close(dev, ref, label=""):
    assert len(dev) == len(ref), f"{label} contact count {len(dev)} vs {len(ref)}"
    assert np.array_equal(dev["a"], ref["a"]) and np.array_equal(dev["b"], ref["b"]), label
    assert np.array_equal(dev["point_count"], ref["point_count"]), label
    assert rel_close(dev["normal"], ref["normal"], POS_TOL).all(), label
    for i in range(len(ref)):
        n = int(ref["point_count"][i])
        dp, rp = dev["points"][i][:n], ref["points"][i][:n]
        assert np.array_equal(dp["feature"], rp["feature"]), (label, i)
        assert rel_close(dp["position"], rp["position"], POS_TOL).all(), (label, i)
        assert rel_close(dp["depth"], rp["depth"], POS_TOL, floor=1e-2).all(), (label, i)
        impulse_scale = max(1e-3, float(np.max(np.abs(rp["accumulated_normal"]))))
        assert (np.abs(dp["accumulated_normal"] - rp["accumulated_normal"])
                <= IMPULSE_TOL * impulse_scale).all(), (label, i, dp["accumulated_normal"],
                                                        rp["accumulated_normal"])


def assert_counters_equal(dev, ref, label=""):
    md, mr = dev.measured(), ref.measured()
    diff = {k: (md[k], mr[k]) for k in DISCRETE if md[k] != mr[k]}
    assert not diff, f"{label} counters differ: {diff}"


def start_from(scene, states):
    """The same scene restarted from the given states (a fresh world: no archive)."""
    import copy
    s = copy.copy(scene)
    s.states = states.copy()
    return s


def single_step_cases():
    hulls = ellipsoid_hulls(12, seed=4)
    return {
        "tower16": (tower(16), config(), 0, None),
        "cube_drop_300_mid_fall": (cube_drop(300, seed=5, footprint=12.0, low=1.0, high=12.0),
                                   config(solve_iterations=4), 90, None),
        "mixed_pile_240": (mixed_pile(240, seed=2), config(), 60, None),
        "convex_pile_300": (mixed_pile(300, seed=6, spacing=1.6, kinds=5, hull_sources=12),
                            config(), 75, hulls),
    }


@pytest.mark.parametrize("name", list(single_step_cases()))
def test_single_step_from_identical_state(device, oracle, name):
    bodies, cfg, warm, hulls = single_step_cases()[name]
    print(f"[world parity] single step {name} after {warm} oracle steps", flush=True)
    scene = encode(bodies, cfg, hulls)
    pre = oracle.world(scene)
    for _ in range(warm):
        pre.step(DT)
    start = start_from(scene, pre.states())
    ref = oracle.world(start)
    dev = device_world(start)
    ref.step(DT)
    dev.step(DT)
    assert np.array_equal(dev.pairs(), ref.pairs()), "candidate pairs must be bit-exact"
    assert_counters_equal(dev, ref, name)
    assert_contacts_close(dev.inspect_contacts(), ref.inspect_contacts(), name)
    assert_states_close(dev.states(), ref.states(), label=name)


@pytest.mark.parametrize("name,frames", [("tower16", 120), ("box_stack_plane", 120),
                                         ("spheres_capsules", 90), ("hulls_cylinders", 90)])
def test_short_rollout_tracks_oracle(device, oracle, name, frames):
    if name == "tower16":
        bodies, cfg = tower(16), config()
    elif name == "box_stack_plane":
        bodies = [BodyDesc.plane()] + [BodyDesc.cuboid((0.5, 0.5, 0.5), position=(0.02 * i, 0.55 + i * 1.1, 0.0))
                                       for i in range(6)]
        cfg = config()
    elif name == "spheres_capsules":
        bodies = [BodyDesc.plane()]
        for i in range(12):
            pos = (0.3 * (i % 3), 1.0 + 1.3 * i, 0.2 * (i % 2))
            bodies.append(BodyDesc.sphere(0.4, position=pos) if i % 2 == 0
                          else BodyDesc.capsule(0.25, 0.5, position=pos))
        cfg = gravity_config()
    else:
        hulls = ellipsoid_hulls(4, seed=8)
        bodies = [BodyDesc.plane()]
        for i in range(12):
            pos = (0.25 * (i % 3), 1.0 + 1.4 * i, 0.15 * (i % 2))
            bodies.append(BodyDesc.hull(i % 4, position=pos) if i % 2 == 0
                          else BodyDesc.cylinder(0.3, 0.4, position=pos))
        cfg = config()
    print(f"[world parity] rollout {name} {frames} frames", flush=True)
    scene = encode(bodies, cfg, hulls if name == "hulls_cylinders" else None)
    ref, dev = oracle.world(scene), device_world(scene)
    for frame in range(frames):
        ref.step(DT)
        dev.step(DT)
        if frame % 10 == 9:
            assert np.array_equal(dev.pairs(), ref.pairs()), (name, frame)
            assert_counters_equal(dev, ref, f"{name} frame {frame}")
            # a rollout compounds the single-step tolerance; 1e-4 relative over 120 frames
            assert_states_close(dev.states(), ref.states(), pos_tol=1e-4, vel_tol=1e-3,
                                label=f"{name} frame {frame}")


@pytest.mark.parametrize("name,frames", [("tower16", 120), ("box_stack_plane", 120),
                                         ("spheres_capsules", 90), ("hulls_cylinders", 90),
                                         ("convex_pile_300", 60)])
def test_rollout_is_bit_identical_to_oracle(device, oracle, name, frames):
    """The device evaluates every float operation of the step in the oracle's order (no
    contraction, IEEE division and square root, correctly rounded sines and cosines), so over
    these rollouts its states and contacts equal the oracle's bit for bit."""
    hulls = None
    if name == "tower16":
        bodies, cfg = tower(16), config()
    elif name == "box_stack_plane":
        bodies = [BodyDesc.plane()] + [BodyDesc.cuboid((0.5, 0.5, 0.5), position=(0.02 * i, 0.55 + i * 1.1, 0.0))
                                       for i in range(6)]
        cfg = config()
    elif name == "spheres_capsules":
        bodies = [BodyDesc.plane()]
        for i in range(12):
            pos = (0.3 * (i % 3), 1.0 + 1.3 * i, 0.2 * (i % 2))
            bodies.append(BodyDesc.sphere(0.4, position=pos) if i % 2 == 0
                          else BodyDesc.capsule(0.25, 0.5, position=pos))
        cfg = gravity_config()
    elif name == "hulls_cylinders":
        hulls = ellipsoid_hulls(4, seed=8)
        bodies = [BodyDesc.plane()]
        for i in range(12):
            pos = (0.25 * (i % 3), 1.0 + 1.4 * i, 0.15 * (i % 2))
            bodies.append(BodyDesc.hull(i % 4, position=pos) if i % 2 == 0
                          else BodyDesc.cylinder(0.3, 0.4, position=pos))
        cfg = config()
    else:
        hulls = ellipsoid_hulls(12, seed=4)
        bodies, cfg = mixed_pile(300, seed=6, spacing=1.6, kinds=5, hull_sources=12), config()
    print(f"[world parity] bitwise rollout {name} {frames} frames", flush=True)
    scene = encode(bodies, cfg, hulls)
    ref, dev = oracle.world(scene), device_world(scene)
    for frame in range(frames):
        ref.step(DT)
        dev.step(DT)
        if frame % 15 == 14 or frame == frames - 1:
            assert dev.states().tobytes() == ref.states().tobytes(), (name, frame)
            assert dev.inspect_contacts().tobytes() == ref.inspect_contacts().tobytes(), (name, frame)


@pytest.mark.parametrize("solve,position,substeps", [(1, 1, 2), (1, 1, 1), (2, 1, 3), (3, 2, 1)])
def test_iteration_counts_bit_identical(device, oracle, solve, position, substeps):
    """Iteration counts at the edges of the solver's schedule (the parameters admit one of each
    at least): one velocity iteration (its apply is the substep's last and writes the velocities
    back), one position iteration (the first is the last: the pose goes from the body state to
    the body state), one and three substeps. The device's states and contacts equal the oracle's
    bit for bit over a settling stack and a few falling cubes."""
    bodies = [BodyDesc.plane()] + [BodyDesc.cuboid((0.5, 0.5, 0.5), position=(0.03 * i, 0.55 + i * 1.1, 0.0))
                                   for i in range(5)]
    bodies += [BodyDesc.cuboid((0.4, 0.3, 0.5), position=(2.0 + 0.9 * i, 1.5 + 0.7 * i, 0.1 * i)) for i in range(4)]
    cfg = config(solve_iterations=solve, position_iterations=position, substeps=substeps)
    print(f"[world parity] iterations {solve}/{position}/{substeps}", flush=True)
    scene = encode(bodies, cfg)
    ref, dev = oracle.world(scene), device_world(scene)
    for frame in range(90):
        ref.step(DT)
        dev.step(DT)
        if frame % 15 == 14:
            assert dev.states().tobytes() == ref.states().tobytes(), (solve, position, substeps, frame)
            assert dev.inspect_contacts().tobytes() == ref.inspect_contacts().tobytes(), (solve, position, substeps, frame)


def lowest_points(states, half=0.5):
    corners = np.array([[sx, sy, sz] for sx in (-half, half) for sy in (-half, half)
                        for sz in (-half, half)])
    return np.array([s["position"][1] + min(qrot(s["orientation"], c)[1] for c in corners)
                     for s in states])


def pile_statistics(world, scene):
    st = world.states()[:scene.dynamic_count]
    inv_mass = scene.descs["inverse_mass"][:scene.dynamic_count]
    mass = 1.0 / inv_mass
    ke = 0.5 * float(np.sum(mass * np.sum(st["velocity"].astype(np.float64) ** 2, axis=1)))
    low = lowest_points(st)
    m = world.measured()
    return {"penetration_max": float(max(0.0, -low.min())),
            "penetration_mean": float(np.mean(np.maximum(0.0, -low))),
            "kinetic_energy": ke, "mean_height": float(np.mean(st["position"][:, 1])),
            "resting": m["RESTING"], "contacts": m["CONTACTS"], "sleeping": int(st["sleeping"].sum())}


def test_cube_drop_long_run_statistics(device, oracle):
    """The 1,000-cube drop over its full 300 steps with 4 solver iterations: the device and
    the oracle end in statistically matching piles."""
    print("[world parity] 1000-cube drop, 300 steps", flush=True)
    scene = encode(cube_drop(1000, seed=0), config(solve_iterations=4))
    ref, dev = oracle.world(scene), device_world(scene)
    for frame in range(300):
        ref.step(DT)
        dev.step(DT)
        if frame == 29:
            # early frames are exact-path parity: free fall and first impacts
            assert np.array_equal(dev.pairs(), ref.pairs())
            assert_states_close(dev.states(), ref.states(), pos_tol=1e-4, vel_tol=1e-3,
                                label="frame 30")
    a, b = pile_statistics(dev, scene), pile_statistics(ref, scene)
    print("[world parity] device", a, "oracle", b, flush=True)
    assert a["penetration_max"] <= 0.02 + 0.005 + 0.02
    assert abs(a["penetration_mean"] - b["penetration_mean"]) <= 2e-3
    assert abs(a["mean_height"] - b["mean_height"]) <= 0.05 * max(1.0, b["mean_height"])
    assert abs(a["kinetic_energy"] - b["kinetic_energy"]) <= 0.25 * max(b["kinetic_energy"], 50.0)
    assert abs(a["contacts"] - b["contacts"]) <= 0.1 * b["contacts"] + 10
    assert abs(a["resting"] - b["resting"]) <= 0.1 * max(b["resting"], 10) + 10


def test_convex_pile_long_run_statistics(device, oracle):
    """800 spheres, boxes, capsules, hulls and cylinders dropped in four layers onto a plane and
    stepped 300 times: the device and the oracle follow the same path for the first frames and
    end in statistically matching piles. Both throw a few bodies out of the pile: when two
    stacked boxes' parallel faces overlap almost equally, box_box_sat's reference face
    (narrowphase.wgsl:334-350) flips between steps, the new features relay no impulse
    (identity.wgsl feature_carries_over) while the floor's impact impulse is warm started, and
    the Jacobi blocks launch the pair. Over seeds 13-15 the oracle throws 1.6-5.6 % of the
    bodies, so the bounds below compare statistics of the kept bodies."""
    print("[world parity] 800-body convex pile, 300 steps", flush=True)
    hulls = ellipsoid_hulls(8, seed=21)
    scene = encode(mixed_pile(800, seed=13, spacing=1.9, kinds=5, hull_sources=8), config(), hulls)
    ref, dev = oracle.world(scene), device_world(scene)
    d = scene.dynamic_count
    for frame in range(300):
        ref.step(DT)
        dev.step(DT)
        if frame == 29:
            assert np.array_equal(dev.pairs(), ref.pairs())
            assert_states_close(dev.states(), ref.states(), pos_tol=1e-4, vel_tol=1e-3, label="frame 30")

    def stats(world):
        st = world.states()[:d]
        v = np.linalg.norm(st["velocity"].astype(np.float64), axis=1)
        y = st["position"][:, 1].astype(np.float64)
        kept = (v < 5.0) & (y > 0.0) & (y < 20.0)
        return {"thrown": float(1.0 - kept.mean()), "median_v": float(np.median(v)),
                "mean_y": float(y[kept].mean()), "contacts": world.measured()["CONTACTS"]}

    a, b = stats(dev), stats(ref)
    print("[world parity] device", a, "oracle", b, flush=True)
    assert a["thrown"] <= 0.1 and b["thrown"] <= 0.1
    assert abs(a["thrown"] - b["thrown"]) <= 0.06
    assert a["median_v"] < 0.5 and b["median_v"] < 0.5
    assert abs(a["mean_y"] - b["mean_y"]) <= 0.1
    assert abs(a["contacts"] - b["contacts"]) <= 0.15 * b["contacts"]


def test_device_step_is_bit_deterministic(device):
    # determinism.rs:61-90: two worlds from the same records step bit-identically
    scene = encode(cube_drop(400, seed=9, footprint=10.0, low=1.0, high=15.0), config())
    outs = []
    for _ in range(2):
        world = device_world(scene)
        for _ in range(90):
            world.step(DT)
        outs.append((world.states().tobytes(), world.inspect_contacts().tobytes()))
        world.close()
    assert outs[0] == outs[1]


def test_graph_replays_match_eager_steps(device):
    """The captured step replayed many times, with reads (eager work) between replays, gives
    bit-identical states and contacts to the same steps launched eagerly."""
    hulls = ellipsoid_hulls(6, seed=3)
    scene = encode(mixed_pile(160, seed=8, spacing=1.5, kinds=5, hull_sources=6), config(), hulls)
    eager, graphed = device_world(scene), device_world(scene)
    eager.set_graphs(False)
    for frame in range(48):
        eager.step(DT)
        graphed.step(DT)
        if frame % 6 == 5:
            assert eager.states().tobytes() == graphed.states().tobytes(), frame
            assert eager.inspect_contacts().tobytes() == graphed.inspect_contacts().tobytes(), frame
            assert eager.measured() == graphed.measured(), frame
    eager.close()
    graphed.close()


def test_world_rejects_bad_records(device):
    from dynamis_amd._native import DynamisError
    scene = encode([BodyDesc.sphere(0.5), BodyDesc.sphere(0.5, position=(0, 2, 0))])
    bad = start_from(scene, scene.states)
    bad.states["body_id"] = 0  # duplicate ids
    with pytest.raises(DynamisError):
        device_world(bad)
    bad = start_from(scene, scene.states)
    bad.owners = scene.owners.copy()
    bad.owners[0] = 7
    with pytest.raises(DynamisError):
        device_world(bad)
