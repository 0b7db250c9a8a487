"""The full-size workloads on the device: the 100,000-body mixed convex pile stepped its whole
600 steps, and the 4,096-world batch stepped its whole 240 steps.

The pile runs the paths only a large single world reaches: the island seeding pass over a pile
that is one island, the gathered key-order records of long coarse grid runs, the autotuned
0.8 M-entry and 3.2 M-pair sorts. Every step is checked for faults and refusals; one step in
the middle of the run is compared with the oracle's step from the same states (pairs and counters
exact, states within 1e-5 relative, contacts within the per-step tolerances); the settled pile
is checked for bounded penetration, drained kinetic energy and a bounded fraction of thrown
bodies.

The batch runs worlds of 8 to 2,048 bodies side by side; sampled worlds, among them the
smallest and the largest, must step bit for bit as they do alone.
"""

from __future__ import annotations

import numpy as np
import pytest

from dynamis_amd.world import Batch, World, default_config, encode
from tests.test_gpu_world import (assert_contacts_close, assert_counters_equal, assert_states_close,
                                  start_from)
from tests.test_capacity import check_plan
from tests.worlds import DT

pytestmark = pytest.mark.gpu

FATAL = ("ENTRY_FAULTS", "ISLAND_FAULTS", "LIVE_FAULTS", "ROW_FAULTS")
FAULTS = FATAL + ("REFUSED_PAIRS", "REFUSED_CONTACTS", "REFUSED_RESTING")
PILE_STEPS = 600
PILE_COMPARE_AT = 300
BATCH_STEPS = 240


def pile_scene():
    from dynamis_amd.scenes import convex_pile
    bodies, hulls = convex_pile(100_000, seed=3)
    return encode(bodies, default_config(), hulls)


def assert_no_faults(world, label, counters=FAULTS):
    m = world.measured()
    bad = {k: m[k] for k in counters if m[k]}
    assert not bad, f"{label}: {bad}"


def contact_penetration(contacts) -> tuple[float, float]:
    """Deepest and mean penetration over the live manifold points (depth < 0 overlaps)."""
    depths = []
    for c in contacts:
        n = int(c["point_count"])
        depths.append(c["points"]["depth"][:n])
    d = np.concatenate(depths) if depths else np.zeros(1, np.float32)
    pen = np.maximum(0.0, -d.astype(np.float64))
    return float(pen.max()), float(pen.mean())


def test_convex_pile_100k_full_run(device, oracle):
    print("[configs] 100k convex pile: building the scene", flush=True)
    scene = pile_scene()
    d = scene.dynamic_count
    world = World()
    world.create(scene)
    ke_start = None
    mass = 1.0 / scene.descs["inverse_mass"][:d].astype(np.float64)
    plan = []
    for step in range(PILE_STEPS):
        world.step(DT)
        assert_no_faults(world, f"pile step {step}", FATAL)
        m = world.measured()
        plan.append((world.capacities(), m["PAIRS"], max(m["CONTACTS"], m["RESTING"] + m["REFUSED_RESTING"]),
                     m["REFUSED_PAIRS"], m["REFUSED_CONTACTS"]))
        if step == 120:
            v = world.states()["velocity"][:d].astype(np.float64)
            ke_start = 0.5 * float(np.sum(mass * np.sum(v * v, axis=1)))
        if step == PILE_COMPARE_AT - 1:
            print(f"[configs] pile: one step from the states after {step + 1} device steps, device "
                  f"vs oracle", flush=True)
            start = start_from(scene, world.states())
            ref = oracle.world(start)
            dev = World()
            dev.create(start)
            ref.step(DT)
            dev.step(DT)
            assert np.array_equal(dev.pairs(), ref.pairs()), "candidate pairs must be bit-exact"
            assert_counters_equal(dev, ref, "pile mid-run step")
            assert_contacts_close(dev.inspect_contacts(), ref.inspect_contacts(), "pile mid-run step")
            assert_states_close(dev.states(), ref.states(), label="pile mid-run step")
            dev.close()
            ref.close()
            print("[configs] pile: mid-run step matches the oracle", flush=True)
        if step % 100 == 99:
            print(f"[configs] pile step {step + 1}: {world.measured()['CONTACTS']} contacts", flush=True)
    # bodies thrown by the reference's manifold row mass (solver_contact_block.wgsl:14-28 can come
    # out non-positive for a lopsided manifold) sweep bounds across the pile and outgrow the pair
    # stream; every refusal is followed by the planned growth two steps later
    check_plan(plan)
    refusing = [t for t, r in enumerate(plan) if r[3] or r[4]]
    print(f"[configs] pile: capacities {world.capacities()}, steps with refusals {refusing}", flush=True)
    assert not (plan[-1][3] or plan[-1][4])
    st = world.states()[:d]
    m = world.measured()
    v = np.linalg.norm(st["velocity"].astype(np.float64), axis=1)
    pos = st["position"].astype(np.float64)
    inside = (np.abs(pos[:, 0]) < 30.0) & (np.abs(pos[:, 2]) < 30.0) & (pos[:, 1] > 0.0)
    kept = inside & (v < 5.0)
    thrown = float(1.0 - kept.mean())
    ke_kept = 0.5 * float(np.sum(mass[kept] * v[kept] ** 2))
    pen_max, pen_mean = contact_penetration(world.inspect_contacts())
    print(f"[configs] pile after {PILE_STEPS} steps: thrown {thrown:.4f}, median speed "
          f"{np.median(v):.4f}, kinetic energy of the kept bodies {ke_kept:.1f} J (step 120: {ke_start:.1f} J), "
          f"penetration max {pen_max:.4f} mean {pen_mean:.5f}, contacts {m['CONTACTS']}, islands "
          f"{m['ISLANDS']}", flush=True)
    assert thrown <= 0.1, thrown
    # the pile still creeps at 10 s (bodies fell from up to 70 m), far below its falling speeds
    assert np.median(v) < 2.5
    # the kept bodies have drained the energy they had while falling into the pile
    assert ke_kept < 0.1 * ke_start, (ke_kept, ke_start)
    # contacts resolve to within a few centimetres of the slop (0.005) and margin (0.02)
    assert pen_mean < 0.02, pen_mean
    assert pen_max < 0.25, pen_max
    assert m["CONTACTS"] > 2 * d
    world.close()


def test_batch_4096_worlds_full_run(device):
    from dynamis_amd.scenes import batch_worlds
    print("[configs] 4096-world batch: building the scenes", flush=True)
    scenes = batch_worlds(4096, seed=4)
    sizes = np.array([s.dynamic_count for s in scenes])
    sample = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), *range(0, 4096, 512),
                     int(np.argsort(sizes)[len(sizes) // 2])})
    print(f"[configs] batch: {int(sizes.sum())} bodies, worlds of {sizes.min()}-{sizes.max()}; "
          f"sampled worlds {sample} of sizes {[int(sizes[i]) for i in sample]}", flush=True)
    batch = Batch(scenes)
    alone = []
    for i in sample:
        w = World()
        w.create(scenes[i])
        alone.append(w)
    for step in range(BATCH_STEPS):
        batch.step(DT)
        for w in alone:
            w.step(DT)
        if step % 20 == 19:
            assert_no_faults(batch, f"batch step {step}")
        if step in (59, 159, BATCH_STEPS - 1):
            states, contacts = batch.states(), batch.inspect_contacts()
            for i, w in zip(sample, alone):
                assert batch.world_states(i, states).tobytes() == w.states().tobytes(), (step, i)
                assert batch.world_contacts(i, contacts).tobytes() == w.inspect_contacts().tobytes(), (step, i)
            print(f"[configs] batch step {step + 1}: sampled worlds bit-identical to themselves alone; "
                  f"{batch.measured()['CONTACTS']} contacts", flush=True)
    for w in alone:
        w.close()
    batch.close()
