"""Capacity planning on the device (dyn_world_capacity, dyn_world_reserve): started below a cube
drop's demand, the device refuses, grows and steps exactly as the oracle does (the same lag, the
same growth rule, the same refusals), bit for bit; growth across captured graphs; snapshots are
tied to the capacities they were taken at; a differentiable rollout holds its world's capacities.
"""

from __future__ import annotations

import numpy as np
import pytest

from dynamis_amd import adjoint as adj
from dynamis_amd import scenes
from dynamis_amd._native import DynamisError
from dynamis_amd.scenes import cube_drop
from dynamis_amd.world import Batch, World, default_config, encode

from .test_capacity import check_plan, trace

pytestmark = pytest.mark.gpu

DT = np.float32(1.0 / 60.0)


def device_world(scene, contacts, pairs, graphs=True):
    w = World()
    w.create(scene, contacts, pairs)
    w.set_graphs(graphs)
    return w


@pytest.mark.parametrize("graphs", [True, False])
def test_device_grows_and_refuses_as_the_oracle_does(device, oracle, graphs):
    scene = encode(cube_drop(200, seed=1), default_config())
    ref = oracle.world(scene, 64, 256)
    dev = device_world(scene, 64, 256, graphs)
    rows_ref = trace(ref, 150)
    rows_dev = trace(dev, 150)
    check_plan(rows_dev)
    assert rows_dev == rows_ref
    assert dev.capacities() == ref.capacities() and dev.capacities()[2] >= 2
    assert dev.states().tobytes() == ref.states().tobytes()
    assert dev.inspect_contacts().tobytes() == ref.inspect_contacts().tobytes()
    assert any(r[3] or r[4] for r in rows_dev), "the drop outgrew its starting capacities"
    assert dev.refusals() == [] and not (rows_dev[-1][3] or rows_dev[-1][4])
    dev.close()
    ref.close()


def test_refusals_name_the_shortfall(device):
    scene = encode(cube_drop(200, seed=1), default_config())
    dev = device_world(scene, 64, 256)
    seen = set()
    for _ in range(80):
        dev.step(DT)
        for r in dev.refusals():
            assert r["shortfall"] == "physics" and r["count"] > 0
            seen.add(r["counter"])
    assert seen <= {"REFUSED_PAIRS", "REFUSED_CONTACTS"} and seen
    dev.close()


def test_reserve_and_snapshot_capacities(device):
    scene = encode(cube_drop(100, seed=3), default_config())
    dev = device_world(scene, None, None)
    for _ in range(30):
        dev.step(DT)
    snap = dev.snapshot()
    before = dev.states().copy()
    pairs, contacts, grows = dev.capacities()
    dev.reserve(2 * pairs, 3 * contacts)
    assert dev.capacities() == (2 * pairs, 3 * contacts, grows + 1)
    with pytest.raises(DynamisError, match="other capacities"):
        dev.restore(snap)
    # the durable archives survived the growth: the step from here equals the step without it
    again = device_world(scene, None, None)
    for _ in range(30):
        again.step(DT)
    assert again.states().tobytes() == before.tobytes()
    for _ in range(20):
        dev.step(DT)
        again.step(DT)
    assert dev.states().tobytes() == again.states().tobytes()
    assert dev.inspect_contacts().tobytes() == again.inspect_contacts().tobytes()
    snap.free()
    dev.close()
    again.close()


def test_differentiable_rollout_holds_capacities(device):
    sc, cfg, _ = scenes.adjoint_worlds(2, 8, seed=7)
    batch = Batch(sc, cfg)
    a = adj.Adjoint(batch, 8, 4)
    try:
        pairs, contacts, _ = batch.capacities()
        with pytest.raises(DynamisError, match="holds the world's capacities"):
            batch.reserve(2 * pairs, 2 * contacts)
    finally:
        a.close()
    batch.reserve(2 * pairs, 2 * contacts)
    assert batch.capacities()[:2] == (2 * pairs, 2 * contacts)
    batch.close()
