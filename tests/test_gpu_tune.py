"""The autotuner on the device: every sort tiling it may choose gives the same stable order, and a
world whose sorts were tuned steps bit for bit as an untuned one; the per-GPU cache is reused."""

from __future__ import annotations

import numpy as np
import pytest

from dynamis_amd import scenes, tune
from dynamis_amd.sort import RadixSort, SortChannels, units_for
from dynamis_amd.world import World, encode
from tests.sort_cases import expected_order, xorshift

pytestmark = pytest.mark.gpu

TILINGS = [(8, 512), (8, 1024), (8, 2048), (16, 512), (16, 1024), (16, 2048)]


@pytest.mark.parametrize("items,groups", TILINGS)
def test_every_tiling_sorts_the_same(device, items, groups):
    n = 300_000
    major, minor = xorshift(n, 11), xorshift(n, 12)
    payload = np.arange(n, dtype=np.uint32)
    try:
        tune.load(np.array([[0, tune.bucket(n), items, groups]], np.uint32))
        channels = SortChannels.upload(major, minor, payload)
        RadixSort().sort(channels, units_for(n), 4, 4)
        m, k, p = channels.download(n)
    finally:
        tune.load(np.zeros((0, 4), np.uint32))
    assert np.array_equal(p, expected_order(major, minor, 4, 4))
    assert np.array_equal(m, major[p]) and np.array_equal(k, minor[p])


def test_autotuned_world_steps_identically(device, tmp_path):
    def run(tuned):
        world = World()
        world.create(encode(scenes.cube_drop(2000, seed=3)))
        info = tune.autotune_world(world, tmp_path) if tuned else None
        for _ in range(30):
            world.step(1 / 60)
        return world.states(), info

    try:
        plain, _ = run(False)
        tuned, info = run(True)
        assert not info["from_cache"]
        assert all(t in TILINGS for t in info["tilings"].values())
        assert tuned.tobytes() == plain.tobytes()
        again, info2 = run(True)
        assert info2["from_cache"] and info2["tilings"] == info["tilings"]
        assert again.tobytes() == plain.tobytes()
    finally:
        tune.load(np.zeros((0, 4), np.uint32))


@pytest.mark.parametrize("threads", [64, 128])
def test_solver_workgroup_sizes_step_identically(device, threads):
    """The solver substeps' kernels in workgroups of 64 or 128 give the 256-thread result bit for
    bit, over a settling drop with many contacts."""
    scene = encode(scenes.cube_drop(1500, seed=4))

    def run(size):
        world = World()
        world.create(scene)
        if size:
            world.set_solver_threads(size)
        for _ in range(60):
            world.step(1 / 60)
        out = world.states(), world.inspect_contacts()
        world.close()
        return out

    (s0, c0), (s1, c1) = run(None), run(threads)
    assert len(c0) > 500
    assert s1.tobytes() == s0.tobytes() and c1.tobytes() == c0.tobytes()


def test_solver_tuning_leaves_the_world_unchanged(device):
    scene = encode(scenes.cube_drop(1500, seed=5))
    plain, tuned = World(), World()
    plain.create(scene)
    tuned.create(scene)
    for _ in range(40):
        plain.step(1 / 60)
        tuned.step(1 / 60)
    before = tuned.states().copy()
    threads, ms = tuned.tune_solver(3)
    assert threads in (64, 128, 256) and all(m > 0 for m in ms)
    assert tuned.states().tobytes() == before.tobytes()
    for _ in range(20):
        plain.step(1 / 60)
        tuned.step(1 / 60)
    assert tuned.states().tobytes() == plain.states().tobytes()
    assert tuned.inspect_contacts().tobytes() == plain.inspect_contacts().tobytes()
    plain.close()
    tuned.close()
